// sn_kernels.h -- HIP kernels of the Metropolis dipole-lattice path (gfx950).
//
// All arithmetic is IEEE double.  The dipolar part of site_energy
// (src/starrynight-montecarlo-core.c:91-117) is evaluated as a local field:
//   dE_dd = L_i (p' - p) . h_i,   h_i = sum_r L_j (p_j - 3 n (n . p_j)) / d^3,
// which is the reference's per-neighbour expression with the common factor
// (p' - p) pulled out (it is linear in the proposal), and the cage term as
//   dE_cage = -J (p' - p) . g_i,  g_i = sum_{|r|=1} p_j.
#pragma once
#include "sn_device.h"

namespace sn {

constexpr int kBlock = 256;   // threads per workgroup of the tile sweep

struct SweepArgs {
    double4 *lat;
    const RepDev *reps;
    const TileDev *tiles;
    const ShapeDev *shapes;
    const int *class_start;   // per shape: ncol + 1 entries (offsets into class_ri)
    const int *class_ri;      // per shape: LDS region index of each tile site, by fine colour
    const int *chunk_col;     // per shape: nchunks + 1 colour indices
    const PairDev *pairs;
    const unsigned *sweep_ctr;
    unsigned long long *counters;  // 2 per replica: accepted, rejected
    int npairs, rc, two_d, max_region;
    int max_tile_sites, max_cls;   // sizes of the LDS copies of the shape tables
};

// DPP lane exchange of a double (two 32-bit moves); CTRL is a dpp_ctrl code.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_update_dpp(0, lo, CTRL, 0xF, 0xF, false);
    hi = __builtin_amdgcn_update_dpp(0, hi, CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double swz_xor16_f64(double v)
{
    // ds_swizzle bitmask mode: and 0x1F, or 0, xor 0x10 (within 32 lanes)
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_ds_swizzle(lo, 0x401F);
    hi = __builtin_amdgcn_ds_swizzle(hi, 0x401F);
    return __hiloint2double(hi, lo);
}

// Sum over aligned groups of G lanes, result in every lane of the group:
// butterfly of row_mirror (i <-> 15-i), row_half_mirror (i <-> 7-i) and
// quad permutations xor 2, xor 1 (plus a swizzle xor 16 for G = 32).
template <int G>
__device__ __forceinline__ double group_sum_dpp(double v)
{
    if (G >= 32) v += swz_xor16_f64(v);
    if (G >= 16) v += dpp_f64<0x140>(v);
    if (G >= 8) v += dpp_f64<0x141>(v);
    if (G >= 4) v += dpp_f64<0x4E>(v);
    if (G >= 2) v += dpp_f64<0xB1>(v);
    return v;
}

// One workgroup updates one tile: stage tile + halo in LDS, draw the
// proposals of all tile sites, run the fine colours one after another (G lanes
// share one site's neighbour sum, NPL +-r pairs per lane), write the tile back.
template <int G, int NPL, bool MIXED>
__global__ __launch_bounds__(kBlock) void tile_sweep_kernel(SweepArgs a, int first_tile, int ntiles)
{
    extern __shared__ double smem[];
    const int tid = threadIdx.x;
    const int b = blockIdx.x;
    // Consecutive tiles (spatial neighbours sharing halo sites) go to the same
    // XCD: the hardware deals workgroups to the 8 XCDs round-robin.
    const int t = (ntiles % 8 == 0) ? (b % 8) * (ntiles / 8) + b / 8 : b;
    const TileDev tile = a.tiles[first_tile + t];
    const RepDev rep = a.reps[tile.replica];
    const ShapeDev sh = a.shapes[tile.shape];
    const int hx = a.rc, hy = a.rc, hz = a.two_d ? 0 : a.rc;
    const int RY = sh.ty + 2 * hy, RZ = sh.tz + 2 * hz;
    const int RX = sh.tx + 2 * hx;
    const int region = RX * RY * RZ;

    double *sx = smem;
    double *sy = sx + a.max_region;
    double *sz = sy + a.max_region;
    double *sl = sz + a.max_region;                        // lengths (MIXED only)
    double *prop = MIXED ? sl + a.max_region : sl;         // 4 x kBlock: px, py, pz, threshold
    int *s_ri = (int *)(prop + 4 * kBlock);                // region index of class sites
    int *s_cls = s_ri + a.max_tile_sites;                  // class starts, then chunk colours
    int *s_chunk = s_cls + a.max_cls + 1;

    // ---- shape tables into LDS ----
    const int ntile = sh.tx * sh.ty * sh.tz;
    for (int i = tid; i < ntile; i += kBlock) s_ri[i] = a.class_ri[sh.sites_off + i];
    for (int i = tid; i <= sh.ncol; i += kBlock) s_cls[i] = a.class_start[sh.class_off + i];
    for (int i = tid; i <= sh.nchunks; i += kBlock) s_chunk[i] = a.chunk_col[sh.chunk_off + i];

    // ---- stage tile + halo (periodic wrap), 4 loads in flight per lane ----
    const double4 *lat = a.lat + rep.base;
    {
        // mixed-radix position of element tid, and the step of kBlock elements
        int iz = tid % RZ, q = tid / RZ, iy = q % RY, ix = q / RY;
        const int sz_ = kBlock % RZ, q2 = kBlock / RZ, sy_ = q2 % RY, sx_ = q2 / RY;
        for (int base = tid; base < region; base += 4 * kBlock) {
            double4 v[4];
            int idx[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                idx[k] = base + k * kBlock;
                if (idx[k] < region) {
                    int gx = wrap_once(tile.x0 - hx + ix, rep.X);
                    int gy = wrap_once(tile.y0 - hy + iy, rep.Y);
                    int gz = wrap_once(tile.z0 - hz + iz, rep.Z);
                    v[k] = lat[((long long)gx * rep.Y + gy) * rep.Z + gz];
                }
                iz += sz_; iy += sy_; ix += sx_;
                if (iz >= RZ) { iz -= RZ; iy++; }
                if (iy >= RY) { iy -= RY; ix++; }
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (idx[k] < region) {
                    sx[idx[k]] = v[k].x; sy[idx[k]] = v[k].y; sz[idx[k]] = v[k].z;
                    if (MIXED) sl[idx[k]] = v[k].w;
                }
            }
        }
    }

    // ---- this lane's share of the neighbour pairs, held in registers ----
    const int part = tid % G;
    const int group = tid / G;
    constexpr int ngroups = kBlock / G;
    int dl[NPL];
    double nx[NPL], ny[NPL], nz[NPL], w[NPL], w3[NPL];
#pragma unroll
    for (int m = 0; m < NPL; m++) {
        int j = part + m * G;
        if (j < a.npairs) {
            PairDev p = a.pairs[j];
            dl[m] = (p.dx * RY + p.dy) * RZ + p.dz;
            nx[m] = p.nx; ny[m] = p.ny; nz[m] = p.nz; w[m] = p.w; w3[m] = -3.0 * p.w;
        } else {
            dl[m] = 0; nx[m] = 0.0; ny[m] = 0.0; nz[m] = 0.0; w[m] = 0.0; w3[m] = 0.0;
        }
    }
    // nearest-neighbour pairs come first in the table, so they sit in slot 0
    const double nnf = (part < a.npairs && a.pairs[part].nn) ? 1.0 : 0.0;

    const unsigned sweep = a.sweep_ctr[tile.replica];
    const double J = rep.J, K = rep.K, E0 = rep.E[0], E1 = rep.E[1], E2 = rep.E[2];
    const long long tile_base = ((long long)tile.x0 * rep.Y + tile.y0) * rep.Z + tile.z0;
    unsigned nacc = 0, nrej = 0;
    __syncthreads();

    for (int ch = 0; ch < sh.nchunks; ch++) {
        const int c0 = s_chunk[ch], c1 = s_chunk[ch + 1];
        const int s0 = s_cls[c0], s1 = s_cls[c1];
        // ---- proposals of this chunk's sites: one site per lane ----
        if (tid < s1 - s0) {
            int ri = s_ri[s0 + tid];
            int iz = ri % RZ, q = ri / RZ;
            int iy = q % RY, ix = q / RY;
            unsigned gsite = (unsigned)(tile_base + ((long long)(ix - hx) * rep.Y + (iy - hy)) * rep.Z + (iz - hz));
            double px, py, pz, u;
            counter_draw(rep.seed, (uint32_t)tile.replica, sweep, gsite, rep.constrain_to_x, px, py, pz, u);
            prop[tid] = px;
            prop[kBlock + tid] = py;
            prop[2 * kBlock + tid] = pz;
            prop[3 * kBlock + tid] = metropolis_threshold(u, rep.beta);
        }
        __syncthreads();
        for (int c = c0; c < c1; c++) {
            const int cs = s_cls[c], ce = s_cls[c + 1];
            for (int sb = cs; sb < ce; sb += ngroups) {
                const int si = sb + group;
                const bool active = si < ce;
                const int ri = s_ri[active ? si : cs];

                double hx_ = 0.0, hy_ = 0.0, hz_ = 0.0;
                double gx_ = 0.0, gy_ = 0.0, gz_ = 0.0;
#pragma unroll
                for (int m = 0; m < NPL; m++) {
                    const int pa = ri + dl[m], pb = ri - dl[m];
                    const double ax = sx[pa], ay = sy[pa], az = sz[pa];
                    const double bx = sx[pb], by = sy[pb], bz = sz[pb];
                    const double ux = ax + bx, uy = ay + by, uz = az + bz;
                    double mx, my, mz;
                    if (MIXED) {
                        const double la = sl[pa], lb = sl[pb];
                        mx = la * ax + lb * bx; my = la * ay + lb * by; mz = la * az + lb * bz;
                    } else {
                        mx = ux; my = uy; mz = uz;
                    }
                    const double s = nx[m] * mx + ny[m] * my + nz[m] * mz;
                    const double ts = w3[m] * s;
                    hx_ += w[m] * mx + ts * nx[m];
                    hy_ += w[m] * my + ts * ny[m];
                    hz_ += w[m] * mz + ts * nz[m];
                    if (m == 0) { gx_ = nnf * ux; gy_ = nnf * uy; gz_ = nnf * uz; }
                }
                const int pq = active ? si - s0 : 0;
                const double px = prop[pq], py = prop[kBlock + pq], pz = prop[2 * kBlock + pq];
                const double thr = prop[3 * kBlock + pq];
                const double ox = sx[ri], oy = sy[ri], oz = sz[ri];
                const double Li = MIXED ? sl[ri] : rep.L0;
                const double coef = MIXED ? Li : Li * rep.L0;
                const double dx = px - ox, dy = py - oy, dz = pz - oz;
                double partial = coef * (dx * hx_ + dy * hy_ + dz * hz_) - J * (dx * gx_ + dy * gy_ + dz * gz_);
                partial = group_sum_dpp<G>(partial);
                double dE = partial + (dx * E0 + dy * E1 + dz * E2);
                if (K > 0.0) dE += -K * (fabs(px) - fabs(ox) + fabs(py) - fabs(oy));
                const bool acc = dE < thr;
                if (active && part == 0 && Li != 0.0) {
                    if (acc) {
                        sx[ri] = px; sy[ri] = py; sz[ri] = pz;
                        nacc++;
                    } else nrej++;
                }
            }
            __syncthreads();
        }
    }

    // ---- write the tile back ----
    double4 *out = a.lat + rep.base;
    for (int i = tid; i < ntile; i += kBlock) {
        int lz = i % sh.tz, q = i / sh.tz;
        int ly = q % sh.ty, lx = q / sh.ty;
        int ri = ((lx + hx) * RY + (ly + hy)) * RZ + (lz + hz);
        double4 v;
        v.x = sx[ri]; v.y = sy[ri]; v.z = sz[ri];
        v.w = MIXED ? sl[ri] : rep.L0;
        out[tile_base + ((long long)lx * rep.Y + ly) * rep.Z + lz] = v;
    }
    // ---- counters: one atomic per wave ----
    unsigned long long ca = nacc, cr = nrej;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        ca += __shfl_xor(ca, o, 64);
        cr += __shfl_xor(cr, o, 64);
    }
    if ((tid & 63) == 0 && (ca | cr)) {
        atomicAdd(&a.counters[2 * tile.replica], ca);
        atomicAdd(&a.counters[2 * tile.replica + 1], cr);
    }
}

__global__ void bump_sweep_kernel(unsigned *sweep_ctr, int R)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R) sweep_ctr[r] += 1u;
}

// Local field pieces of one site by direct gathers from HBM (used by the
// site_energy batch, the replay and the energy / adjoint kernels).
struct FieldOut {
    double hx, hy, hz;  // sum_r L_j (p_j - 3 n (n.p_j)) / d^3
    double gx, gy, gz;  // sum_{|r|=1} p_j
};

__device__ __forceinline__ void site_field(const double4 *lat, const RepDev &rep, const PairDev *offs,
                                           int noff, int x, int y, int z, int first, int stride,
                                           FieldOut &f)
{
    f.hx = f.hy = f.hz = f.gx = f.gy = f.gz = 0.0;
    for (int i = first; i < noff; i += stride) {
        const PairDev o = offs[i];
        int gx = wrap_once(x + o.dx, rep.X), gy = wrap_once(y + o.dy, rep.Y), gz = wrap_once(z + o.dz, rep.Z);
        double4 t = lat[((long long)gx * rep.Y + gy) * rep.Z + gz];
        double s = o.nx * t.x + o.ny * t.y + o.nz * t.z;
        double wl = o.w * t.w;
        double ts = -3.0 * wl * s;
        f.hx += wl * t.x + ts * o.nx;
        f.hy += wl * t.y + ts * o.ny;
        f.hz += wl * t.z + ts * o.nz;
        if (o.nn) { f.gx += t.x; f.gy += t.y; f.gz += t.z; }
    }
}

__global__ void site_energy_batch_kernel(const double4 *lat, const RepDev *reps, int replica,
                                         const PairDev *offs, int noff, long long n,
                                         const int *site, const double *prop, double *dE)
{
    long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    const RepDev rep = reps[replica];
    const double4 *L = lat + rep.base;
    int s = site[k];
    int z = s % rep.Z, q = s / rep.Z, y = q % rep.Y, x = q / rep.Y;
    double4 o = L[s];
    FieldOut f;
    site_field(L, rep, offs, noff, x, y, z, 0, 1, f);
    double px = prop[3 * k], py = prop[3 * k + 1], pz = prop[3 * k + 2];
    double dx = px - o.x, dy = py - o.y, dz = pz - o.z;
    double e = o.w * (dx * f.hx + dy * f.hy + dz * f.hz) - rep.J * (dx * f.gx + dy * f.gy + dz * f.gz) +
               (dx * rep.E[0] + dy * rep.E[1] + dz * rep.E[2]);
    if (rep.K > 0.0) e += -rep.K * (fabs(px) - fabs(o.x) + fabs(py) - fabs(o.y));
    dE[k] = e;
}

// Sequential replay of recorded moves: one workgroup, one move at a time;
// lanes split the neighbour list, a workgroup reduction gives h, g and the
// rounding scale, lane 0 decides and writes.
__global__ __launch_bounds__(256) void replay_kernel(double4 *lat, const RepDev *reps, int replica,
                                                     const PairDev *offs, int noff, long long n,
                                                     const int *site, const double *prop,
                                                     const double *u, const unsigned char *u_drawn,
                                                     const unsigned char *ref_accept,
                                                     unsigned char *decision,
                                                     unsigned long long *stats)
{
    __shared__ double red[7][256];
    const int tid = threadIdx.x;
    const RepDev rep = reps[replica];
    double4 *L = lat + rep.base;
    unsigned long long mism = 0, ties = 0, nacc = 0, nrej = 0;
    for (long long k = 0; k < n; k++) {
        const int s = site[k];
        const int z = s % rep.Z, q = s / rep.Z, y = q % rep.Y, x = q / rep.Y;
        const double4 o = L[s];
        const double px = prop[3 * k], py = prop[3 * k + 1], pz = prop[3 * k + 2];
        double v[7] = {0, 0, 0, 0, 0, 0, 0};
        for (int i = tid; i < noff; i += 256) {
            const PairDev of = offs[i];
            int gx = wrap_once(x + of.dx, rep.X), gy = wrap_once(y + of.dy, rep.Y), gz = wrap_once(z + of.dz, rep.Z);
            double4 t = L[((long long)gx * rep.Y + gy) * rep.Z + gz];
            double s_ = of.nx * t.x + of.ny * t.y + of.nz * t.z;
            double wl = of.w * t.w;
            double ts = -3.0 * wl * s_;
            v[0] += wl * t.x + ts * of.nx;
            v[1] += wl * t.y + ts * of.ny;
            v[2] += wl * t.z + ts * of.nz;
            if (of.nn) { v[3] += t.x; v[4] += t.y; v[5] += t.z; }
            // rounding scale of the reference's float terms for this neighbour
            double pn = fabs(px * t.x + py * t.y + pz * t.z), po = fabs(o.x * t.x + o.y * t.y + o.z * t.z);
            double an = fabs(of.nx * px + of.ny * py + of.nz * pz), ao = fabs(of.nx * o.x + of.ny * o.y + of.nz * o.z);
            v[6] += fabs(o.w * t.w) * of.w * (pn + po + 3.0 * fabs(s_) * (an + ao));
            if (of.nn) v[6] += fabs(rep.J) * (pn + po);
        }
        for (int c = 0; c < 7; c++) red[c][tid] = v[c];
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if (tid < st)
                for (int c = 0; c < 7; c++) red[c][tid] += red[c][tid + st];
            __syncthreads();
        }
        if (tid == 0) {
            const double dx = px - o.x, dy = py - o.y, dz = pz - o.z;
            double e = o.w * (dx * red[0][0] + dy * red[1][0] + dz * red[2][0]) -
                       rep.J * (dx * red[3][0] + dy * red[4][0] + dz * red[5][0]) +
                       (dx * rep.E[0] + dy * rep.E[1] + dz * rep.E[2]);
            double scale = red[6][0] + fabs(px * rep.E[0] + py * rep.E[1] + pz * rep.E[2]) +
                           fabs(o.x * rep.E[0] + o.y * rep.E[1] + o.z * rep.E[2]);
            if (rep.K > 0.0) {
                e += -rep.K * (fabs(px) - fabs(o.x) + fabs(py) - fabs(o.y));
                scale += rep.K * (fabs(px) + fabs(o.x) + fabs(py) + fabs(o.y));
            }
            // float rounding bound of the reference's dE (terms in float,
            // summed in double, result rounded to float)
            const double delta = 4e-6 * scale + 2e-7 * fabs(e);
            const int ref = ref_accept[k] ? 1 : 0;
            int mine = -1;   // -1: undecidable without a u the reference never drew
            bool tie = false;
            if (e < 0.0) mine = 1;
            else if (u_drawn[k]) mine = exp(-e * rep.beta) > u[k];
            if (fabs(e) <= delta) tie = true;
            if (u_drawn[k] && !__builtin_isinf(rep.beta)) {
                double g = exp(-e * rep.beta);
                double bound = rep.beta * delta * exp(-(e - delta) * rep.beta);
                if (fabs(g - u[k]) <= bound) tie = true;
            }
            int apply = ref;
            if (mine != ref) {
                if (tie) ties++;
                else mism++;
            }
            if (o.w != 0.0) {
                if (apply) { L[s] = make_double4(px, py, pz, o.w); nacc++; }
                else nrej++;
            }
            decision[k] = (unsigned char)apply;
        }
        __syncthreads();
    }
    if (tid == 0) { stats[0] = mism; stats[1] = ties; stats[2] = nacc; stats[3] = nrej; }
}

// Energy and adjoint: one site per lane; per-site gradient dE/dp_i and
// per-workgroup partial sums of
//   {E_dd, E_cage, E_field, E_K, dE/dJ, dE/dEx, dE/dEy, dE/dEz, dE/ds, dE/dK}.
constexpr int kNPart = 10;

__global__ __launch_bounds__(256) void energy_grad_kernel(const double4 *lat, const RepDev *reps,
                                                          int replica, const PairDev *offs, int noff,
                                                          double *grad, double *partials)
{
    __shared__ double red[kNPart][256];
    const int tid = threadIdx.x;
    const RepDev rep = reps[replica];
    const double4 *L = lat + rep.base;
    const long long N = (long long)rep.X * rep.Y * rep.Z;
    long long s = (long long)blockIdx.x * blockDim.x + tid;
    double v[kNPart];
    for (int c = 0; c < kNPart; c++) v[c] = 0.0;
    if (s < N) {
        int z = (int)(s % rep.Z), q = (int)(s / rep.Z), y = q % rep.Y, x = q / rep.Y;
        double4 p = L[s];
        FieldOut f;
        site_field(L, rep, offs, noff, x, y, z, 0, 1, f);
        double ph = p.x * f.hx + p.y * f.hy + p.z * f.hz;
        double pg = p.x * f.gx + p.y * f.gy + p.z * f.gz;
        double gx = p.w * f.hx - rep.J * f.gx + rep.E[0];
        double gy = p.w * f.hy - rep.J * f.gy + rep.E[1];
        double gz = p.w * f.hz - rep.J * f.gz + rep.E[2];
        double ak = fabs(p.x) + fabs(p.y);
        if (rep.K > 0.0) {
            gx -= rep.K * ((p.x > 0.0) - (p.x < 0.0));
            gy -= rep.K * ((p.y > 0.0) - (p.y < 0.0));
        }
        if (grad) { grad[3 * s] = gx; grad[3 * s + 1] = gy; grad[3 * s + 2] = gz; }
        v[0] = 0.5 * p.w * ph;
        v[1] = -0.5 * rep.J * pg;
        v[2] = p.x * rep.E[0] + p.y * rep.E[1] + p.z * rep.E[2];
        v[3] = rep.K > 0.0 ? -rep.K * ak : 0.0;
        v[4] = -0.5 * pg;
        v[5] = p.x; v[6] = p.y; v[7] = p.z;
        v[8] = p.w * ph;
        v[9] = rep.K > 0.0 ? -ak : 0.0;
    }
    for (int c = 0; c < kNPart; c++) red[c][tid] = v[c];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (tid < st)
            for (int c = 0; c < kNPart; c++) red[c][tid] += red[c][tid + st];
        __syncthreads();
    }
    if (tid < kNPart) partials[(long long)blockIdx.x * kNPart + tid] = red[tid][0];
}

// Fixed-order sum of per-block partials (deterministic): one workgroup,
// Neumaier-compensated per lane, then a compensated tree.
__global__ __launch_bounds__(256) void reduce_partials_kernel(const double *partials, int nblocks,
                                                              int width, double *out)
{
    __shared__ double s_[256], c_[256];
    const int tid = threadIdx.x;
    for (int col = 0; col < width; col++) {
        double s = 0.0, c = 0.0;
        for (int b = tid; b < nblocks; b += 256) {
            double v = partials[(long long)b * width + col];
            double t = s + v;
            c += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
            s = t;
        }
        s_[tid] = s; c_[tid] = c;
        __syncthreads();
        for (int st = 128; st > 0; st >>= 1) {
            if (tid < st) {
                double a = s_[tid], v = s_[tid + st];
                double t = a + v;
                c_[tid] += c_[tid + st] + ((fabs(a) >= fabs(v)) ? (a - t) + v : (v - t) + a);
                s_[tid] = t;
            }
            __syncthreads();
        }
        if (tid == 0) out[col] = s_[0] + c_[0];
        __syncthreads();
    }
}

// Sum of dipole vectors (partials per block, then reduce_partials_kernel).
__global__ __launch_bounds__(256) void moment_kernel(const double4 *lat, const RepDev *reps, int replica,
                                                     double *partials)
{
    __shared__ double red[3][256];
    const int tid = threadIdx.x;
    const RepDev rep = reps[replica];
    const long long N = (long long)rep.X * rep.Y * rep.Z;
    double v0 = 0, v1 = 0, v2 = 0;
    for (long long s = (long long)blockIdx.x * blockDim.x + tid; s < N; s += (long long)gridDim.x * blockDim.x) {
        double4 p = lat[rep.base + s];
        v0 += p.x; v1 += p.y; v2 += p.z;
    }
    red[0][tid] = v0; red[1][tid] = v1; red[2][tid] = v2;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (tid < st) for (int c = 0; c < 3; c++) red[c][tid] += red[c][tid + st];
        __syncthreads();
    }
    if (tid < 3) partials[(long long)blockIdx.x * 3 + tid] = red[tid][0];
}

// Random initial lattice on the device: unit dipole from the counter RNG
// (stream tag 0xFFFFFFFF in the sweep slot) and a length drawn from the
// species table (cumulative prevalence walk, as solid_solution does).
__global__ void init_random_kernel(double4 *lat, const RepDev *reps, int replica, int nspec,
                                   const double *len, const double *cum)
{
    const RepDev rep = reps[replica];
    const long long N = (long long)rep.X * rep.Y * rep.Z;
    long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N) return;
    double px, py, pz, u;
    counter_draw(rep.seed, (uint32_t)replica, 0xFFFFFFFFu, (uint32_t)s, 0, px, py, pz, u);
    int k = 0;
    while (k + 1 < nspec && u >= cum[k]) k++;
    lat[rep.base + s] = make_double4(px, py, pz, len[k]);
}

}  // namespace sn
