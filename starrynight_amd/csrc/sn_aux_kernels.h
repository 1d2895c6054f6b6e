// sn_aux_kernels.h -- gather kernels beside the tile sweep (gfx950, FP64):
// sweep-counter bump, the site_energy batch, sequential replay, energy /
// adjoint, reductions, moment and the random initial lattice.  Included by
// sn_api.hip only (non-template kernels: one definition).
#pragma once
#include "sn_device.h"

namespace sn {
__global__ void bump_sweep_kernel(unsigned *sweep_ctr, int R, unsigned n)
{
    int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < R) sweep_ctr[r] += n;
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
    // an index outside the replica reads nothing and yields NaN
    if (s < 0 || (long long)s >= (long long)rep.X * rep.Y * rep.Z) {
        dE[k] = __builtin_nan("");
        return;
    }
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

// Fixed-order sum of per-block partials (deterministic).  Workgroup (g, y)
// sums columns [g*kReduceCols, (g+1)*kReduceCols) of rows [seg[y], seg[y+1])
// (all nblocks rows when seg is null) of the nblocks x width table into
// out[y * width + col]: every
// lane Neumaier-sums its rows of all its columns (their loads in flight
// together), a compensated xor butterfly reduces each wave, and lanes 0..
// kReduceCols-1 combine the four waves in order.
constexpr int kReduceCols = 16;

__device__ __forceinline__ void neumaier_add(double &s, double &c, double v)
{
    const double t = s + v;
    c += (fabs(s) >= fabs(v)) ? (s - t) + v : (v - t) + s;
    s = t;
}

__global__ __launch_bounds__(256) void reduce_partials_kernel(const double *partials, int nblocks,
                                                              int width, double *out, const int *seg)
{
    __shared__ double s_w[kReduceCols][4], c_w[kReduceCols][4];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int col0 = blockIdx.x * kReduceCols;
    const int ncol = min(kReduceCols, width - col0);
    const int b0 = seg ? seg[blockIdx.y] : 0, b1 = seg ? seg[blockIdx.y + 1] : nblocks;
    out += (long long)blockIdx.y * width;
    double s[kReduceCols], c[kReduceCols];
#pragma unroll
    for (int k = 0; k < kReduceCols; k++) { s[k] = 0.0; c[k] = 0.0; }
    for (int b = b0 + tid; b < b1; b += 256) {
        const double *row = partials + (long long)b * width + col0;
#pragma unroll
        for (int k = 0; k < kReduceCols; k++)
            if (k < ncol) neumaier_add(s[k], c[k], row[k]);
    }
#pragma unroll
    for (int k = 0; k < kReduceCols; k++) {
        if (k < ncol) {   // uniform over the workgroup
#pragma unroll
            for (int o = 32; o >= 1; o >>= 1) {
                const double v = __shfl_xor(s[k], o, 64), cv = __shfl_xor(c[k], o, 64);
                c[k] += cv;
                neumaier_add(s[k], c[k], v);
            }
            if (lane == 0) { s_w[k][wave] = s[k]; c_w[k][wave] = c[k]; }
        }
    }
    __syncthreads();
    if (tid < ncol) {
        double S = s_w[tid][0], Cc = c_w[tid][0];
        for (int w = 1; w < 4; w++) {
            Cc += c_w[tid][w];
            neumaier_add(S, Cc, s_w[tid][w]);
        }
        out[col0 + tid] = S + Cc;
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

// Scan of a lattice written from device memory: per workgroup {count of
// non-finite values, min length, max length} (the checks sn_set_lattice makes
// on the host).
__global__ __launch_bounds__(256) void lattice_scan_kernel(const double4 *lat, long long N, double *partials)
{
    __shared__ double red[3][256];
    const int tid = threadIdx.x;
    double bad = 0.0, lo = INFINITY, hi = -INFINITY;
    for (long long s = (long long)blockIdx.x * blockDim.x + tid; s < N; s += (long long)gridDim.x * blockDim.x) {
        const double4 p = lat[s];
        if (!isfinite(p.x) || !isfinite(p.y) || !isfinite(p.z) || !isfinite(p.w)) bad += 1.0;
        lo = fmin(lo, p.w);
        hi = fmax(hi, p.w);
    }
    red[0][tid] = bad; red[1][tid] = lo; red[2][tid] = hi;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (tid < st) {
            red[0][tid] += red[0][tid + st];
            red[1][tid] = fmin(red[1][tid], red[1][tid + st]);
            red[2][tid] = fmax(red[2][tid], red[2][tid + st]);
        }
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

// radial_order_parameter (src/starrynight-analysis.c:528-598) in FP64.  Over
// every site i and every offset r with |r|^2 <= 80 (the dz range runs even
// when Z == 1, where it wraps onto the same plane, as in the reference):
//   FE[r^2]  += p_i.p_j,   AFE[r^2] += p_i.p_j - 3 (n.p_i)(n.p_j),  j = i + r
// with n = r/|r| (n = 0 at r = 0).  Offsets arrive sorted by r^2 with
// bin_start[82]; the loop over them is uniform across the wave, so the
// offset table is read through scalar loads.  One lane per site; each wave
// reduces a bin and its lane 0 adds to that wave's own LDS row, so no LDS
// atomics; a block writes 162 partial sums.
struct RadialOff {
    int dx, dy, dz, pad;
    double nx, ny, nz, pad2;
};
constexpr int kRadialBins = 81;

__global__ __launch_bounds__(256) void radial_order_kernel(const double4 *lat, const RepDev *reps, int replica,
                                                           const RadialOff *offs, const int *bin_start,
                                                           double *partials)
{
    __shared__ double hist[4][2 * kRadialBins];
    const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
    for (int i = tid; i < 4 * 2 * kRadialBins; i += 256) (&hist[0][0])[i] = 0.0;
    __syncthreads();
    const RepDev rep = reps[replica];
    const int X = rep.X, Y = rep.Y, Z = rep.Z;
    const long long N = (long long)X * Y * Z;
    // x + dx + 9X >= 0 for |dx| <= 9: one unsigned modulo wraps any axis length
    const unsigned bx = 9u * X, by = 9u * Y, bz = 9u * Z;
    for (long long s0 = (long long)blockIdx.x * 256; s0 < N; s0 += (long long)gridDim.x * 256) {
        const long long s = s0 + tid;
        const bool act = s < N;
        const long long sc = act ? s : 0;
        const int z = (int)(sc % Z), y = (int)((sc / Z) % Y), x = (int)(sc / ((long long)Y * Z));
        double4 a = lat[rep.base + sc];
        if (!act) a = make_double4(0.0, 0.0, 0.0, 0.0);   // contributes exact zeros
        for (int b = 0; b < kRadialBins; b++) {
            double fe = 0.0, afe = 0.0;
            const int k1 = bin_start[b + 1];
            for (int k = bin_start[b]; k < k1; k++) {
                const RadialOff o = offs[k];
                const unsigned xx = (unsigned)(x + o.dx + (int)bx) % (unsigned)X;
                const unsigned yy = (unsigned)(y + o.dy + (int)by) % (unsigned)Y;
                const unsigned zz = (unsigned)(z + o.dz + (int)bz) % (unsigned)Z;
                const double4 q = lat[rep.base + ((long long)xx * Y + yy) * Z + zz];
                const double f = a.x * q.x + a.y * q.y + a.z * q.z;
                const double na = o.nx * a.x + o.ny * a.y + o.nz * a.z;
                const double nq = o.nx * q.x + o.ny * q.y + o.nz * q.z;
                fe += f;
                afe += f - 3.0 * na * nq;
            }
            fe = group_sum<64>(fe);
            afe = group_sum<64>(afe);
            if (lane == 0) {
                hist[wave][2 * b] += fe;
                hist[wave][2 * b + 1] += afe;
            }
        }
    }
    __syncthreads();
    for (int i = tid; i < 2 * kRadialBins; i += 256)
        partials[(long long)blockIdx.x * 2 * kRadialBins + i] = (hist[0][i] + hist[1][i]) + (hist[2][i] + hist[3][i]);
}

// dipole_potential (src/starrynight-analysis.c:65-94) and dipole_electricfield
// (src/starrynight-analysis.c:393-457) at every site of a replica, FP64.  One
// lane per site; the offset loop is uniform across the wave, so the table is
// read through scalar loads.  Potential offsets carry a = r / |r|^3,
// field offsets n = r / |r| and w = 1 / |r|^3.
struct StencilOff {
    int dx, dy, dz, pad;
    double ax, ay, az, w;
};

template <bool FIELD>
__global__ __launch_bounds__(256) void stencil_observable_kernel(const double4 *lat, const RepDev *reps,
                                                                 int replica, const StencilOff *offs, int noff,
                                                                 double *out)
{
    const RepDev rep = reps[replica];
    const int X = rep.X, Y = rep.Y, Z = rep.Z;
    const long long N = (long long)X * Y * Z;
    const long long s = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= N) return;
    const int z = (int)(s % Z), y = (int)((s / Z) % Y), x = (int)(s / ((long long)Y * Z));
    // x + dx + 9X >= 0 for |dx| <= 9 X: one unsigned modulo wraps any axis
    const unsigned bx = 9u * X, by = 9u * Y, bz = 9u * Z;
    double ex = 0.0, ey = 0.0, ez = 0.0;
    for (int k = 0; k < noff; k++) {
        const StencilOff o = offs[k];
        const unsigned xx = (unsigned)(x + o.dx + (int)bx) % (unsigned)X;
        const unsigned yy = (unsigned)(y + o.dy + (int)by) % (unsigned)Y;
        const unsigned zz = (unsigned)(z + o.dz + (int)bz) % (unsigned)Z;
        const double4 q = lat[rep.base + ((long long)xx * Y + yy) * Z + zz];
        if (FIELD) {
            const double radial = o.ax * q.x + o.ay * q.y + o.az * q.z;
            ex += (3.0 * o.ax * radial - q.x) * o.w;
            ey += (3.0 * o.ay * radial - q.y) * o.w;
            ez += (3.0 * o.az * radial - q.z) * o.w;
        } else {
            ex += q.w * (q.x * o.ax + q.y * o.ay + q.z * o.az);
        }
    }
    if (FIELD) {
        const double4 a = lat[rep.base + s];
        out[3 * s] = ex - a.x / 3.0;
        out[3 * s + 1] = ey - a.y / 3.0;
        out[3 * s + 2] = ez - a.z / 3.0;
    } else {
        out[s] = ex;
    }
}

// recombination_calculator's sums (src/starrynight-analysis.c:97-164) over a
// potential map: per workgroup partial sums of {exp(-b p), exp(b p),
// 1 / (exp(b p) + 1), 1 / (exp(-b p) + 1), their product}, p = scale * pot.
constexpr int kRecombParts = 5;

__global__ __launch_bounds__(256) void recombination_kernel(const double *pot, long long N, double scale,
                                                            double beta, double *partials)
{
    __shared__ double red[kRecombParts][256];
    const int tid = threadIdx.x;
    double v[kRecombParts] = {0, 0, 0, 0, 0};
    for (long long i = (long long)blockIdx.x * 256 + tid; i < N; i += (long long)gridDim.x * 256) {
        const double p = scale * pot[i];
        const double ep = exp(p * beta), em = exp(-p * beta);
        const double fe = 1.0 / (ep + 1.0), fh = 1.0 / (em + 1.0);
        v[0] += em; v[1] += ep; v[2] += fe; v[3] += fh; v[4] += fe * fh;
    }
    for (int c = 0; c < kRecombParts; c++) red[c][tid] = v[c];
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (tid < st)
            for (int c = 0; c < kRecombParts; c++) red[c][tid] += red[c][tid + st];
        __syncthreads();
    }
    if (tid < kRecombParts) partials[(long long)blockIdx.x * kRecombParts + tid] = red[tid][0];
}

}  // namespace sn
