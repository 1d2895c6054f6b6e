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
    unsigned long long *counters;  // 2 per tile: accepted, rejected (accumulated over sweeps)
    const int *gsteps;        // per shape (gstep_off): per step, barrier flag + BLK / G entries
                              // (region index | proposal slot << 16, or -1)
    const int *gchunk;        // per shape (gchunk_off): first step of each proposal chunk (nchunks + 1)
    int npairs, rc, two_d, max_region;
    int max_tile_sites, max_cls;   // sizes of the LDS copies of the shape tables
    int max_gsteps, max_chunks;
    int prop_stride;          // proposal slots per chunk (<= kBlock)
};

// DPP lane exchange of a double (two 32-bit moves); CTRL is a dpp_ctrl code.
template <int CTRL>
__device__ __forceinline__ double dpp_f64(double v)
{
    // every control used here permutes lanes within a row, so each lane has a
    // source and no old value is needed
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __builtin_amdgcn_mov_dpp(lo, CTRL, 0xF, 0xF, true);
    hi = __builtin_amdgcn_mov_dpp(hi, CTRL, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}

// v + v[lane ^ 16] (ROWS = 16) or v + v[lane ^ 32] (ROWS = 32) in every lane,
// by gfx950's v_permlane16_swap / v_permlane32_swap on both halves of the
// double: the two results of a swap hold the even and the odd row of each
// row pair (no LDS round trip, unlike ds_swizzle or ds_bpermute).  The sum is
// the same in both lanes of a pair, bit for bit (addition commutes).
template <int ROWS>
__device__ __forceinline__ double pair_sum_f64(double v)
{
    const unsigned lo = (unsigned)__double2loint(v), hi = (unsigned)__double2hiint(v);
    if constexpr (ROWS == 16) {
        const auto l = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
        const auto h = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
        return __hiloint2double((int)h[0], (int)l[0]) + __hiloint2double((int)h[1], (int)l[1]);
    } else {
        const auto l = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
        const auto h = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
        return __hiloint2double((int)h[0], (int)l[0]) + __hiloint2double((int)h[1], (int)l[1]);
    }
}

// Sum over aligned groups of G lanes, result in every lane of the group:
// butterfly of row_mirror (i <-> 15-i), row_half_mirror (i <-> 7-i) and
// quad permutations xor 2, xor 1, plus row-pair swaps for G >= 32 (lane xor
// 16) and G = 64 (lane xor 32).
template <int G>
__device__ __forceinline__ double group_sum_dpp(double v)
{
    static_assert(G == 2 || G == 4 || G == 8 || G == 16 || G == 32 || G == 64, "G is a power of two <= 64");
    if (G >= 64) v = pair_sum_f64<32>(v);
    if (G >= 32) v = pair_sum_f64<16>(v);
    if (G >= 16) v += dpp_f64<0x140>(v);
    if (G >= 8) v += dpp_f64<0x141>(v);
    if (G >= 4) v += dpp_f64<0x4E>(v);
    if (G >= 2) v += dpp_f64<0xB1>(v);
    return v;
}

// One workgroup updates one tile: stage tile + halo in LDS, draw the
// proposals of all tile sites, run the fine colours one after another (G lanes
// share one site's neighbour sum, NPL +-r pairs per lane), write the tile back.
__device__ __forceinline__ int s_gchunk_total(const ShapeDev &sh) { return sh.gsteps_n; }

// BLK threads per workgroup (256 or 512); proposals are drawn in chunks of at
// most prop_stride (<= kBlock) sites.  The staged x, y, z arrays are read by
// separate instructions: a read2 of x and y at a stride that is a multiple of
// the bank cycle puts both halves in one bank (measured: 2.9 conflict cycles
// per LDS instruction against 1.3, and 9% slower at 8^3 tiles).
// MRC > 0 fixes the stride of the staged arrays at compile time (the launch's
// max_region must equal it, and its prop_stride kBlock): the y and z reads then
// take immediate LDS offsets from the x address instead of an address add each
// (measured 2% at 8^3 tiles), and so do the proposal reads.
// -3 w n of a lane's pairs (present only where the kernel keeps it)
template <int N, bool ON> struct PairTerm { double x[N], y[N], z[N]; };
template <int N> struct PairTerm<N, false> {};

template <int G, int NPL, bool MIXED, int BLK, int MRC = 0>
__global__ __launch_bounds__(BLK) void tile_sweep_kernel(SweepArgs a, int first_tile, int ntiles)
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
    const int region = RX * RY * RZ;                       // sites staged
    const int RZs = gather_zstride(RZ);                    // z stride in LDS (>= RZ)

    const int MR = MRC > 0 ? MRC : a.max_region;          // stride of the staged arrays
    // Uniform lengths: x, y and z arrays (the fixed-stride instances reach y
    // and z by immediate offsets from the x address).  Mixed lengths: (x, y)
    // pairs and (z, L) pairs, so two 16-byte reads bring a neighbour's
    // dipole and length (measured at DipoleCutOff 4, half vacant: three
    // 8-byte reads and a length read 2.49e9, (z, L) pairs 2.68e9, both pairs
    // 2.81e9 site-updates/s; the uniform instances lost 2% with (x, y) pairs).
    constexpr int ZS = MIXED ? 2 : 1;                      // doubles per site in the x, y and z arrays
    double *sx = smem;                                     // x of site i at sx[ZS * i]
    double *sy = MIXED ? sx + 1 : sx + MR;                 // y of site i at sy[ZS * i]
    double *sz = sx + 2 * MR;                              // z of site i at sz[ZS * i]
    double *sl = sz + 1;                                   // L of site i at sl[2 * i] (MIXED only)
    double *prop = sz + ZS * MR;                           // 4 x prop_stride: px, py, pz, threshold
    const int PS = MRC > 0 ? kBlock : a.prop_stride;        // (fixed-stride launches draw chunks of kBlock)
    int *s_ri = (int *)(prop + 4 * PS);                    // region index of class sites
    int *s_cls = s_ri + a.max_tile_sites;                  // class starts, then chunk colours
    int *s_chunk = s_cls + a.max_cls + 1;
    int *s_gchunk = s_chunk + a.max_chunks + 1;            // first step of each chunk
    int *s_gstep = s_gchunk + a.max_chunks + 1;            // step table
    constexpr int SW = BLK / G + 1;                        // ints per step

    // ---- shape tables into LDS ----
    const int ntile = sh.tx * sh.ty * sh.tz;
    for (int i = tid; i < ntile; i += BLK) s_ri[i] = a.class_ri[sh.sites_off + i];
    for (int i = tid; i <= sh.ncol; i += BLK) s_cls[i] = a.class_start[sh.class_off + i];
    for (int i = tid; i <= sh.nchunks; i += BLK) {
        s_chunk[i] = a.chunk_col[sh.chunk_off + i];
        s_gchunk[i] = a.gchunk[sh.gchunk_off + i];
    }
    for (int i = tid; i < s_gchunk_total(sh) * SW; i += BLK) s_gstep[i] = a.gsteps[sh.gstep_off + i];

    // ---- this lane's share of the neighbour pairs, held in registers (loaded
    // first: their latency overlaps the staging) ----
    const int part = tid % G;
    const int group = tid / G;
    int dl[NPL];
    double nx[NPL], ny[NPL], nz[NPL], w[NPL], w3[NPL];
#pragma unroll
    for (int m = 0; m < NPL; m++) {
        int j = part + m * G;
        if (j < a.npairs) {
            PairDev p = a.pairs[j];
            dl[m] = (p.dx * RY + p.dy) * RZs + p.dz;
            nx[m] = p.nx; ny[m] = p.ny; nz[m] = p.nz; w[m] = p.w; w3[m] = -3.0 * p.w;
        } else {
            dl[m] = 0; nx[m] = 0.0; ny[m] = 0.0; nz[m] = 0.0; w[m] = 0.0; w3[m] = 0.0;
        }
    }
    // TN: t = -3 w n per pair in registers, one multiply fewer per pair for
    // three more registers.  Only the fixed-stride instances take it: their
    // LDS (>= 66 KB) allows two workgroups per CU, so up to 256 VGPRs cost no
    // occupancy, while smaller tiles run three workgroups per CU (<= 168).
    constexpr bool TN = MRC > 0;
    PairTerm<NPL, TN> t3;
    if constexpr (TN) {
#pragma unroll
        for (int m = 0; m < NPL; m++) { t3.x[m] = w3[m] * nx[m]; t3.y[m] = w3[m] * ny[m]; t3.z[m] = w3[m] * nz[m]; }
    }
    // nearest-neighbour pairs come first in the table, so they sit in slot 0
    const double nnf = (part < a.npairs && a.pairs[part].nn) ? 1.0 : 0.0;

    // proposal and Metropolis threshold of the chunk's site k (region index ri)
    const unsigned sweep = a.sweep_ctr[tile.replica];
    const long long tile_base = ((long long)tile.x0 * rep.Y + tile.y0) * rep.Z + tile.z0;
    auto draw = [&](int k, int ri) __attribute__((always_inline)) {
        const int iz = ri % RZs, q = ri / RZs;
        const int iy = q % RY, ix = q / RY;
        const unsigned gsite = (unsigned)(tile_base + ((long long)(ix - hx) * rep.Y + (iy - hy)) * rep.Z + (iz - hz));
        double px, py, pz, u;
        counter_draw(rep.seed, (uint32_t)tile.replica, sweep, gsite, rep.constrain_to_x, px, py, pz, u);
        prop[k] = px;
        prop[PS + k] = py;
        prop[2 * PS + k] = pz;
        prop[3 * PS + k] = metropolis_threshold(u, rep.beta);
    };
    // the first chunk's proposals are drawn while the staging loads are in
    // flight (its site list is read from global memory: the LDS copy is not
    // ready before the barrier)
    const int n0 = a.class_start[sh.class_off + a.chunk_col[sh.chunk_off + 1]];
    const int ri0 = tid < n0 ? a.class_ri[sh.sites_off + tid] : 0;
    bool drawn = false;

    // ---- stage tile + halo (periodic wrap), 4 loads in flight per lane ----
    const double4 *lat = a.lat + rep.base;
    {
        // mixed-radix position of element tid, and the step of BLK elements
        int iz = tid % RZ, q = tid / RZ, iy = q % RY, ix = q / RY;
        const int sz_ = BLK % RZ, q2 = BLK / RZ, sy_ = q2 % RY, sx_ = q2 / RY;
        for (int base = tid; base < region; base += 4 * BLK) {
            double4 v[4];
            int idx[4], pidx[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                idx[k] = base + k * BLK;
                pidx[k] = RZs == RZ ? idx[k] : (ix * RY + iy) * RZs + iz;
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
            if (!drawn) {
                // after this batch's loads have issued (the fence keeps them
                // above it), before their data is needed
                drawn = true;
                int r0 = ri0;
                asm volatile("" : "+v"(r0) : : "memory");
                if (tid < n0) draw(tid, r0);
            }
#pragma unroll
            for (int k = 0; k < 4; k++) {
                if (idx[k] < region) {
                    if (MIXED) {
                        double2 xy, zl;
                        xy.x = v[k].x; xy.y = v[k].y; zl.x = v[k].z; zl.y = v[k].w;
                        ((double2 *)sx)[pidx[k]] = xy;
                        ((double2 *)sz)[pidx[k]] = zl;
                    } else {
                        sx[pidx[k]] = v[k].x; sy[pidx[k]] = v[k].y; sz[pidx[k]] = v[k].z;
                    }
                }
            }
        }
    }
    if (!drawn && tid < n0) draw(tid, ri0);   // lanes past a small region's end

    const double J = rep.J, K = rep.K, E0 = rep.E[0], E1 = rep.E[1], E2 = rep.E[2];
    // Uniform lengths: the cage-strain sum rides on the nearest-neighbour
    // pair's weight, h' = h - J / L0^2 g, so the partial is L0^2 dp.h' and no
    // separate g is accumulated.  Mixed lengths keep g (L_j sits inside h).
    const double wn0 = (!MIXED && nnf != 0.0 && rep.L0 != 0.0) ? w[0] - J / (rep.L0 * rep.L0) : w[0];
    unsigned nacc = 0, nrej = 0;
    __syncthreads();
    if (MIXED) {
        // vacant sites (length 0) leave the tile's step table: the reference
        // returns from MC_move for them before drawing or counting
        // (src/starrynight-montecarlo-core.c:163).  Each step's remaining
        // sites then move to its first lane groups (one lane per step), so
        // whole waves skip the step.  Measured on 128^3 DipoleCutOff 4, half
        // vacant, 512 threads: 1.89e9 -> 2.29e9 (dropping) -> 2.43e9 (packing)
        // site-updates/s.  Packing a colour's sites across its steps instead
        // measured 2.36e9.
        for (int i = tid; i < s_gchunk_total(sh) * SW; i += BLK) {
            const int e_ = s_gstep[i];
            if (i % SW != 0 && e_ >= 0 && sl[2 * (e_ & 0xFFFF)] == 0.0) s_gstep[i] = -1;
        }
        __syncthreads();
        for (int st_ = tid; st_ < s_gchunk_total(sh); st_ += BLK) {
            int *row = s_gstep + st_ * SW + 1;
            int k = 0;
            for (int g = 0; g < SW - 1; g++) {
                const int e_ = row[g];
                if (e_ >= 0) row[k++] = e_;
            }
            for (; k < SW - 1; k++) row[k] = -1;
        }
        __syncthreads();
    }

    for (int ch = 0; ch < sh.nchunks; ch++) {
        const int c0 = s_chunk[ch], c1 = s_chunk[ch + 1];
        const int s0 = s_cls[c0], s1 = s_cls[c1];
        // ---- proposals of this chunk's sites: one site per lane (the first
        // chunk's were drawn during staging) ----
        if (ch > 0 && tid < s1 - s0) draw(tid, s_ri[s0 + tid]);
        if (ch > 0) __syncthreads();
        // this chunk's steps: every group decides one site per step; a step
        // holds sites of one colour and a colour's last step ends with the
        // barrier.  The next step's entry and flag load with this step's data.
        const int st0 = s_gchunk[ch], st1 = s_gchunk[ch + 1];
        // (the LDS table has one spare row, so the read one past a chunk's last
        // step stays in the allocation; its value is never used)
        int e = s_gstep[st0 * SW + 1 + group];
        int flag = s_gstep[st0 * SW];
        for (int st = st0; st < st1; st++) {
            // a group without a site this step (a short colour) skips the step:
            // its lanes issue no LDS reads (the group's lanes are whole DPP rows,
            // so the reduction of the active groups is unaffected)
            const bool active = e >= 0;
            const int ri = e & 0xFFFF;
            const int pq = e >> 16;
            const int e_next = s_gstep[(st + 1) * SW + 1 + group];
            const int flag_next = s_gstep[(st + 1) * SW];
            if (active) {
                double hx_, hy_, hz_;
                double gx_, gy_, gz_;
                // every LDS read of the step is issued before the first use, so
                // the step waits for LDS latency once, not once per pair
                double ax[NPL], ay[NPL], az[NPL], bx[NPL], by[NPL], bz[NPL], la[NPL], lb[NPL];
                const double *cx = sx + ri, *cy = sy + ri, *cz = sz + ri;
                const double2 *cxy = (const double2 *)sx + ri, *czl = (const double2 *)sz + ri;
#pragma unroll
                for (int m = 0; m < NPL; m++) {
                    if (MIXED) {
                        const double2 pa = cxy[dl[m]], pb = cxy[-dl[m]];
                        const double2 za = czl[dl[m]], zb = czl[-dl[m]];
                        ax[m] = pa.x; ay[m] = pa.y; az[m] = za.x; la[m] = za.y;
                        bx[m] = pb.x; by[m] = pb.y; bz[m] = zb.x; lb[m] = zb.y;
                    } else {
                        ax[m] = cx[dl[m]]; ay[m] = cy[dl[m]]; az[m] = cz[dl[m]];
                        bx[m] = cx[-dl[m]]; by[m] = cy[-dl[m]]; bz[m] = cz[-dl[m]];
                    }
                }
                const double px = prop[pq], py = prop[PS + pq], pz = prop[2 * PS + pq];
                const double thr = prop[3 * PS + pq];
                const double ox = sx[ZS * ri], oy = sy[ZS * ri], oz = sz[ZS * ri];
                const double Li = MIXED ? sl[2 * ri] : rep.L0;
                asm volatile("" ::: "memory");
#pragma unroll
                for (int m = 0; m < NPL; m++) {
                    const double ux = ax[m] + bx[m], uy = ay[m] + by[m], uz = az[m] + bz[m];
                    double mx, my, mz;
                    if (MIXED) {
                        mx = la[m] * ax[m] + lb[m] * bx[m];
                        my = la[m] * ay[m] + lb[m] * by[m];
                        mz = la[m] * az[m] + lb[m] * bz[m];
                    } else {
                        mx = ux; my = uy; mz = uz;
                    }
                    const double s = nx[m] * mx + ny[m] * my + nz[m] * mz;
                    // the first pair starts the sums (no adds of +0.0); each
                    // component is two dependent FMAs
                    if constexpr (TN) {
                        if (m == 0) {
                            hx_ = fma(s, t3.x[m], wn0 * mx);
                            hy_ = fma(s, t3.y[m], wn0 * my);
                            hz_ = fma(s, t3.z[m], wn0 * mz);
                            if (MIXED) { gx_ = nnf * ux; gy_ = nnf * uy; gz_ = nnf * uz; }
                        } else {
                            hx_ = fma(s, t3.x[m], fma(w[m], mx, hx_));
                            hy_ = fma(s, t3.y[m], fma(w[m], my, hy_));
                            hz_ = fma(s, t3.z[m], fma(w[m], mz, hz_));
                        }
                        continue;
                    }
                    const double ts = w3[m] * s;
                    if (m == 0) {
                        hx_ = fma(ts, nx[m], wn0 * mx);
                        hy_ = fma(ts, ny[m], wn0 * my);
                        hz_ = fma(ts, nz[m], wn0 * mz);
                        if (MIXED) { gx_ = nnf * ux; gy_ = nnf * uy; gz_ = nnf * uz; }
                    } else {
                        hx_ = fma(ts, nx[m], fma(w[m], mx, hx_));
                        hy_ = fma(ts, ny[m], fma(w[m], my, hy_));
                        hz_ = fma(ts, nz[m], fma(w[m], mz, hz_));
                    }
                }
                const double coef = MIXED ? Li : Li * rep.L0;
                const double dx = px - ox, dy = py - oy, dz = pz - oz;
                double partial = coef * (dx * hx_ + dy * hy_ + dz * hz_);
                if (MIXED) partial -= J * (dx * gx_ + dy * gy_ + dz * gz_);
                partial = group_sum_dpp<G>(partial);
                double dE = partial + (dx * E0 + dy * E1 + dz * E2);
                if (K > 0.0) dE += -K * (fabs(px) - fabs(ox) + fabs(py) - fabs(oy));
                const bool acc = dE < thr;
                if (part == 0 && Li != 0.0) {
                    if (acc) {
                        sx[ZS * ri] = px; sy[ZS * ri] = py; sz[ZS * ri] = pz;
                        nacc++;
                    } else nrej++;
                }
            }
            if (flag) __syncthreads();
            e = e_next;
            flag = flag_next;
        }
    }

    // ---- write the tile back ----
    double4 *out = a.lat + rep.base;
    for (int i = tid; i < ntile; i += BLK) {
        int lz = i % sh.tz, q = i / sh.tz;
        int ly = q % sh.ty, lx = q / sh.ty;
        int ri = ((lx + hx) * RY + (ly + hy)) * RZs + (lz + hz);
        double4 v;
        v.x = sx[ZS * ri]; v.y = sy[ZS * ri]; v.z = sz[ZS * ri];
        v.w = MIXED ? sl[2 * ri] : rep.L0;
        out[tile_base + ((long long)lx * rep.Y + ly) * rep.Z + lz] = v;
    }
    // ---- counters: this tile's own slot (one workgroup per tile per launch) ----
    unsigned long long ca = nacc, cr = nrej;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        ca += __shfl_xor(ca, o, 64);
        cr += __shfl_xor(cr, o, 64);
    }
    __shared__ unsigned long long s_cnt[2 * (BLK / 64)];
    if ((tid & 63) == 0) { s_cnt[2 * (tid >> 6)] = ca; s_cnt[2 * (tid >> 6) + 1] = cr; }
    __syncthreads();
    if (tid == 0) {
        unsigned long long A_ = 0, R_ = 0;
        for (int w_ = 0; w_ < BLK / 64; w_++) { A_ += s_cnt[2 * w_]; R_ += s_cnt[2 * w_ + 1]; }
        a.counters[2 * (first_tile + t)] += A_;
        a.counters[2 * (first_tile + t) + 1] += R_;
    }
}

}  // namespace sn
