// sn_torus.h -- whole-lattice sweep of small replicas in one workgroup (gfx950, FP64).
//
// A replica small enough for LDS (its x, y, z (, length) arrays at once) is
// swept by one workgroup, sweep after sweep, inside one launch: no tiles, no
// halo staging, no launch per coarse colour.  The sites are coloured by a
// sublattice that holds the replica's box vectors and whose shortest vector is
// longer than DipoleCutOff, so the sites of one colour are more than rc apart
// through the periodic boundary and update together; colours follow one
// another with a workgroup barrier.  Neighbour indices wrap periodically on
// the fly (the reference's % X wrap, src/starrynight-montecarlo-core.c:97).
// Per site the arithmetic is the tile kernels': G lanes share a site, each
// lane takes NPL of the +-r pairs, one scalar per lane is reduced over the
// group, and the Metropolis threshold and proposal come from the same
// counter RNG draw of (seed, replica, sweep, site).
#pragma once
#include "sn_kernels.h"

namespace sn {

struct TorusArgs {
    double4 *lat;
    const RepDev *reps;
    const PairDev *pairs;
    int npairs;
    const int *sites;          // per replica (at site_off[r]): sites by colour, packed x | y << 10 | z << 20
    const int *site_off;       // R entries
    const int *col_start;      // per replica: ncol + 1 colour starts (at r * (ncol + 1))
    int ncol;
    const unsigned *sweep_ctr;
    unsigned long long *counters;  // 2 per replica: accepted, rejected (accumulated)
    int max_sites;             // stride of the staged arrays
    int prop_cap;              // proposal slots (>= the largest colour): colours are drawn in
                               // runs of at most this many sites
};

// (c + d) mod L for 0 <= c < L and |d| < L, without branches: of c + d and
// c + d + s (s = -L for d >= 0, +L for d < 0) one lies in [0, L) and the other
// is negative or >= L, so as unsigned values the smaller one is the answer.
__device__ __forceinline__ int wrap_pm(int t, int s) { return (int)min((unsigned)t, (unsigned)(t + s)); }

template <int G, int NPL, bool MIXED, int BLK>
__global__ __launch_bounds__(BLK) void torus_sweep_kernel(TorusArgs t, int nsweeps)
{
    extern __shared__ double smem[];
    const int tid = threadIdx.x, r = blockIdx.x;
    const RepDev rep = t.reps[r];
    const int X = rep.X, Y = rep.Y, Z = rep.Z, N = X * Y * Z, YZ = Y * Z;   // N < 2^24 (LDS)
    const int MS = t.max_sites, PS = t.prop_cap;
    double *sx = smem, *sy = sx + MS, *sz = sy + MS, *sl = sz + MS;
    double *prop = MIXED ? sl + MS : sl;    // 4 x PS: px, py, pz, threshold of the colour's sites
    int *sites = (int *)(prop + 4 * PS);    // the replica's site table (MS ints)
    const int *cs = t.col_start + r * (t.ncol + 1);   // colour starts (uniform: scalar loads)
    const int ncol = t.ncol;
    double4 *lat = t.lat + rep.base;

    // ---- the whole replica and its site table into LDS ----
    const int *gsites = t.sites + t.site_off[r];
    for (int i = tid; i < N; i += BLK) {
        const double4 v = lat[i];
        sx[i] = v.x; sy[i] = v.y; sz[i] = v.z;
        if (MIXED) sl[i] = v.w;
        sites[i] = gsites[i];
    }

    // ---- this lane's share of the +-r pairs ----
    const int part = tid % G, group = tid / G;
    constexpr int NG = BLK / G;
    int pdx[NPL], pdy[NPL], pdz[NPL], swx[NPL], swy[NPL], swz[NPL];   // offsets, wrap steps of +d
    double nx[NPL], ny[NPL], nz[NPL], w[NPL], w3[NPL];
#pragma unroll
    for (int m = 0; m < NPL; m++) {
        const int j = part + m * G;
        if (j < t.npairs) {
            const PairDev p = t.pairs[j];
            pdx[m] = p.dx; pdy[m] = p.dy; pdz[m] = p.dz;
            swx[m] = p.dx >= 0 ? -X : X; swy[m] = p.dy >= 0 ? -Y : Y; swz[m] = p.dz >= 0 ? -Z : Z;
            nx[m] = p.nx; ny[m] = p.ny; nz[m] = p.nz; w[m] = p.w; w3[m] = -3.0 * p.w;
        } else {
            pdx[m] = 0; pdy[m] = 0; pdz[m] = 0; swx[m] = -X; swy[m] = -Y; swz[m] = -Z;
            nx[m] = 0.0; ny[m] = 0.0; nz[m] = 0.0; w[m] = 0.0; w3[m] = 0.0;
        }
    }
    // nearest-neighbour pairs come first in the table, so they sit in slot 0
    const double nnf = (part < t.npairs && t.pairs[part].nn) ? 1.0 : 0.0;
    const double J = rep.J, K = rep.K, E0 = rep.E[0], E1 = rep.E[1], E2 = rep.E[2];
    // uniform lengths: the cage-strain sum rides on the nearest-neighbour weight
    const double wn0 = (!MIXED && nnf != 0.0 && rep.L0 != 0.0) ? w[0] - J / (rep.L0 * rep.L0) : w[0];
    unsigned nacc = 0, nrej = 0;
    __syncthreads();

    // One round: the group's site of colour k starting at site b (its entry
    // q = b + group, active while q < the colour's end).  The site table never
    // changes, so with LOOKAHEAD a round's site and its neighbours' indices
    // are found during the previous round, and after a colour's barrier a
    // round goes straight to its LDS reads.  Measured (site-updates/s, 64 or
    // 256 replicas): +20-26% at 4 pairs a lane (12^3 at DipoleCutOff 1-3, 2D
    // 30 x 30 at 2, 4, 5), -3 to -5% at 1-2 pairs a lane (25 x 25, 16 x 16 x 8
    // at DipoleCutOff 3, 12^3 at 4), which find their site after the barrier.
    constexpr bool LOOKAHEAD = NPL >= 4;
    struct Round {
        bool act;
        int si, pq, ia[NPL], ib[NPL];
    };
    // the site entry of the round at (k, b) (the site table never changes)
    auto entry = [&](int k, int b) __attribute__((always_inline)) {
        const int q = b + group;
        return q < cs[k + 1] ? sites[q] : -1;
    };
    auto prepare = [&](int pk, int pq) __attribute__((always_inline)) {
        Round R;
        R.act = pk >= 0;
        const int x = pk & 1023, y = (pk >> 10) & 1023, z = (pk >> 20) & 1023;
        R.si = __umul24(x, YZ) + __umul24(y, Z) + z;
#pragma unroll
        for (int m = 0; m < NPL; m++) {
            // periodic wrap of x +- dx etc. (|d| <= rc < length; -d wraps
            // the other way, so its step is -s)
            const int xa = wrap_pm(x + pdx[m], swx[m]), xb = wrap_pm(x - pdx[m], -swx[m]);
            const int ya = wrap_pm(y + pdy[m], swy[m]), yb = wrap_pm(y - pdy[m], -swy[m]);
            const int za = wrap_pm(z + pdz[m], swz[m]), zb = wrap_pm(z - pdz[m], -swz[m]);
            R.ia[m] = __umul24(xa, YZ) + __umul24(ya, Z) + za;
            R.ib[m] = __umul24(xb, YZ) + __umul24(yb, Z) + zb;
        }
        R.pq = pq;
        return R;
    };

    for (int s = 0; s < nsweeps; s++) {
        const unsigned sweep = t.sweep_ctr[r] + (unsigned)s;
        for (int k0 = 0; k0 < ncol;) {
            // ---- proposals of a run of colours [k0, k1) holding at most PS sites ----
            int k1 = k0 + 1;
            while (k1 < ncol && cs[k1 + 1] - cs[k0] <= PS) k1++;
            const int p0 = cs[k0];
            for (int i = tid; i < cs[k1] - p0; i += BLK) {
                double px, py, pz, u;
                const int pk = sites[p0 + i];
                const int si = __umul24(pk & 1023, YZ) + __umul24((pk >> 10) & 1023, Z) + (pk >> 20);
                counter_draw(rep.seed, (uint32_t)r, sweep, (unsigned)si, rep.constrain_to_x, px, py, pz, u);
                prop[i] = px;
                prop[PS + i] = py;
                prop[2 * PS + i] = pz;
                prop[3 * PS + i] = metropolis_threshold(u, rep.beta);
            }
            __syncthreads();
            // ---- the run's colours, NG sites at a time (a colour's sites are
            // independent of each other); a barrier ends each colour ----
            // (empty colours, the padding of replicas with fewer colours, come
            // after all of a replica's sites: a lookahead that reaches one
            // finds no site, and neither do the colours after it)
            Round cur;
            if constexpr (LOOKAHEAD) cur = prepare(entry(k0, p0), group);
            for (int k = k0; k < k1; k++) {
                const int c1 = cs[k + 1];
                for (int b0 = cs[k]; b0 < c1; b0 += NG) {
                    const bool last = b0 + NG >= c1;           // the colour's last round
                    const int kn = last ? k + 1 : k, bn = last ? c1 : b0 + NG;
                    int pk_next = -1;
                    if constexpr (LOOKAHEAD) pk_next = kn < k1 ? entry(kn, bn) : -1;
                    else cur = prepare(entry(k, b0), b0 + group - p0);
                    if (cur.act) {   // whole groups (DPP rows) are active or not
                        double hx_, hy_, hz_, gx_ = 0.0, gy_ = 0.0, gz_ = 0.0;
                        double ax[NPL], ay[NPL], az[NPL], bx[NPL], by[NPL], bz[NPL], la[NPL], lb[NPL];
#pragma unroll
                        for (int m = 0; m < NPL; m++) {
                            const int ia = cur.ia[m], ib = cur.ib[m];
                            ax[m] = sx[ia]; ay[m] = sy[ia]; az[m] = sz[ia];
                            bx[m] = sx[ib]; by[m] = sy[ib]; bz[m] = sz[ib];
                            if (MIXED) { la[m] = sl[ia]; lb[m] = sl[ib]; }
                        }
                        const int pq = cur.pq, si = cur.si;
                        const double px = prop[pq], py = prop[PS + pq], pz = prop[2 * PS + pq];
                        const double thr = prop[3 * PS + pq];
                        const double ox = sx[si], oy = sy[si], oz = sz[si];
                        const double Li = MIXED ? sl[si] : rep.L0;
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
                            const double sdot = nx[m] * mx + ny[m] * my + nz[m] * mz;
                            const double ts = w3[m] * sdot;
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
                        if (part == 0 && Li != 0.0) {   // vacancies never move
                            const bool acc = dE < thr;
                            if (acc) { sx[si] = px; sy[si] = py; sz[si] = pz; }
                            nacc += acc ? 1u : 0u;
                            nrej += acc ? 0u : 1u;
                        }
                    }
                    if constexpr (LOOKAHEAD) cur = prepare(pk_next, bn + group - p0);
                    if (last) __syncthreads();   // the colour is done before the next one reads it
                }
            }
            k0 = k1;
        }
    }

    // ---- write back and count ----
    for (int i = tid; i < N; i += BLK) {
        double4 v;
        v.x = sx[i]; v.y = sy[i]; v.z = sz[i];
        v.w = MIXED ? sl[i] : rep.L0;
        lat[i] = v;
    }
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
        t.counters[2 * r] += A_;
        t.counters[2 * r + 1] += R_;
    }
}

}  // namespace sn
