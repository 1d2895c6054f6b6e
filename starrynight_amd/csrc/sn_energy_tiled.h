// sn_energy_tiled.h -- energy and adjoint over LDS-staged tiles (gfx950, FP64).
//
// The Hamiltonian whose single-site differences are the reference's
// site_energy (src/starrynight-montecarlo-core.c:91-134):
//   E = 1/2 sum_i L_i p_i . h_i - J/2 sum_i p_i . g_i + sum_i p_i . E - [K>0] K sum_i (|p_ix| + |p_iy|)
// with h_i = sum_r L_j (p_j - 3 n (n . p_j)) / d^3 and g_i = sum_{|r|=1} p_j.
// The adjoint is dE/dp_i = L_i h_i - J g_i + E - [K>0] K (sgn p_ix, sgn p_iy, 0)
// (the local field is its negative), and
//   dE/dJ = -1/2 sum p.g,  dE/dE = sum p,  dE/ds = 2 E_dd (lengths scaled by s),
//   dE/dK = -sum (|p_x| + |p_y|)  (K > 0).
//
// One workgroup per tile of `tile_list` (every tile of one replica, or of all
// replicas, in one launch: nothing is written to the lattice, so tiles of all
// coarse colours run together).
// The tile and its halo are staged in LDS with the field sweep's padded
// layout, and every tile site's h and g come from the same compile-time
// unrolled column stencil (tile_fields), two z-adjacent sites per lane.
// Per-workgroup partial sums of the energy parts and parameter derivatives
// go to `partials` (kEnergyParts per workgroup, reduced afterwards).
#pragma once
#include "sn_sweep_field.h"

namespace sn {

constexpr int kEnergyParts = 10;   // {E_dd, E_cage, E_field, E_K, dJ, dEx, dEy, dEz, ds, dK}

struct EnergyArgs {
    const double4 *lat;
    const RepDev *reps;
    const TileDev *tiles;
    const ShapeDev *shapes;
    const int *tile_list;     // the launch's tiles (indices into `tiles`), grouped by replica
    double *grad;             // 3 per site (row rep.base - grad_origin + site) or null
    long long grad_origin;    // lattice row of grad's first row
    double *partials;         // kEnergyParts per workgroup
    int region_doubles;       // stride between the staged x, y, z (, length) arrays
};

template <int RC, bool TWO_D, bool MIXED>
__global__ __launch_bounds__(kBlock) void tile_energy_kernel(EnergyArgs ea)
{
    constexpr int HZ = TWO_D ? 0 : RC;
    constexpr int S = TWO_D ? 1 : 2;
    extern __shared__ double smem[];
    __shared__ double red[kEnergyParts][kBlock / 64];
    const int tid = threadIdx.x;
    const TileDev tile = ea.tiles[ea.tile_list[blockIdx.x]];
    const RepDev rep = ea.reps[tile.replica];
    const ShapeDev sh = ea.shapes[tile.shape];
    const int RY = sh.ty + 2 * RC, RZ = sh.tz + 2 * HZ, RX = sh.tx + 2 * RC;
    const int region = RX * RY * RZ;
    const int RZp = TWO_D ? RZ : (RZ | 1);
    const int sY = RZp, sX = RY * RZp;
    const int MR = ea.region_doubles;
    double *sx = smem, *sy = sx + MR, *sz = sy + MR, *sl = sz + MR;

    // ---- stage tile + halo (periodic wrap), 4 loads in flight per lane ----
    const double4 *lat = ea.lat + rep.base;
    {
        int iz = tid % RZ, q = tid / RZ, iy = q % RY, ix = q / RY;
        const int dz_ = kBlock % RZ, q2 = kBlock / RZ, dy_ = q2 % RY, dx_ = q2 / RY;
        for (int base = tid; base < region; base += 4 * kBlock) {
            double4 v[4];
            int idx[4], pidx[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                idx[k] = base + k * kBlock;
                pidx[k] = (ix * RY + iy) * RZp + iz;
                if (idx[k] < region) {
                    const int gx = wrap_once(tile.x0 - RC + ix, rep.X);
                    const int gy = wrap_once(tile.y0 - RC + iy, rep.Y);
                    const int gz = wrap_once(tile.z0 - HZ + iz, rep.Z);
                    v[k] = lat[((long long)gx * rep.Y + gy) * rep.Z + gz];
                }
                iz += dz_; iy += dy_; ix += dx_;
                if (iz >= RZ) { iz -= RZ; iy++; }
                if (iy >= RY) { iy -= RY; ix++; }
            }
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (idx[k] < region) {
                    const int d = pidx[k];
                    sx[d] = v[k].x; sy[d] = v[k].y; sz[d] = v[k].z;
                    if (MIXED) sl[d] = v[k].w;
                }
        }
    }
    __syncthreads();

    // ---- fields, gradient and partial sums of every tile site ----
    const double J = rep.J, K = rep.K;
    const double Lh = MIXED ? 1.0 : rep.L0;   // h without the L_j factor when lengths are uniform
    double v[kEnergyParts];
#pragma unroll
    for (int c = 0; c < kEnergyParts; c++) v[c] = 0.0;
    auto site = [&](const Acc &A, const double p[4], int lx, int ly, int lz) __attribute__((always_inline)) {
        const double hx = Lh * A.hx, hy = Lh * A.hy, hz = Lh * A.hz;
        const double Li = MIXED ? p[3] : rep.L0;
        const double ph = p[0] * hx + p[1] * hy + p[2] * hz;
        const double pg = p[0] * A.gx + p[1] * A.gy + p[2] * A.gz;
        double gx = Li * hx - J * A.gx + rep.E[0];
        double gy = Li * hy - J * A.gy + rep.E[1];
        const double gz = Li * hz - J * A.gz + rep.E[2];
        const double ak = fabs(p[0]) + fabs(p[1]);
        if (K > 0.0) {
            gx -= K * ((p[0] > 0.0) - (p[0] < 0.0));
            gy -= K * ((p[1] > 0.0) - (p[1] < 0.0));
        }
        if (ea.grad) {
            const long long s = rep.base - ea.grad_origin +
                                ((long long)(tile.x0 + lx) * rep.Y + (tile.y0 + ly)) * rep.Z + (tile.z0 + lz);
            ea.grad[3 * s] = gx; ea.grad[3 * s + 1] = gy; ea.grad[3 * s + 2] = gz;
        }
        v[0] += 0.5 * Li * ph;
        v[1] += -0.5 * J * pg;
        v[2] += p[0] * rep.E[0] + p[1] * rep.E[1] + p[2] * rep.E[2];
        v[3] += K > 0.0 ? -K * ak : 0.0;
        v[4] += -0.5 * pg;
        v[5] += p[0]; v[6] += p[1]; v[7] += p[2];
        v[8] += Li * ph;
        v[9] += K > 0.0 ? -ak : 0.0;
    };
    const int tzp = (sh.tz + S - 1) / S;
    const int nitems = sh.tx * sh.ty * tzp;
    for (int item = tid; item < nitems; item += kBlock) {
        const int lx = item % sh.tx, q = item / sh.tx, ly = q % sh.ty, zp = q / sh.ty;
        const int lz = zp * S;
        const int riA = ((lx + RC) * RY + (ly + RC)) * RZp + (lz + HZ);
        Acc A, B;
        double pA[4], pB[4];
        tile_fields<RC, TWO_D, MIXED>(sx, sy, sz, sl, riA, sX, sY, A, B, pA, pB);
        site(A, pA, lx, ly, lz);
        if (S == 2 && lz + 1 < sh.tz) site(B, pB, lx, ly, lz + 1);
    }

    // ---- workgroup partial sums: wave shuffles, then the waves through LDS ----
#pragma unroll
    for (int c = 0; c < kEnergyParts; c++) {
        double x = v[c];
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) x += __shfl_xor(x, o, 64);
        if ((tid & 63) == 0) red[c][tid >> 6] = x;
    }
    __syncthreads();
    if (tid < kEnergyParts) {
        double x = 0.0;
        for (int w = 0; w < kBlock / 64; w++) x += red[tid][w];
        ea.partials[(long long)blockIdx.x * kEnergyParts + tid] = x;
    }
}

}  // namespace sn
