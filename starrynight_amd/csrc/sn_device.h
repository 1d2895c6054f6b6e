// sn_device.h -- device-side building blocks shared by the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sn {

// Per-replica parameters, resident in HBM so that sn_set_params reaches a
// captured graph without re-capture.
struct RepDev {
    int X, Y, Z, mixed;
    long long base;        // first site of this replica in the lattice buffer
    double beta, J, K, L0; // L0: the common dipole length when !mixed
    double E[3];
    unsigned long long seed;
    int constrain_to_x, pad;
};

struct TileDev {
    int replica, x0, y0, z0, shape, pad0, pad1, pad2;
};

struct ShapeDev {
    int tx, ty, tz, ncol;
    int nchunks, class_off, chunk_off, sites_off;
    int nsteps, step_off;     // field kernel's step table (see sn_sweep_field.h)
    int typ, tzp, coord_off;  // field kernel's padded tile-order strides and coordinate table
    int gstep_off, gchunk_off, gsteps_n;   // gather kernel's step table (see sn_kernels.h)
};

// z stride of the gather kernel's staged region.  A row of a multiple of 8
// sites (16 for 8-site tiles at DipoleCutOff 4) is padded by one: with a
// stride of 16 doubles every neighbour at the same dz lands in the same LDS
// bank, whatever its dx and dy.
__host__ __device__ inline int gather_zstride(int rz) { return (rz > 1 && rz % 8 == 0) ? rz + 1 : rz; }

// One +-r neighbour pair (sweep kernel) or one offset (gather kernels).
struct PairDev {
    int dx, dy, dz, nn;
    double nx, ny, nz, w;  // unit vector r/|r| and 1/|r|^3
};

// Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11).  Each round's two
// 32 x 32 -> 64-bit products are formed as 64-bit multiplies (one
// v_mad_u64_u32 each) rather than separate high and low multiplies.
__device__ __forceinline__ void philox4x32_10(uint32_t c[4], uint32_t k0, uint32_t k1)
{
#pragma unroll
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0], p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0; c[1] = lo1; c[2] = n2; c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// Proposal and acceptance uniform for one visit of `site` in `sweep`:
// counter (site, sweep, replica, k), key = seed.  u = 53-bit uniform from words
// 2, 3 of k = 0; Marsaglia (1972) disc point from words 0, 1 of k = 0, then
// words 0, 1 and 2, 3 of k = 1, 2, ... (two candidates per retry call: a
// wave waits for its slowest lane, so fewer calls per rejection shorten every
// draw), with x = a 2^-26, a odd: the disc test is exact in integers and
// 1 - s, 1 - 2 s are exact, so only one sqrt and two products round.
__device__ __forceinline__ void counter_draw(unsigned long long seed, uint32_t replica,
                                             uint32_t sweep, uint32_t site, int constrain_to_x,
                                             double &px, double &py, double &pz, double &u)
{
    const uint32_t k0 = (uint32_t)seed, k1 = (uint32_t)(seed >> 32);
    uint32_t w[4] = {site, sweep, replica, 0u};
    philox4x32_10(w, k0, k1);
    u = ((double)(w[2] >> 5) * 67108864.0 + (double)(w[3] >> 6)) * (1.0 / 9007199254740992.0);
    if (constrain_to_x) {
        uint32_t i = w[0] % 6u;
        px = (i == 0) ? 1.0 : (i == 1) ? -1.0 : 0.0;
        py = (i == 2) ? 1.0 : (i == 3) ? -1.0 : 0.0;
        pz = (i == 4) ? 1.0 : (i == 5) ? -1.0 : 0.0;
        return;
    }
    // the point of candidate words (a0, a1) if it lies in the disc
    auto disc = [&](uint32_t a0, uint32_t a1) -> bool {
        const long long a = 2 * (long long)(a0 >> 6) + 1 - (1ll << 26);
        const long long b = 2 * (long long)(a1 >> 6) + 1 - (1ll << 26);
        const long long s = a * a + b * b;
        if (s >= (1ll << 52)) return false;
        const double x1 = (double)a * (1.0 / 67108864.0);
        const double x2 = (double)b * (1.0 / 67108864.0);
        const double S = (double)s * (1.0 / 4503599627370496.0);
        const double r = __builtin_sqrt(1.0 - S);
        px = (2.0 * x1) * r;
        py = (2.0 * x2) * r;
        pz = 1.0 - 2.0 * S;
        return true;
    };
    if (disc(w[0], w[1])) return;
    for (uint32_t k = 1; k <= 32; k++) {
        uint32_t c[4] = {site, sweep, replica, k};
        philox4x32_10(c, k0, k1);
        if (disc(c[0], c[1]) || disc(c[2], c[3])) return;
    }
    px = 0.0; py = 0.0; pz = 1.0;   // 65 rejections in a row: p ~ 1e-43
}

// Metropolis threshold: exp(-beta dE) > u  <=>  dE < -ln(u)/beta, and the
// dE < 0 branch of src/starrynight-montecarlo-core.c:179 is covered because
// the threshold is >= 0.  beta = +inf (T = 0) gives 0: only dE < 0 accepted.
__device__ __forceinline__ double metropolis_threshold(double u, double beta)
{
    if (u <= 0.0 || beta <= 0.0) return __builtin_inf();
    if (__builtin_isinf(beta)) return 0.0;
    return -log(u) / beta;
}

__device__ __forceinline__ int wrap_once(int v, int L)
{
    v += (v < 0) ? L : 0;
    v -= (v >= L) ? L : 0;
    return v;
}

template <int G>
__device__ __forceinline__ double group_sum(double v)
{
#pragma unroll
    for (int o = 1; o < G; o <<= 1) v += __shfl_xor(v, o, 64);
    return v;
}

}  // namespace sn
