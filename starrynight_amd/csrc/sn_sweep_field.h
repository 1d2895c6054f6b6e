// sn_sweep_field.h -- field-update tile sweep (the production sweep kernel).
//
// One workgroup owns one tile for one coarse colour.  Instead of gathering the
// 122 (rc = 3) neighbours of every site again in every fine phase, it
//   1. stages tile + halo in LDS and computes the local field of every tile
//      site once, all sites in parallel: a compile-time unrolled stencil over
//      the +-r pairs (constants folded, zero components dropped), two sites per
//      lane along z sharing every column load;
//   2. runs the fine colours: each site of the colour decides from its field
//      (dE = (p' - p) . f_i, the reference's site_energy with the field
//      factored out, src/starrynight-montecarlo-core.c:91-134) and, when it
//      accepts, adds T_r (p' - p) to the field of every tile site within the
//      cutoff (LDS fp64 atomics); halo sites are frozen for the whole coarse
//      colour, so their contribution never changes;
//   3. writes the tile back.
// Per site: f_i = c_i h_i - J g_i + E with h_i = sum_r T_r m_{i+r},
// T_r = (I - 3 n n^T) / d^3, m_j = L_j p_j (or p_j with c_i = L0^2 when every
// length equals L0), g_i = sum_{|r|=1} p_{i+r}.
#pragma once
#include <utility>

#include "sn_kernels.h"

namespace sn {

constexpr int cisqrt(int v)
{
    int r = 0;
    while ((r + 1) * (r + 1) <= v) r++;
    return r;
}

constexpr double csqrt(double x)
{
    double r = x > 1.0 ? x : 1.0;
    for (int i = 0; i < 64; i++) r = 0.5 * (r + x / r);
    return r;
}

template <int Lo, typename F, int... I>
__device__ __forceinline__ void static_for_impl(F &&f, std::integer_sequence<int, I...>)
{
    (f(std::integral_constant<int, Lo + I>{}), ...);
}
// calls f(integral_constant<int, k>) for k = Lo .. Hi (inclusive); nothing when Hi < Lo
template <int Lo, int Hi, typename F>
__device__ __forceinline__ void static_for(F &&f)
{
    if constexpr (Hi >= Lo) static_for_impl<Lo>(f, std::make_integer_sequence<int, Hi - Lo + 1>{});
}

struct Acc {
    double hx, hy, hz, gx, gy, gz;
};

__device__ __forceinline__ void acc_zero(Acc &a) { a.hx = a.hy = a.hz = a.gx = a.gy = a.gz = 0.0; }

// Column boundary: the accumulators pass through an empty asm (and memory is
// clobbered), so one column's loads and arithmetic finish before the next
// column starts.  Without it the scheduler evaluates every column's loads and
// independent arithmetic first and runs the accumulation chains last, which
// keeps hundreds of values live and spills.
__device__ __forceinline__ void column_fence(Acc &a, Acc &b)
{
    asm volatile("" : "+v"(a.hx), "+v"(a.hy), "+v"(a.hz), "+v"(a.gx), "+v"(a.gy), "+v"(a.gz),
                 "+v"(b.hx), "+v"(b.hy), "+v"(b.hz), "+v"(b.gx), "+v"(b.gy), "+v"(b.gz)
                 :
                 : "memory");
}

// h += T_r M for the offset r = (DX, DY, DZ), every constant folded.
template <int DX, int DY, int DZ>
__device__ __forceinline__ void apply_T(double mx, double my, double mz, Acc &a)
{
    constexpr int r2 = DX * DX + DY * DY + DZ * DZ;
    constexpr double d = csqrt((double)r2);
    constexpr double w = 1.0 / (d * d * d);
    constexpr double nx = DX / d, ny = DY / d, nz = DZ / d;
    constexpr double m3w = -3.0 * w;
    double s;
    if constexpr (DX != 0) {
        s = nx * mx;
        if constexpr (DY != 0) s = fma(ny, my, s);
        if constexpr (DZ != 0) s = fma(nz, mz, s);
    } else if constexpr (DY != 0) {
        s = ny * my;
        if constexpr (DZ != 0) s = fma(nz, mz, s);
    } else {
        s = nz * mz;
    }
    const double t = m3w * s;
    a.hx = fma(w, mx, a.hx);
    if constexpr (DX != 0) a.hx = fma(t, nx, a.hx);
    a.hy = fma(w, my, a.hy);
    if constexpr (DY != 0) a.hy = fma(t, ny, a.hy);
    a.hz = fma(w, mz, a.hz);
    if constexpr (DZ != 0) a.hz = fma(t, nz, a.hz);
}

// Pair (+r, -r) contribution to one site: M = m(+r) + m(-r) (or the
// length-weighted sum), nearest neighbours also add the plain sum to g.
template <int DX, int DY, int DZ, bool MIXED>
__device__ __forceinline__ void pair_term(double ax, double ay, double az, double al, double bx, double by,
                                          double bz, double bl, Acc &acc)
{
    double mx, my, mz;
    if constexpr (MIXED) {
        mx = fma(al, ax, bl * bx); my = fma(al, ay, bl * by); mz = fma(al, az, bl * bz);
    } else {
        mx = ax + bx; my = ay + by; mz = az + bz;
    }
    apply_T<DX, DY, DZ>(mx, my, mz, acc);
    if constexpr (DX * DX + DY * DY + DZ * DZ == 1) {
        if constexpr (MIXED) { acc.gx += ax + bx; acc.gy += ay + by; acc.gz += az + bz; }
        else { acc.gx += mx; acc.gy += my; acc.gz += mz; }
    }
}

// 3D: the column pair (+DX,+DY) / (-DX,-DY) for sites A = (x, y, z) and
// B = (x, y, z + 1); column offsets run along z with stride 1 in the region.
template <int RC, int DX, int DY, bool MIXED>
__device__ __forceinline__ void column_pair_3d(const double *sx, const double *sy, const double *sz,
                                               const double *sl, int riA, int sX, int sY, Acc &A, Acc &B)
{
    constexpr int H = cisqrt(RC * RC - DX * DX - DY * DY);
    constexpr int NZ = 2 * H + 2;  // dz = -H .. H + 1
    const int P0 = riA + DX * sX + DY * sY - H;
    const int Q0 = riA - DX * sX - DY * sY - H;
    double px[NZ], py[NZ], pz[NZ], qx[NZ], qy[NZ], qz[NZ], pl[NZ], ql[NZ];
#pragma unroll
    for (int k = 0; k < NZ; k++) {
        px[k] = sx[P0 + k]; py[k] = sy[P0 + k]; pz[k] = sz[P0 + k];
        qx[k] = sx[Q0 + k]; qy[k] = sy[Q0 + k]; qz[k] = sz[Q0 + k];
        if constexpr (MIXED) { pl[k] = sl[P0 + k]; ql[k] = sl[Q0 + k]; }
        else { pl[k] = 1.0; ql[k] = 1.0; }
    }
    static_for<-H, H>([&](auto dzc) __attribute__((always_inline)) {
        constexpr int DZ = decltype(dzc)::value;
        constexpr int iA = DZ + H, jA = -DZ + H;          // P[DZ], Q[-DZ]
        constexpr int iB = DZ + 1 + H, jB = 1 - DZ + H;   // P[DZ+1], Q[1-DZ]
        pair_term<DX, DY, DZ, MIXED>(px[iA], py[iA], pz[iA], pl[iA], qx[jA], qy[jA], qz[jA], ql[jA], A);
        pair_term<DX, DY, DZ, MIXED>(px[iB], py[iB], pz[iB], pl[iB], qx[jB], qy[jB], qz[jB], ql[jB], B);
    });
}

// 3D centre column (0, 0, dz), dz = 1..RC, for A and B; also returns p_A, p_B.
template <int RC, bool MIXED>
__device__ __forceinline__ void centre_column_3d(const double *sx, const double *sy, const double *sz,
                                                 const double *sl, int riA, Acc &A, Acc &B, double pA[4],
                                                 double pB[4])
{
    constexpr int NZ = 2 * RC + 2;
    const int C0 = riA - RC;
    double cx[NZ], cy[NZ], cz[NZ], cl[NZ];
#pragma unroll
    for (int k = 0; k < NZ; k++) {
        cx[k] = sx[C0 + k]; cy[k] = sy[C0 + k]; cz[k] = sz[C0 + k];
        cl[k] = MIXED ? sl[C0 + k] : 1.0;
    }
    static_for<1, RC>([&](auto dzc) __attribute__((always_inline)) {
        constexpr int DZ = decltype(dzc)::value;
        constexpr int iA = DZ + RC, jA = -DZ + RC, iB = DZ + 1 + RC, jB = 1 - DZ + RC;
        pair_term<0, 0, DZ, MIXED>(cx[iA], cy[iA], cz[iA], cl[iA], cx[jA], cy[jA], cz[jA], cl[jA], A);
        pair_term<0, 0, DZ, MIXED>(cx[iB], cy[iB], cz[iB], cl[iB], cx[jB], cy[jB], cz[jB], cl[jB], B);
    });
    pA[0] = cx[RC]; pA[1] = cy[RC]; pA[2] = cz[RC]; pA[3] = cl[RC];
    pB[0] = cx[RC + 1]; pB[1] = cy[RC + 1]; pB[2] = cz[RC + 1]; pB[3] = cl[RC + 1];
}

// 2D (Z == 1): one site per item; columns run along y (stride 1).
template <int RC, int DX, bool MIXED>
__device__ __forceinline__ void column_pair_2d(const double *sx, const double *sy, const double *sz,
                                               const double *sl, int ri, int sX, Acc &A)
{
    constexpr int H = cisqrt(RC * RC - DX * DX);
    constexpr int NY = 2 * H + 1;
    const int P0 = ri + DX * sX - H, Q0 = ri - DX * sX - H;
    double px[NY], py[NY], pz[NY], qx[NY], qy[NY], qz[NY], pl[NY], ql[NY];
#pragma unroll
    for (int k = 0; k < NY; k++) {
        px[k] = sx[P0 + k]; py[k] = sy[P0 + k]; pz[k] = sz[P0 + k];
        qx[k] = sx[Q0 + k]; qy[k] = sy[Q0 + k]; qz[k] = sz[Q0 + k];
        pl[k] = MIXED ? sl[P0 + k] : 1.0;
        ql[k] = MIXED ? sl[Q0 + k] : 1.0;
    }
    static_for<-H, H>([&](auto dyc) __attribute__((always_inline)) {
        constexpr int DY = decltype(dyc)::value;
        constexpr int i = DY + H, j = -DY + H;
        pair_term<DX, DY, 0, MIXED>(px[i], py[i], pz[i], pl[i], qx[j], qy[j], qz[j], ql[j], A);
    });
}

template <int RC, bool MIXED>
__device__ __forceinline__ void centre_column_2d(const double *sx, const double *sy, const double *sz,
                                                 const double *sl, int ri, Acc &A, double pA[4])
{
    constexpr int NY = 2 * RC + 1;
    const int C0 = ri - RC;
    double cx[NY], cy[NY], cz[NY], cl[NY];
#pragma unroll
    for (int k = 0; k < NY; k++) {
        cx[k] = sx[C0 + k]; cy[k] = sy[C0 + k]; cz[k] = sz[C0 + k];
        cl[k] = MIXED ? sl[C0 + k] : 1.0;
    }
    static_for<1, RC>([&](auto dyc) __attribute__((always_inline)) {
        constexpr int DY = decltype(dyc)::value;
        constexpr int i = DY + RC, j = -DY + RC;
        pair_term<0, DY, 0, MIXED>(cx[i], cy[i], cz[i], cl[i], cx[j], cy[j], cz[j], cl[j], A);
    });
    pA[0] = cx[RC]; pA[1] = cy[RC]; pA[2] = cz[RC]; pA[3] = cl[RC];
}

// Local fields of sites A = region index riA and (3D) B = riA + 1.
template <int RC, bool TWO_D, bool MIXED>
__device__ __forceinline__ void tile_fields(const double *sx, const double *sy, const double *sz,
                                            const double *sl, int riA, int sX, int sY, Acc &A, Acc &B,
                                            double pA[4], double pB[4])
{
    acc_zero(A);
    acc_zero(B);
    if constexpr (TWO_D) {
        centre_column_2d<RC, MIXED>(sx, sy, sz, sl, riA, A, pA);
        static_for<1, RC>([&](auto dxc) __attribute__((always_inline)) {
            column_fence(A, B);
            column_pair_2d<RC, decltype(dxc)::value, MIXED>(sx, sy, sz, sl, riA, sX, A);
        });
    } else {
        centre_column_3d<RC, MIXED>(sx, sy, sz, sl, riA, A, B, pA, pB);
        // columns with DX > 0, or DX == 0 and DY > 0, inside the cutoff disc
        static_for<0, RC>([&](auto dxc) __attribute__((always_inline)) {
            constexpr int DX = decltype(dxc)::value;
            static_for<-RC, RC>([&](auto dyc) __attribute__((always_inline)) {
                constexpr int DY = decltype(dyc)::value;
                if constexpr ((DX > 0 || DY > 0) && DX * DX + DY * DY <= RC * RC) {
                    column_fence(A, B);
                    column_pair_3d<RC, DX, DY, MIXED>(sx, sy, sz, sl, riA, sX, sY, A, B);
                }
            });
        });
    }
    column_fence(A, B);
}

template <int RC, bool TWO_D>
constexpr int n_pairs()
{
    int n = 0;
    for (int dx = -RC; dx <= RC; dx++)
        for (int dy = -RC; dy <= RC; dy++)
            for (int dz = (TWO_D ? 0 : -RC); dz <= (TWO_D ? 0 : RC); dz++) {
                int r2 = dx * dx + dy * dy + dz * dz;
                if (r2 > 0 && r2 <= RC * RC) n++;
            }
    return n / 2;
}

// largest tile (sites) a workgroup takes: each lane gathers at most two
// items of two sites (3D) or one site (2D)
constexpr int kFieldMaxTile3D = 4 * kBlock, kFieldMaxTile2D = 2 * kBlock;

struct FieldArgs {
    SweepArgs s;
    const int *steps;         // per shape (step_off): per step, barrier flag + NGROUPS tile indices
    const int *tile_coord;    // per shape (coord_off), by padded tile index: lx | ly << 10 | lz << 20, -1 in the padding
    int tile_doubles;         // stride of the tile-order arrays (largest padded tile)
    unsigned long long *tile_counts;  // 2 per tile (accepted, rejected), accumulated
    int phaseA_doubles;       // LDS doubles reserved for the staged region / compact arrays
    int region_doubles;       // stride between the staged x, y, z (, length) arrays
};

template <int RC, bool TWO_D, bool MIXED, int G>
__global__ __launch_bounds__(kBlock) void field_sweep_kernel(FieldArgs fa, int first_tile, int ntiles)
{
    constexpr int HZ = TWO_D ? 0 : RC;
    constexpr int S = TWO_D ? 1 : 2;                 // sites per gather item (along z)
    constexpr int NP = n_pairs<RC, TWO_D>();
    constexpr int NGROUPS = kBlock / G;
    const SweepArgs &a = fa.s;
    extern __shared__ double smem[];
    const int tid = threadIdx.x;
    const int b = blockIdx.x;
    const int t = (ntiles % 8 == 0) ? (b % 8) * (ntiles / 8) + b / 8 : b;
    const int tile_id = first_tile + t;
    const TileDev tile = a.tiles[tile_id];
    const RepDev rep = a.reps[tile.replica];
    const ShapeDev sh = a.shapes[tile.shape];
    const int RY = sh.ty + 2 * RC, RZ = sh.tz + 2 * HZ, RX = sh.tx + 2 * RC;
    const int region = RX * RY * RZ;
    // z rows are padded to an odd length in 3D: with the x-fastest item order
    // below, a wave's column loads then spread evenly over the LDS banks
    const int RZp = TWO_D ? RZ : (RZ | 1);
    const int sY = RZp, sX = RY * RZp;
    // tile-order arrays use the padded index (lx * TYP + ly) * TZP + lz: the
    // padded strides spread a scatter's targets over the LDS banks
    const int TYP = sh.typ, TZP = sh.tzp;
    const int npad = sh.tx * TYP * TZP;
    const int MT = fa.tile_doubles;

    // ---- LDS: [phase A / phase B area][tables].  Region arrays are padded by
    // two doubles: the phantom second site of the last item of an odd-length
    // tile column reads one element past its column.
    const int MR = fa.region_doubles;
    double *sx = smem;
    double *sy = sx + MR;
    double *sz = sy + MR;
    double *sl = sz + MR;                               // MIXED only
    // phase B (after the fields are computed) reuses the same area, tile order
    double *cfx = smem, *cfy = cfx + MT, *cfz = cfy + MT;
    double *cpx = cfz + MT, *cpy = cpx + MT, *cpz = cpy + MT;
    double *cqx = cpz + MT, *cqy = cqx + MT, *cqz = cqy + MT, *cthr = cqz + MT;
    double *cl = cthr + MT;                             // MIXED only
    int *s_coord = (int *)(smem + fa.phaseA_doubles);   // tile order
    int *s_step = s_coord + MT;                           // step table, (NGROUPS + 1) ints per step

    for (int i = tid; i < npad; i += kBlock) s_coord[i] = fa.tile_coord[sh.coord_off + i];
    for (int i = tid; i < sh.nsteps * (NGROUPS + 1); i += kBlock) s_step[i] = fa.steps[sh.step_off + i];

    // ---- stage tile + halo; the proposals of this lane's tile sites are drawn
    // while the staging loads are in flight and kept in registers until the
    // tile-order arrays exist ----
    const unsigned sweep = a.sweep_ctr[tile.replica];
    const long long tile_base = ((long long)tile.x0 * rep.Y + tile.y0) * rep.Z + tile.z0;
    const int ntile = sh.tx * sh.ty * sh.tz;
    constexpr int NPROP = (TWO_D ? kFieldMaxTile2D : kFieldMaxTile3D) / kBlock;
    struct Prop {
        double x, y, z, thr;
        int kp;   // padded tile index, -1: no site
    };
    Prop pr[NPROP];
    auto draw = [&](int i, Prop &P) __attribute__((always_inline)) {
        const int c = tid + kBlock * i;   // dense tile index (lx * ty + ly) * tz + lz
        P.kp = -1;
        P.x = P.y = P.z = P.thr = 0.0;
        if (c < ntile) {
            const int lz = c % sh.tz, r_ = c / sh.tz, ly = r_ % sh.ty, lx = r_ / sh.ty;
            const unsigned gsite = (unsigned)(tile_base + ((long long)lx * rep.Y + ly) * rep.Z + lz);
            double u;
            counter_draw(rep.seed, (uint32_t)tile.replica, sweep, gsite, rep.constrain_to_x, P.x, P.y, P.z, u);
            P.thr = metropolis_threshold(u, rep.beta);
            P.kp = (lx * TYP + ly) * TZP + lz;
        }
    };
    const double4 *lat = a.lat + rep.base;
    {
        int iz = tid % RZ, q = tid / RZ, iy = q % RY, ix = q / RY;
        const int dz_ = kBlock % RZ, q2 = kBlock / RZ, dy_ = q2 % RY, dx_ = q2 / RY;
        // one batch: 4 loads per lane issued, `between` runs, then the LDS writes
        auto batch = [&](int base, auto &&between) __attribute__((always_inline)) {
            double4 v[4];
            int idx[4], pidx[4];
#pragma unroll
            for (int k = 0; k < 4; k++) {
                idx[k] = base + k * kBlock;
                pidx[k] = (ix * RY + iy) * RZp + iz;   // padded LDS index
                if (idx[k] < region) {
                    int gx = wrap_once(tile.x0 - RC + ix, rep.X);
                    int gy = wrap_once(tile.y0 - RC + iy, rep.Y);
                    int gz = wrap_once(tile.z0 - HZ + iz, rep.Z);
                    v[k] = lat[((long long)gx * rep.Y + gy) * rep.Z + gz];
                }
                iz += dz_; iy += dy_; ix += dx_;
                if (iz >= RZ) { iz -= RZ; iy++; }
                if (iy >= RY) { iy -= RY; ix++; }
            }
            between();
#pragma unroll
            for (int k = 0; k < 4; k++)
                if (idx[k] < region) {
                    const int d = pidx[k];
                    sx[d] = v[k].x; sy[d] = v[k].y; sz[d] = v[k].z;
                    if (MIXED) sl[d] = v[k].w;
                }
        };
        static_for<0, NPROP - 1>([&](auto ic) __attribute__((always_inline)) {
            constexpr int i = decltype(ic)::value;
            batch(tid + i * 4 * kBlock, [&]() __attribute__((always_inline)) { draw(i, pr[i]); });
        });
        for (int base = tid + NPROP * 4 * kBlock; base < region; base += 4 * kBlock)
            batch(base, []() {});
    }
    __syncthreads();

    // ---- local fields of every tile site (items of S sites along z) ----
    const int tzp = (sh.tz + S - 1) / S;
    const int nitems = sh.tx * sh.ty * tzp;
    // Each lane handles at most two items (tiles up to 2 * kBlock * S sites),
    // written out as two explicit calls: no runtime-indexed result arrays.
    struct ItemOut {
        double f0[3], p0[4], f1[3], p1[4];
        int c0, c1;
    };
    const double J = rep.J;
    const double c_nm = rep.L0 * rep.L0;
    auto gather_item = [&](int item, ItemOut &o) __attribute__((always_inline)) {
        o.c0 = -1;
        o.c1 = -1;
        if (item >= nitems) return;
        int lx = item % sh.tx, q = item / sh.tx, ly = q % sh.ty, zp = q / sh.ty;
        int lz = zp * S;
        int riA = ((lx + RC) * RY + (ly + RC)) * RZp + (lz + HZ);
        Acc A, B;
        double pA[4], pB[4];
        tile_fields<RC, TWO_D, MIXED>(sx, sy, sz, sl, riA, sX, sY, A, B, pA, pB);
        const double cA = MIXED ? pA[3] : c_nm;
        o.f0[0] = cA * A.hx - J * A.gx + rep.E[0];
        o.f0[1] = cA * A.hy - J * A.gy + rep.E[1];
        o.f0[2] = cA * A.hz - J * A.gz + rep.E[2];
        o.p0[0] = pA[0]; o.p0[1] = pA[1]; o.p0[2] = pA[2]; o.p0[3] = pA[3];
        o.c0 = (lx * TYP + ly) * TZP + lz;
        if (S == 2 && lz + 1 < sh.tz) {
            const double cB = MIXED ? pB[3] : c_nm;
            o.f1[0] = cB * B.hx - J * B.gx + rep.E[0];
            o.f1[1] = cB * B.hy - J * B.gy + rep.E[1];
            o.f1[2] = cB * B.hz - J * B.gz + rep.E[2];
            o.p1[0] = pB[0]; o.p1[1] = pB[1]; o.p1[2] = pB[2]; o.p1[3] = pB[3];
            o.c1 = (lx * TYP + ly) * TZP + lz + 1;
        }
    };
    ItemOut o0, o1;
    gather_item(tid, o0);
    asm volatile("" ::: "memory");
    gather_item(tid + kBlock, o1);
    __syncthreads();   // the staged region is dead from here on
    auto put = [&](const ItemOut &o) __attribute__((always_inline)) {
        if (o.c0 >= 0) {
            cfx[o.c0] = o.f0[0]; cfy[o.c0] = o.f0[1]; cfz[o.c0] = o.f0[2];
            cpx[o.c0] = o.p0[0]; cpy[o.c0] = o.p0[1]; cpz[o.c0] = o.p0[2];
            if (MIXED) cl[o.c0] = o.p0[3];
        }
        if (o.c1 >= 0) {
            cfx[o.c1] = o.f1[0]; cfy[o.c1] = o.f1[1]; cfz[o.c1] = o.f1[2];
            cpx[o.c1] = o.p1[0]; cpy[o.c1] = o.p1[1]; cpz[o.c1] = o.p1[2];
            if (MIXED) cl[o.c1] = o.p1[3];
        }
    };
    put(o0);
    put(o1);
    // ---- proposals into the tile-order arrays ----
#pragma unroll
    for (int i = 0; i < NPROP; i++)
        if (pr[i].kp >= 0) {
            const int c = pr[i].kp;
            cqx[c] = pr[i].x; cqy[c] = pr[i].y; cqz[c] = pr[i].z;
            cthr[c] = pr[i].thr;
        }

    // ---- this lane's scatter pairs: pair j = lane + 64 q of the +-r list ----
    // An accepted site i adds to the field of every tile site t = i +- r
    //   c_t T_r (L_i dp) - J dp [|r| = 1],  T_r m = w m - 3 w (n.m) n,
    // written per pair as A dp + (B (n.dp)) n: A = c w - J [nn], B = -3 c w
    // with c = L0^2 for uniform lengths; with mixed lengths A = w, B = -3 w and
    // the result is scaled by c_t L_i and -J dp [nn] added per target.
    constexpr int PPL = (NP + 63) / 64;
    const int lane = tid & 63;
    const int part = tid % G, group = tid / G;
    int pdx[PPL], pdy[PPL], pdz[PPL], poff[PPL];
    double pnx[PPL], pny[PPL], pnz[PPL], pA[PPL], pB[PPL], pjn[PPL];
#pragma unroll
    for (int q = 0; q < PPL; q++) {
        const int j = lane + 64 * q;
        if (j < NP) {
            PairDev p = a.pairs[j];
            pdx[q] = p.dx; pdy[q] = p.dy; pdz[q] = p.dz;
            pnx[q] = p.nx; pny[q] = p.ny; pnz[q] = p.nz;
            pjn[q] = p.nn ? -J : 0.0;
            if constexpr (MIXED) { pA[q] = p.w; pB[q] = -3.0 * p.w; }
            else { pA[q] = c_nm * p.w + pjn[q]; pB[q] = -3.0 * c_nm * p.w; }
        } else {
            pdx[q] = 1 << 20; pdy[q] = 0; pdz[q] = 0;   // never inside a tile
            pnx[q] = pny[q] = pnz[q] = pA[q] = pB[q] = pjn[q] = 0.0;
        }
        poff[q] = (pdx[q] * TYP + pdy[q]) * TZP + pdz[q];
    }
    const double K = rep.K;
    unsigned nacc = 0, nrej = 0;
    __syncthreads();

    // ---- fine colours ----
    // A step is up to NGROUPS sites of one fine colour, one per group of G
    // lanes; the host lists them per tile shape: per step, a barrier flag (set
    // on a colour's last step) and the tile index of each group's site (-1:
    // none).  Each group decides its site; then every wave walks its accepted
    // sites one at a time, all 64 lanes scattering that site's pairs, so a
    // rejected site costs no scatter instructions.
    // The site data of a step (dipole, proposal, threshold, coordinates) do
    // not change while earlier steps run: a site's dipole is written only in
    // its own step, and sites of one colour never interact.  So step s + 1's
    // data are loaded with step s's fields, and step s + 2's site indices with
    // them; after each colour's barrier the field read, the decision and the
    // atomics remain on the critical path.  Targets in colours already done
    // are still updated (their fields are never read again).
    constexpr int SW = NGROUPS + 1;                   // ints per step in the table
    struct Raw {
        int k, pc;
        double ox, oy, oz, qx, qy, qz, thr, Li;
    };
    struct Prep {
        int k, pc;                                    // tile index (>= 0), packed local coordinates
        int live;                                     // 1: a non-vacant site to decide
        double dx, dy, dz, qx, qy, qz, thr, kt, Li;
    };
    auto load_raw = [&](int kk, Raw &R) __attribute__((always_inline)) {
        const int k = kk < 0 ? 0 : kk;
        R.k = kk;
        R.ox = cpx[k]; R.oy = cpy[k]; R.oz = cpz[k];
        R.qx = cqx[k]; R.qy = cqy[k]; R.qz = cqz[k];
        R.thr = cthr[k];
        R.Li = MIXED ? cl[k] : rep.L0;
        R.pc = s_coord[k];
    };
    auto prep = [&](const Raw &R, Prep &P) __attribute__((always_inline)) {
        P.k = R.k < 0 ? 0 : R.k;
        P.pc = R.pc;
        P.live = (R.k >= 0 && R.Li != 0.0) ? 1 : 0;
        P.qx = R.qx; P.qy = R.qy; P.qz = R.qz;
        P.thr = R.thr;
        P.Li = R.Li;
        P.dx = R.qx - R.ox; P.dy = R.qy - R.oy; P.dz = R.qz - R.oz;
        P.kt = (K > 0.0) ? -K * (fabs(R.qx) - fabs(R.ox) + fabs(R.qy) - fabs(R.oy)) : 0.0;
    };
    // the prepared values are final before a barrier (the compiler would
    // otherwise sink their arithmetic past it, onto the next step's critical
    // path)
    auto settle = [&](Prep &P) __attribute__((always_inline)) {
        asm volatile("" : "+v"(P.k), "+v"(P.pc), "+v"(P.live), "+v"(P.dx), "+v"(P.dy), "+v"(P.dz), "+v"(P.qx),
                     "+v"(P.qy), "+v"(P.qz), "+v"(P.thr), "+v"(P.kt), "+v"(P.Li));
    };
    auto readlane_d = [](double v, int l) __attribute__((always_inline)) {
        const long long b = __double_as_longlong(v);
        const int lo = __builtin_amdgcn_readlane((int)b, l);
        const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
        return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
    };

    const int nsteps = sh.nsteps;
    if (nsteps > 0) {
        Prep cur, nxt;
        Raw raw;
        load_raw(s_step[1 + group], raw);
        int k_next = nsteps > 1 ? s_step[SW + 1 + group] : -1;   // site of step s + 1
        int flag = s_step[0];
        prep(raw, cur);
        for (int st = 0; st < nsteps; st++) {
            // loads: this step's fields, step s + 1's site data, step s + 2's site
            const int k = cur.k;
            const double fx = cfx[k], fy = cfy[k], fz = cfz[k];
            const bool more = st + 1 < nsteps;
            int flag_next = 0, k_next2 = -1;
            if (more) {
                load_raw(k_next, raw);
                flag_next = s_step[(st + 1) * SW];
                if (st + 2 < nsteps) k_next2 = s_step[(st + 2) * SW + 1 + group];
            }
            // decide from the current field (every lane of the group alike)
            const double dE = fma(cur.dx, fx, fma(cur.dy, fy, cur.dz * fz)) + cur.kt;
            const bool acc = cur.live && dE < cur.thr;
            if (cur.live && part == 0) {
                if (acc) { cpx[k] = cur.qx; cpy[k] = cur.qy; cpz[k] = cur.qz; nacc++; }
                else nrej++;
            }
            // the wave's accepted sites, one at a time (lane G g leads group g)
            unsigned long long bal = __ballot(acc);
            while (bal) {
                const int src = __ffsll((long long)bal) - 1;
                bal &= ~(((G >= 64) ? ~0ull : ((1ull << G) - 1)) << src);
                const double sdx = readlane_d(cur.dx, src), sdy = readlane_d(cur.dy, src),
                             sdz = readlane_d(cur.dz, src);
                const int sk = __builtin_amdgcn_readlane(cur.k, src);
                const int spc = __builtin_amdgcn_readlane(cur.pc, src);
                const double sL = MIXED ? readlane_d(cur.Li, src) : 1.0;
                const int lx = spc & 1023, ly = (spc >> 10) & 1023, lz = spc >> 20;
#pragma unroll
                for (int q = 0; q < PPL; q++) {
                    const double sp = fma(pnx[q], sdx, fma(pny[q], sdy, pnz[q] * sdz));
                    const double t = pB[q] * sp;
                    const double ux = fma(pA[q], sdx, t * pnx[q]);
                    const double uy = fma(pA[q], sdy, t * pny[q]);
                    const double uz = fma(pA[q], sdz, t * pnz[q]);
#pragma unroll
                    for (int h = 0; h < 2; h++) {
                        const int sg = h ? 1 : -1;
                        const int tx_ = lx + sg * pdx[q], ty_ = ly + sg * pdy[q], tz_ = lz + sg * pdz[q];
                        if ((unsigned)tx_ < (unsigned)sh.tx && (unsigned)ty_ < (unsigned)sh.ty &&
                            (unsigned)tz_ < (unsigned)sh.tz) {
                            const int tc = sk + sg * poff[q];
                            if constexpr (MIXED) {
                                const double ct = cl[tc] * sL;
                                atomicAdd(&cfx[tc], fma(ct, ux, pjn[q] * sdx));
                                atomicAdd(&cfy[tc], fma(ct, uy, pjn[q] * sdy));
                                atomicAdd(&cfz[tc], fma(ct, uz, pjn[q] * sdz));
                            } else {
                                atomicAdd(&cfx[tc], ux);
                                atomicAdd(&cfy[tc], uy);
                                atomicAdd(&cfz[tc], uz);
                            }
                        }
                    }
                }
            }
            if (!more) break;
            prep(raw, nxt);
            settle(nxt);
            if (flag) __syncthreads();
            cur = nxt;
            flag = flag_next;
            k_next = k_next2;
        }
        __syncthreads();
    }

    // ---- write the tile back ----
    double4 *out = a.lat + rep.base;
    for (int c = tid; c < npad; c += kBlock) {
        const int pc = s_coord[c];
        if (pc < 0) continue;
        const int lx = pc & 1023, ly = (pc >> 10) & 1023, lz = pc >> 20;
        double4 v;
        v.x = cpx[c]; v.y = cpy[c]; v.z = cpz[c];
        v.w = MIXED ? cl[c] : rep.L0;
        out[tile_base + ((long long)lx * rep.Y + ly) * rep.Z + lz] = v;
    }
    // ---- counters: this tile's own slot (one workgroup per tile per launch) ----
    unsigned long long ca = nacc, cr = nrej;
#pragma unroll
    for (int o = 32; o >= 1; o >>= 1) {
        ca += __shfl_xor(ca, o, 64);
        cr += __shfl_xor(cr, o, 64);
    }
    __shared__ unsigned long long s_cnt[2 * (kBlock / 64)];
    if ((tid & 63) == 0) { s_cnt[2 * (tid >> 6)] = ca; s_cnt[2 * (tid >> 6) + 1] = cr; }
    __syncthreads();
    if (tid == 0) {
        unsigned long long A_ = 0, R_ = 0;
        for (int w_ = 0; w_ < kBlock / 64; w_++) { A_ += s_cnt[2 * w_]; R_ += s_cnt[2 * w_ + 1]; }
        fa.tile_counts[2 * tile_id] += A_;
        fa.tile_counts[2 * tile_id + 1] += R_;
    }
}

}  // namespace sn
