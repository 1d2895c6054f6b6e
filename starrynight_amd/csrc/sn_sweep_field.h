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

struct FieldArgs {
    SweepArgs s;
    const int *class_tile;    // per shape (sites_off): tile-local index of each site, by fine colour
    const int *tile_coord;    // per shape (sites_off): packed local coords lx | ly << 10 | lz << 20
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
    constexpr int NPL = (NP + G - 1) / G;            // scatter pairs per lane
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
    const int ntile = sh.tx * sh.ty * sh.tz;
    const int MT = a.max_tile_sites;

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
    int *s_ctile = s_coord + MT;                          // class order -> tile index
    int *s_cls = s_ctile + MT;

    for (int i = tid; i < ntile; i += kBlock) {
        s_coord[i] = fa.tile_coord[sh.sites_off + i];
        s_ctile[i] = fa.class_tile[sh.sites_off + i];
    }
    for (int i = tid; i <= sh.ncol; i += kBlock) s_cls[i] = a.class_start[sh.class_off + i];

    // ---- stage tile + halo ----
    const double4 *lat = a.lat + rep.base;
    {
        int iz = tid % RZ, q = tid / RZ, iy = q % RY, ix = q / RY;
        const int dz_ = kBlock % RZ, q2 = kBlock / RZ, dy_ = q2 % RY, dx_ = q2 / RY;
        for (int base = tid; base < region; base += 4 * kBlock) {
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
        o.c0 = (lx * sh.ty + ly) * sh.tz + lz;
        if (S == 2 && lz + 1 < sh.tz) {
            const double cB = MIXED ? pB[3] : c_nm;
            o.f1[0] = cB * B.hx - J * B.gx + rep.E[0];
            o.f1[1] = cB * B.hy - J * B.gy + rep.E[1];
            o.f1[2] = cB * B.hz - J * B.gz + rep.E[2];
            o.p1[0] = pB[0]; o.p1[1] = pB[1]; o.p1[2] = pB[2]; o.p1[3] = pB[3];
            o.c1 = (lx * sh.ty + ly) * sh.tz + lz + 1;
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
    // ---- proposals of every tile site (tile order) ----
    const unsigned sweep = a.sweep_ctr[tile.replica];
    const long long tile_base = ((long long)tile.x0 * rep.Y + tile.y0) * rep.Z + tile.z0;
    for (int c = tid; c < ntile; c += kBlock) {
        const int pc = s_coord[c];
        const int lx = pc & 1023, ly = (pc >> 10) & 1023, lz = pc >> 20;
        unsigned gsite = (unsigned)(tile_base + ((long long)lx * rep.Y + ly) * rep.Z + lz);
        double px, py, pz, u;
        counter_draw(rep.seed, (uint32_t)tile.replica, sweep, gsite, rep.constrain_to_x, px, py, pz, u);
        cqx[c] = px; cqy[c] = py; cqz[c] = pz;
        cthr[c] = metropolis_threshold(u, rep.beta);
    }

    // ---- this lane's scatter pairs ----
    const int part = tid % G, group = tid / G;
    int pdx[NPL], pdy[NPL], pdz[NPL];
    double pnx[NPL], pny[NPL], pnz[NPL], pw[NPL], pnn[NPL];
#pragma unroll
    for (int m = 0; m < NPL; m++) {
        const int j = part + m * G;
        if (j < NP) {
            PairDev p = a.pairs[j];
            pdx[m] = p.dx; pdy[m] = p.dy; pdz[m] = p.dz;
            pnx[m] = p.nx; pny[m] = p.ny; pnz[m] = p.nz; pw[m] = p.w; pnn[m] = p.nn ? 1.0 : 0.0;
        } else {
            pdx[m] = 1 << 20; pdy[m] = 0; pdz[m] = 0;   // never inside a tile
            pnx[m] = pny[m] = pnz[m] = pw[m] = pnn[m] = 0.0;
        }
    }
    const double K = rep.K;
    unsigned nacc = 0, nrej = 0;
    __syncthreads();

    // ---- fine colours ----
    for (int c = 0; c < sh.ncol; c++) {
        const int cs = s_cls[c], ce = s_cls[c + 1];
        for (int sb = cs; sb < ce; sb += NGROUPS) {
            const int si = sb + group;
            const bool active = si < ce;
            const int k = s_ctile[active ? si : cs];   // tile index of the site
            const double fx = cfx[k], fy = cfy[k], fz = cfz[k];
            const double ox = cpx[k], oy = cpy[k], oz = cpz[k];
            const double qx = cqx[k], qy = cqy[k], qz = cqz[k], thr = cthr[k];
            const double Li = MIXED ? cl[k] : rep.L0;
            const double dx = qx - ox, dy = qy - oy, dz = qz - oz;
            double dE = dx * fx + dy * fy + dz * fz;
            if (K > 0.0) dE += -K * (fabs(qx) - fabs(ox) + fabs(qy) - fabs(oy));
            const bool acc = active && Li != 0.0 && dE < thr;
            if (active && part == 0 && Li != 0.0) {
                if (acc) { cpx[k] = qx; cpy[k] = qy; cpz[k] = qz; nacc++; }
                else nrej++;
            }
            if (acc) {
                // f_t += c_t T_r (L_i dp) - J dp [|r| = 1] for tile sites t = i +- r
                const int pc = s_coord[k];
                const int lx = pc & 1023, ly = (pc >> 10) & 1023, lz = pc >> 20;
                const double sc = MIXED ? Li : 1.0;
                const double mx = sc * dx, my = sc * dy, mz = sc * dz;
#pragma unroll
                for (int m = 0; m < NPL; m++) {
                    const double s = pnx[m] * mx + pny[m] * my + pnz[m] * mz;
                    const double tt = -3.0 * pw[m] * s;
                    const double vx = pw[m] * mx + tt * pnx[m];
                    const double vy = pw[m] * my + tt * pny[m];
                    const double vz = pw[m] * mz + tt * pnz[m];
                    const double jn = -J * pnn[m];
#pragma unroll
                    for (int sg = -1; sg <= 1; sg += 2) {
                        const int tx_ = lx + sg * pdx[m], ty_ = ly + sg * pdy[m], tz_ = lz + sg * pdz[m];
                        if ((unsigned)tx_ < (unsigned)sh.tx && (unsigned)ty_ < (unsigned)sh.ty &&
                            (unsigned)tz_ < (unsigned)sh.tz) {
                            const int tc = (tx_ * sh.ty + ty_) * sh.tz + tz_;
                            const double ct = MIXED ? cl[tc] : c_nm;
                            atomicAdd(&cfx[tc], fma(ct, vx, jn * dx));
                            atomicAdd(&cfy[tc], fma(ct, vy, jn * dy));
                            atomicAdd(&cfz[tc], fma(ct, vz, jn * dz));
                        }
                    }
                }
            }
        }
        __syncthreads();
    }

    // ---- write the tile back ----
    double4 *out = a.lat + rep.base;
    for (int c = tid; c < ntile; c += kBlock) {
        const int pc = s_coord[c];
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
