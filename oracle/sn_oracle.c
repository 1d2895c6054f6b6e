/*
 * sn_oracle.c -- CPU oracle (test infrastructure only; see sn_oracle.h).
 *
 * Compile with -ffp-contract=off and without -ffast-math / -march: the float
 * restatement must round every float operation exactly as the reference's
 * x86-64 SSE build does (src/starrynight-config.c:189-199 has no FMA).
 *
 * Parity pins (tests/test_oracle_*.py):
 *   or_mt_*            numpy's MT19937 (init_genrand seeding), the C++ standard's
 *                      10000th output of std::mt19937(5489) = 4123659995, and the
 *                      reference's own src/mt19937ar-cok.c built into oracle/_ref.
 *   or_philox4x32_10   Random123 known-answer vectors and rocRAND's
 *                      philox4x32_10 engine (tests/golden/philox_rocrand.json).
 *   or_gen_neighbour   closed forms: count of integer points in the ball
 *                      (28 / 122 / 256), dx,dy,dz ascending, dz == 0 iff Z == 1.
 *   or_site_energy_*   ferroelectric flip closed form (12J - 2Ex in 3D), the
 *                      textbook two-dipole energy, dE(p->p) = 0, antisymmetry,
 *                      translation / cubic-rotation invariance, brute-force
 *                      E_tot(new) - E_tot(old); f32 vs f64 agreement ~1e-6.
 *   or_energy/gradient central finite differences; dE = (p'-p).dE/dp_i (K = 0).
 *   or_counter_draw    |p| = 1, E[p] = 0, E[p p^T] = I/3, u uniform.
 *   or_radial_order    count column = N * r3(r^2) (the RDF .dat files in data/ hold it).
 */
#include "sn_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------ */
/* MT19937: Matsumoto & Nishimura (1998).  Restates init_genrand / next_state /
 * genrand_int32 / genrand_real1 / genrand_real2 (src/mt19937ar-cok.c:63-197).
 * The Cokus refill-all-at-once variant produces the canonical stream. */
#define MT_N 624
#define MT_M 397

void or_mt_seed(or_mt19937 *g, uint32_t s)
{
    g->mt[0] = s;
    for (int j = 1; j < MT_N; j++)
        g->mt[j] = 1812433253u * (g->mt[j - 1] ^ (g->mt[j - 1] >> 30)) + (uint32_t)j;
    g->idx = MT_N;
}

static void mt_refill(or_mt19937 *g)
{
    uint32_t *s = g->mt;
    for (int k = 0; k < MT_N; k++) {
        uint32_t y = (s[k] & 0x80000000u) | (s[(k + 1) % MT_N] & 0x7fffffffu);
        uint32_t v = s[(k + MT_M) % MT_N] ^ (y >> 1);
        if (s[(k + 1) % MT_N] & 1u) v ^= 0x9908b0dfu;
        s[k] = v;
    }
    g->idx = 0;
}

uint32_t or_mt_u32(or_mt19937 *g)
{
    if (g->idx >= MT_N) mt_refill(g);
    uint32_t y = g->mt[g->idx++];
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

double or_mt_real1(or_mt19937 *g) { return (double)or_mt_u32(g) * (1.0 / 4294967295.0); }
double or_mt_real2(or_mt19937 *g) { return (double)or_mt_u32(g) * (1.0 / 4294967296.0); }

or_mt19937 *or_mt_new(uint32_t s)
{
    or_mt19937 *g = (or_mt19937 *)malloc(sizeof(*g));
    if (g) or_mt_seed(g, s);
    return g;
}
void or_mt_free(or_mt19937 *g) { free(g); }
void or_mt_fill_u32(or_mt19937 *g, int64_t n, uint32_t *out)
{
    for (int64_t i = 0; i < n; i++) out[i] = or_mt_u32(g);
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11): 10 rounds of the 4x32 S-box, Weyl key schedule. */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* ------------------------------------------------------------------------ */
/* gen_neighbour (src/starrynight-montecarlo-core.c:38-72): every integer
 * offset r != 0 with float |r| <= rc, dx, dy, dz ascending, dz fixed to 0 when
 * Z == 1.  d is the float sqrt of the float sum, as in the reference (:54). */
int or_gen_neighbour(int rc, int Z, or_neighbour *out, int cap)
{
    int zc = (Z == 1) ? 0 : rc;
    int n = 0;
    for (int dx = -rc; dx <= rc; dx++)
        for (int dy = -rc; dy <= rc; dy++)
            for (int dz = -zc; dz <= zc; dz++) {
                if (dx == 0 && dy == 0 && dz == 0) continue;
                float s = (float)dx * dx + dy * dy + dz * dz;
                float d = (float)sqrt((double)s);
                if (d > (float)rc) continue;
                if (n >= cap) return -1;
                out[n].dx = dx; out[n].dy = dy; out[n].dz = dz; out[n].d = d;
                n++;
            }
    return n;
}

static inline int64_t site_index(const or_model *m, int x, int y, int z)
{
    return ((int64_t)x * m->Y + y) * m->Z + z;
}
static inline int wrap(int v, int L) { return (L + v) % L; }

/* ------------------------------------------------------------------------ */
/* Float-faithful restatement. */

/* dot (src/starrynight-config.c:189-199): float accumulate, x then y then z. */
static inline float dotf(const float *a, const float *b)
{
    float sum = 0.0f;
    sum += a[0] * b[0];
    sum += a[1] * b[1];
    sum += a[2] * b[2];
    return sum;
}

/* site_energy (src/starrynight-montecarlo-core.c:76-141). */
double or_site_energy_f32(const or_model *m, const or_neighbour *nb, int nnb, const float *lat,
                          int x, int y, int z, const float newp[4], const float oldp[4])
{
    double dE = 0.0;
    for (int i = 0; i < nnb; i++) {
        int dx = nb[i].dx, dy = nb[i].dy, dz = nb[i].dz;
        float d = nb[i].d;
        const float *t = lat + 4 * site_index(m, wrap(x + dx, m->X), wrap(y + dy, m->Y),
                                              wrap(z + dz, m->Z));
        float n[3] = {(float)dx / d, (float)dy / d, (float)dz / d};
        float ll = oldp[3] * t[3];
        float a = dotf(newp, t) - 3 * dotf(n, newp) * dotf(n, t);
        float b = dotf(oldp, t) - 3 * dotf(n, oldp) * dotf(n, t);
        float term = ll * (a - b) / (d * d * d);
        dE += term;
        if ((dx * dx + dy * dy + dz * dz) == 1)
            dE += -m->J * dotf(newp, t) + m->J * dotf(oldp, t);
    }
    float Ef[3] = {(float)m->E[0], (float)m->E[1], (float)m->E[2]};
    dE += +dotf(newp, Ef) - dotf(oldp, Ef);
    if (m->K > 0.0) {
        float ex[3] = {1.0f, 0.0f, 0.0f}, ey[3] = {0.0f, 1.0f, 0.0f};
        dE += -m->K * fabs(dotf(newp, ex)) + m->K * fabs(dotf(oldp, ex));
        dE += -m->K * fabs(dotf(newp, ey)) + m->K * fabs(dotf(oldp, ey));
    }
    return dE;
}

/* random_sphere_point (src/starrynight-config.c:203-227), DIM == 3 branch. */
void or_random_sphere_point_f32(or_mt19937 *g, float p[3])
{
    float x1, x2;
    do {
        x1 = 2.0 * or_mt_real1(g) - 1.0;
        x2 = 2.0 * or_mt_real1(g) - 1.0;
    } while (x1 * x1 + x2 * x2 > 1.0);
    p[0] = 2 * x1 * sqrt(1 - x1 * x1 - x2 * x2);
    p[1] = 2 * x2 * sqrt(1 - x1 * x1 - x2 * x2);
    p[2] = 1.0 - 2.0 * (x1 * x1 + x2 * x2);
}

static inline int rand_int(or_mt19937 *g, int span)
{
    /* rand_int (src/starrynight-montecarlo-core.c:18-21) */
    return (int)((unsigned long)or_mt_u32(g) % (unsigned long)span);
}

/* random_X_point (src/starrynight-config.c:230-263) */
void or_random_X_point_f32(or_mt19937 *g, float p[3])
{
    static const int ax[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    int i = rand_int(g, 6);
    p[0] = (float)ax[i][0]; p[1] = (float)ax[i][1]; p[2] = (float)ax[i][2];
}

/* initialise_lattice_random (src/starrynight-lattice.c:25-37) */
void or_init_lattice_random_f32(or_mt19937 *g, int X, int Y, int Z, float *lat)
{
    int64_t N = (int64_t)X * Y * Z;
    for (int64_t i = 0; i < N; i++) or_random_sphere_point_f32(g, lat + 4 * i);
}

/* The species walk of solid_solution (src/starrynight-lattice.c:154-159):
 * the float sample steps down the list while it exceeds the entry's
 * prevalence, subtracting it in float.  Returns the species index, or -1
 * where the reference would read past its n-entry table. */
int or_species_walk_f32(float sample, int n, const float *prev)
{
    int i;
    for (i = 0; sample > prev[i]; sample -= prev[i], i++)
        if (i + 1 >= n) return -1;
    return i;
}

/* solid_solution (src/starrynight-lattice.c:139-171): one genrand_real1 per
 * site in x, y, z order, narrowed to float (:143, :152).  Returns -1 where the
 * reference would walk past the species list (prevalence summing below 1). */
int or_solid_solution_f32(or_mt19937 *g, int X, int Y, int Z, int n, const float *len,
                          const float *prev, float *lat)
{
    int64_t N = (int64_t)X * Y * Z;
    for (int64_t s = 0; s < N; s++) {
        float sample = or_mt_real1(g);
        int i = or_species_walk_f32(sample, n, prev);
        if (i < 0) return -1;
        lat[4 * s + 3] = len[i];
    }
    return 0;
}

/* MC_moves / MC_move (src/starrynight-montecarlo-core.c:143-191). */
int64_t or_mc_moves_f32(const or_model *m, const or_neighbour *nb, int nnb, float *lat,
                        or_mt19937 *g, int64_t moves, uint64_t *acc, uint64_t *rej,
                        or_move_rec *trace)
{
    for (int64_t k = 0; k < moves; k++) {
        int x = rand_int(g, m->X);
        int y = rand_int(g, m->Y);
        int z = rand_int(g, m->Z);
        int64_t s = site_index(m, x, y, z);
        float *old = lat + 4 * s;
        or_move_rec *r = trace ? trace + k : NULL;
        if (r) { memset(r, 0, sizeof(*r)); r->site = (int32_t)s; }
        if (old[3] == 0.0f) {
            if (r) r->vacancy = 1;
            continue;
        }
        float np[4];
        if (m->constrain_to_x) or_random_X_point_f32(g, np);
        else or_random_sphere_point_f32(g, np);
        np[3] = old[3];
        float dE = (float)or_site_energy_f32(m, nb, nnb, lat, x, y, z, np, old);
        int accept;
        double u = 0.0;
        int drawn = 0;
        if (dE < 0.0) accept = 1;
        else {
            u = or_mt_real2(g);
            drawn = 1;
            accept = exp(-dE * m->beta) > u;
        }
        if (r) {
            r->prop[0] = np[0]; r->prop[1] = np[1]; r->prop[2] = np[2];
            r->dE = dE; r->u = u; r->u_drawn = drawn; r->accepted = accept;
        }
        if (accept) {
            old[0] = np[0]; old[1] = np[1]; old[2] = np[2];
            if (acc) (*acc)++;
        } else if (rej) (*rej)++;
    }
    return moves;
}

/* ------------------------------------------------------------------------ */
/* FP64 restatement. */

static inline double dotd(const double *a, const double *b)
{
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

double or_site_energy_f64(const or_model *m, const or_neighbour *nb, int nnb, const double *lat,
                          int x, int y, int z, const double newp[3], double *scale)
{
    const double *old = lat + 4 * site_index(m, x, y, z);
    double dE = 0.0, sc = 0.0;
    for (int i = 0; i < nnb; i++) {
        int dx = nb[i].dx, dy = nb[i].dy, dz = nb[i].dz;
        double d = sqrt((double)(dx * dx + dy * dy + dz * dz));
        const double *t = lat + 4 * site_index(m, wrap(x + dx, m->X), wrap(y + dy, m->Y),
                                               wrap(z + dz, m->Z));
        double n[3] = {dx / d, dy / d, dz / d};
        double ll = old[3] * t[3];
        double a = dotd(newp, t) - 3.0 * dotd(n, newp) * dotd(n, t);
        double b = dotd(old, t) - 3.0 * dotd(n, old) * dotd(n, t);
        double d3 = d * d * d;
        dE += ll * (a - b) / d3;
        sc += fabs(ll) * (fabs(dotd(newp, t)) + fabs(dotd(old, t)) +
                          3.0 * fabs(dotd(n, t)) * (fabs(dotd(n, newp)) + fabs(dotd(n, old)))) / d3;
        if (dx * dx + dy * dy + dz * dz == 1) {
            dE += -m->J * dotd(newp, t) + m->J * dotd(old, t);
            sc += fabs(m->J) * (fabs(dotd(newp, t)) + fabs(dotd(old, t)));
        }
    }
    dE += dotd(newp, m->E) - dotd(old, m->E);
    sc += fabs(dotd(newp, m->E)) + fabs(dotd(old, m->E));
    if (m->K > 0.0) {
        dE += -m->K * (fabs(newp[0]) - fabs(old[0]) + fabs(newp[1]) - fabs(old[1]));
        sc += m->K * (fabs(newp[0]) + fabs(old[0]) + fabs(newp[1]) + fabs(old[1]));
    }
    if (scale) *scale = sc;
    return dE;
}

void or_site_energy_batch_f64(const or_model *m, const or_neighbour *nb, int nnb,
                              const double *lat, int64_t n, const int32_t *site,
                              const double *prop, double *dE, double *scale)
{
    for (int64_t k = 0; k < n; k++) {
        int s = site[k];
        int z = s % m->Z, y = (s / m->Z) % m->Y, x = s / (m->Z * m->Y);
        dE[k] = or_site_energy_f64(m, nb, nnb, lat, x, y, z, prop + 3 * k, scale ? scale + k : NULL);
    }
}

/* Marsaglia (1972) sphere point in double from MT real1 draws. */
static void sphere_point_f64(or_mt19937 *g, double p[3])
{
    double x1, x2, s;
    do {
        x1 = 2.0 * or_mt_real1(g) - 1.0;
        x2 = 2.0 * or_mt_real1(g) - 1.0;
        s = x1 * x1 + x2 * x2;
    } while (s > 1.0);
    double r = sqrt(1.0 - s);
    p[0] = 2.0 * x1 * r;
    p[1] = 2.0 * x2 * r;
    p[2] = 1.0 - 2.0 * s;
}

int64_t or_mc_moves_f64(const or_model *m, const or_neighbour *nb, int nnb, double *lat,
                        or_mt19937 *g, int64_t moves, uint64_t *acc, uint64_t *rej)
{
    static const double ax[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
    for (int64_t k = 0; k < moves; k++) {
        int x = rand_int(g, m->X);
        int y = rand_int(g, m->Y);
        int z = rand_int(g, m->Z);
        double *old = lat + 4 * site_index(m, x, y, z);
        if (old[3] == 0.0) continue;
        double np[3];
        if (m->constrain_to_x) {
            int i = rand_int(g, 6);
            np[0] = ax[i][0]; np[1] = ax[i][1]; np[2] = ax[i][2];
        } else sphere_point_f64(g, np);
        double dE = or_site_energy_f64(m, nb, nnb, lat, x, y, z, np, NULL);
        if (dE < 0.0 || exp(-dE * m->beta) > or_mt_real2(g)) {
            old[0] = np[0]; old[1] = np[1]; old[2] = np[2];
            if (acc) (*acc)++;
        } else if (rej) (*rej)++;
    }
    return moves;
}

/* the disc point of candidate words (a0, a1), if it lies in the disc */
static int disc_point(uint32_t a0, uint32_t a1, double p[3])
{
    int64_t a = 2 * (int64_t)(a0 >> 6) + 1 - ((int64_t)1 << 26);
    int64_t b = 2 * (int64_t)(a1 >> 6) + 1 - ((int64_t)1 << 26);
    int64_t s = a * a + b * b;
    if (s >= ((int64_t)1 << 52)) return 0;
    double x1 = (double)a * (1.0 / 67108864.0);
    double x2 = (double)b * (1.0 / 67108864.0);
    double S = (double)s * (1.0 / 4503599627370496.0);
    double r = sqrt(1.0 - S);
    p[0] = (2.0 * x1) * r;
    p[1] = (2.0 * x2) * r;
    p[2] = 1.0 - 2.0 * S;
    return 1;
}

/* ------------------------------------------------------------------------ */
/* Counter-RNG draw for one site visit (the GPU sweep's proposal and u).
 * Philox4x32-10 with key = seed, counter = (site, sweep, replica, k).
 * u = 53-bit uniform from words 2,3 of k = 0.  Marsaglia (1972) disc point
 * from words 0,1 of k = 0, then words 0,1 and 2,3 of k = 1, 2, ... (two
 * candidates per retry call): x = a * 2^-26 with a odd in
 * (-2^26, 2^26), so the disc test a^2 + b^2 < 2^52 is exact in integers and
 * 1 - s, 1 - 2s are exact in double; the only rounded steps are one sqrt and
 * two multiplies, which makes the draw bit-reproducible on any IEEE machine. */
void or_counter_draw(uint64_t seed, uint32_t replica, uint32_t sweep, uint32_t site,
                     int constrain_to_x, double p[3], double *u)
{
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t ctr[4] = {site, sweep, replica, 0u};
    uint32_t w[4];
    or_philox4x32_10(ctr, key, w);
    *u = ((double)(w[2] >> 5) * 67108864.0 + (double)(w[3] >> 6)) * (1.0 / 9007199254740992.0);
    if (constrain_to_x) {
        static const double ax[6][3] = {{1, 0, 0}, {-1, 0, 0}, {0, 1, 0}, {0, -1, 0}, {0, 0, 1}, {0, 0, -1}};
        int i = (int)(w[0] % 6u);
        p[0] = ax[i][0]; p[1] = ax[i][1]; p[2] = ax[i][2];
        return;
    }
    /* candidates: (w0, w1) of k = 0, then (w0, w1) and (w2, w3) of k = 1, 2, ... */
    if (disc_point(w[0], w[1], p)) return;
    for (uint32_t k = 1; k <= 32; k++) {
        ctr[3] = k;
        or_philox4x32_10(ctr, key, w);
        if (disc_point(w[0], w[1], p) || disc_point(w[2], w[3], p)) return;
    }
    p[0] = 0.0; p[1] = 0.0; p[2] = 1.0; /* 65 rejections in a row: p ~ 1e-43 */
}

void or_sweep_ordered_f64(const or_model *m, const or_neighbour *nb, int nnb, double *lat,
                          const int32_t *order, int64_t n, uint64_t seed, uint32_t replica,
                          uint32_t sweep, uint64_t *acc, uint64_t *rej)
{
    for (int64_t k = 0; k < n; k++) {
        int s = order[k];
        double *old = lat + 4 * (int64_t)s;
        if (old[3] == 0.0) continue;
        int z = s % m->Z, y = (s / m->Z) % m->Y, x = s / (m->Z * m->Y);
        double np[3], u;
        or_counter_draw(seed, replica, sweep, (uint32_t)s, m->constrain_to_x, np, &u);
        double dE = or_site_energy_f64(m, nb, nnb, lat, x, y, z, np, NULL);
        if (dE < 0.0 || exp(-dE * m->beta) > u) {
            old[0] = np[0]; old[1] = np[1]; old[2] = np[2];
            if (acc) (*acc)++;
        } else if (rej) (*rej)++;
    }
}

/* ------------------------------------------------------------------------ */
/* Hamiltonian implied by site_energy:
 *   E = 1/2 sum_i sum_r L_i L_j [p_i.p_j - 3(n.p_i)(n.p_j)]/d^3
 *     - 1/2 J sum_i sum_{|r|=1} p_i.p_j + sum_i p_i.E - [K>0] K sum_i (|p_ix|+|p_iy|)
 * so that site_energy(i) = E(p_i -> p') - E(p_i) when no offset aliases. */
typedef struct { double s, c; } ksum;
static inline void kadd(ksum *k, double v)
{
    double t = k->s + v;
    if (fabs(k->s) >= fabs(v)) k->c += (k->s - t) + v;
    else k->c += (v - t) + k->s;
    k->s = t;
}

static void site_fields(const or_model *m, const or_neighbour *nb, int nnb, const double *lat,
                        int x, int y, int z, double h[3], double g[3])
{
    h[0] = h[1] = h[2] = 0.0;
    g[0] = g[1] = g[2] = 0.0;
    for (int i = 0; i < nnb; i++) {
        int dx = nb[i].dx, dy = nb[i].dy, dz = nb[i].dz;
        double d = sqrt((double)(dx * dx + dy * dy + dz * dz));
        const double *t = lat + 4 * site_index(m, wrap(x + dx, m->X), wrap(y + dy, m->Y),
                                               wrap(z + dz, m->Z));
        double n[3] = {dx / d, dy / d, dz / d};
        double w = t[3] / (d * d * d);
        double s = dotd(n, t);
        for (int c = 0; c < 3; c++) h[c] += w * (t[c] - 3.0 * n[c] * s);
        if (dx * dx + dy * dy + dz * dz == 1)
            for (int c = 0; c < 3; c++) g[c] += t[c];
    }
}

double or_energy_f64(const or_model *m, const or_neighbour *nb, int nnb, const double *lat,
                     double parts[4])
{
    ksum kd = {0, 0}, kc = {0, 0}, kf = {0, 0}, kk = {0, 0};
    for (int x = 0; x < m->X; x++)
        for (int y = 0; y < m->Y; y++)
            for (int z = 0; z < m->Z; z++) {
                const double *p = lat + 4 * site_index(m, x, y, z);
                double h[3], g[3];
                site_fields(m, nb, nnb, lat, x, y, z, h, g);
                kadd(&kd, 0.5 * p[3] * dotd(p, h));
                kadd(&kc, -0.5 * m->J * dotd(p, g));
                kadd(&kf, dotd(p, m->E));
                if (m->K > 0.0) kadd(&kk, -m->K * (fabs(p[0]) + fabs(p[1])));
            }
    double P[4] = {kd.s + kd.c, kc.s + kc.c, kf.s + kf.c, kk.s + kk.c};
    if (parts) memcpy(parts, P, sizeof(P));
    ksum t = {0, 0};
    for (int i = 0; i < 4; i++) kadd(&t, P[i]);
    return t.s + t.c;
}

static inline double sgn(double v) { return (v > 0.0) - (v < 0.0); }

void or_gradient_f64(const or_model *m, const or_neighbour *nb, int nnb, const double *lat,
                     double *grad, double dparams[6])
{
    ksum kJ = {0, 0}, kE[3] = {{0, 0}, {0, 0}, {0, 0}}, ks = {0, 0}, kK = {0, 0};
    for (int x = 0; x < m->X; x++)
        for (int y = 0; y < m->Y; y++)
            for (int z = 0; z < m->Z; z++) {
                int64_t si = site_index(m, x, y, z);
                const double *p = lat + 4 * si;
                double h[3], g[3];
                site_fields(m, nb, nnb, lat, x, y, z, h, g);
                for (int c = 0; c < 3; c++) {
                    double v = p[3] * h[c] - m->J * g[c] + m->E[c];
                    if (m->K > 0.0 && c < 2) v -= m->K * sgn(p[c]);
                    if (grad) grad[3 * si + c] = v;
                    kadd(&kE[c], p[c]);
                }
                kadd(&kJ, -0.5 * dotd(p, g));
                kadd(&ks, p[3] * dotd(p, h)); /* 2 * E_dd */
                if (m->K > 0.0) kadd(&kK, -(fabs(p[0]) + fabs(p[1])));
            }
    if (dparams) {
        dparams[0] = kJ.s + kJ.c;
        for (int c = 0; c < 3; c++) dparams[1 + c] = kE[c].s + kE[c].c;
        dparams[4] = ks.s + ks.c;
        dparams[5] = kK.s + kK.c;
    }
}

double or_local_energy_f64(const or_model *m, const or_neighbour *nb, int nnb, const double *lat,
                           int x, int y, int z, const double p[3])
{
    const double *old = lat + 4 * site_index(m, x, y, z);
    double h[3], g[3];
    site_fields(m, nb, nnb, lat, x, y, z, h, g);
    double e = old[3] * dotd(p, h) - m->J * dotd(p, g) + dotd(p, m->E);
    if (m->K > 0.0) e -= m->K * (fabs(p[0]) + fabs(p[1]));
    return e;
}

/* ------------------------------------------------------------------------ */
void or_moment_sum(const double *lat, int64_t N, double M[3])
{
    ksum k[3] = {{0, 0}, {0, 0}, {0, 0}};
    for (int64_t i = 0; i < N; i++)
        for (int c = 0; c < 3; c++) kadd(&k[c], lat[4 * i + c]);
    for (int c = 0; c < 3; c++) M[c] = k[c].s + k[c].c;
}

/* radial_order_parameter (src/starrynight-analysis.c:528-598) in double.  The
 * dz loop runs even when Z == 1, as in the reference. */
void or_radial_order_f64(int X, int Y, int Z, const double *lat, double fe[81], double afe[81],
                         int64_t count[81])
{
    const int C = 9;
    for (int i = 0; i < 81; i++) { fe[i] = 0; afe[i] = 0; count[i] = 0; }
    or_model m = {X, Y, Z, 0, 0, {0, 0, 0}, 0, 0, 0};
    for (int x = 0; x < X; x++)
        for (int y = 0; y < Y; y++)
            for (int z = 0; z < Z; z++) {
                const double *a = lat + 4 * site_index(&m, x, y, z);
                for (int dx = -C; dx <= C; dx++)
                    for (int dy = -C; dy <= C; dy++)
                        for (int dz = -C; dz <= C; dz++) {
                            int d2 = dx * dx + dy * dy + dz * dz;
                            if (d2 > C * C || d2 >= 81) continue;
                            /* + 9L instead of the reference's + L: the same site
                             * where the reference's index is valid (L >= 9, or
                             * Z == 1), a periodic wrap where it is not */
                            const double *b = lat + 4 * site_index(&m, (x + dx + 9 * X) % X,
                                                                   (y + dy + 9 * Y) % Y, (z + dz + 9 * Z) % Z);
                            double f = dotd(a, b);
                            double d = sqrt((double)d2);
                            if (d == 0) d = 1;
                            double n[3] = {dx / d, dy / d, dz / d};
                            fe[d2] += f;
                            afe[d2] += f - 3.0 * dotd(n, a) * dotd(n, b);
                            count[d2]++;
                        }
            }
    for (int i = 0; i < 81; i++)
        if (count[i] > 0) { fe[i] /= (double)count[i]; afe[i] /= (double)count[i]; }
}

/* dipole_potential (src/starrynight-analysis.c:65-94), FP64.  Offsets in
 * [-6,6]^3 with 0 < |r| <= 6; periodic wrap (the reference's (X+x+dx)%X is
 * the same site for X >= 6). */
void or_dipole_potential_f64(int X, int Y, int Z, const double *lat, double *pot)
{
    const int MAX = 6;
    or_model m = {X, Y, Z, 0, 0, {0, 0, 0}, 0, 0, 0};
    for (int x = 0; x < X; x++)
        for (int y = 0; y < Y; y++)
            for (int z = 0; z < Z; z++) {
                double acc = 0.0;
                for (int dx = -MAX; dx <= MAX; dx++)
                    for (int dy = -MAX; dy <= MAX; dy++)
                        for (int dz = -MAX; dz <= MAX; dz++) {
                            int d2 = dx * dx + dy * dy + dz * dz;
                            if (d2 == 0 || d2 > MAX * MAX) continue;
                            const double *b = lat + 4 * site_index(&m, (x + dx + 9 * X) % X,
                                                                   (y + dy + 9 * Y) % Y, (z + dz + 9 * Z) % Z);
                            double d = sqrt((double)d2);
                            acc += b[3] * (b[0] * dx + b[1] * dy + b[2] * dz) / (d * d * d);
                        }
                pot[site_index(&m, x, y, z)] = acc;
            }
}

/* dipole_electricfield (src/starrynight-analysis.c:393-457), FP64 vector. */
void or_dipole_efield_f64(int X, int Y, int Z, int cutoff, const double *lat, double *E)
{
    or_model m = {X, Y, Z, 0, 0, {0, 0, 0}, 0, 0, 0};
    for (int x = 0; x < X; x++)
        for (int y = 0; y < Y; y++)
            for (int z = 0; z < Z; z++) {
                double e[3] = {0, 0, 0};
                for (int dx = -cutoff; dx <= cutoff; dx++)
                    for (int dy = -cutoff; dy <= cutoff; dy++)
                        for (int dz = -cutoff; dz <= cutoff; dz++) {
                            int d2 = dx * dx + dy * dy + dz * dz;
                            if (d2 == 0 || d2 > cutoff * cutoff) continue;
                            const double *b = lat + 4 * site_index(&m, (x + dx + 9 * X) % X,
                                                                   (y + dy + 9 * Y) % Y, (z + dz + 9 * Z) % Z);
                            double d = sqrt((double)d2);
                            double n[3] = {dx / d, dy / d, dz / d};
                            double radial = dotd(n, b);
                            double d3 = d * d * d;
                            for (int k = 0; k < 3; k++) e[k] += (3.0 * n[k] * radial - b[k]) / d3;
                        }
                const double *a = lat + 4 * site_index(&m, x, y, z);
                double *o = E + 3 * site_index(&m, x, y, z);
                for (int k = 0; k < 3; k++) o[k] = e[k] - a[k] / 3.0;
            }
}

/* recombination_calculator (src/starrynight-analysis.c:97-164), FP64:
 * pot_i = 0.165 / 5 * dipole_potential_i (eV), beta = 1 / 0.025 eV^-1,
 * Boltzmann and Fermi-Dirac partition sums, FD densities normalised by their
 * partition sums, R_Boltz = N^2 / (ZBe ZBh), R_FD = N sum_i e_i h_i. */
void or_recombination_f64(int X, int Y, int Z, const double *lat, double out[8])
{
    const int64_t N = (int64_t)X * Y * Z;
    double *pot = (double *)malloc(sizeof(double) * N);
    or_dipole_potential_f64(X, Y, Z, lat, pot);
    const double BETA = 1 / (0.025), scale = 0.165 / 5;
    double ZBe = 0, ZBh = 0, ZFDe = 0, ZFDh = 0;
    for (int64_t i = 0; i < N; i++) {
        const double p = scale * pot[i];
        ZBe += exp(-p * BETA);
        ZBh += exp(p * BETA);
        ZFDe += 1.0 / (exp(p * BETA) + 1.0);
        ZFDh += 1.0 / (exp(-p * BETA) + 1.0);
    }
    double et = 0, ht = 0, rt = 0;
    for (int64_t i = 0; i < N; i++) {
        const double p = scale * pot[i];
        const double e = 1.0 / (exp(p * BETA) + 1.0) / ZFDe, h = 1.0 / (exp(-p * BETA) + 1.0) / ZFDh;
        et += e; ht += h; rt += e * h;
    }
    free(pot);
    out[0] = ZBe; out[1] = ZBh; out[2] = ZFDe; out[3] = ZFDh;
    out[4] = (double)N * (double)N / (ZBe * ZBh);
    out[5] = (double)N * rt;
    out[6] = et; out[7] = ht;
}
