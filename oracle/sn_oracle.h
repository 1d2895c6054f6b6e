/*
 * sn_oracle.h -- CPU oracle for the StarryNight Metropolis dipole-lattice path.
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, __graft_entry__.smoke()
 * and the cpu_baseline leg of bench.py may load, link or run anything built
 * from this directory, and only as the checker (or as the timed CPU baseline).
 * The product library (starrynight_amd/) never links it.
 *
 * Two restatements of the reference hot path live here:
 *   *_f32  follow the reference operation by operation: float lattice, float
 *          dot products without FMA, per-neighbour terms in float summed in
 *          double, dE rounded to float, MT19937 draws in the reference order.
 *          Used for replay traces (decision-for-decision oracle) and as the
 *          CPU baseline.
 *   *_f64  the same algorithm with every quantity in IEEE double; this is
 *          what the FP64 GPU kernels are compared against at 1e-12.
 * Plus the counter-RNG (Philox4x32-10) ordered sweep that restates the GPU
 * colour-sweep semantics, total energy / gradient (the Hamiltonian implied by
 * site_energy) and the observables used by the statistical oracle.
 *
 * Parity pins for every function are listed in the header comment of
 * sn_oracle.c next to the function they check.
 */
#ifndef SN_ORACLE_H
#define SN_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- MT19937 (src/mt19937ar-cok.c:63-197) ---- */
typedef struct {
    uint32_t mt[624];
    int idx;
} or_mt19937;
void     or_mt_seed(or_mt19937 *g, uint32_t s);  /* init_genrand */
uint32_t or_mt_u32(or_mt19937 *g);               /* genrand_int32 */
double   or_mt_real1(or_mt19937 *g);             /* genrand_real1, [0,1] */
double   or_mt_real2(or_mt19937 *g);             /* genrand_real2, [0,1) */
or_mt19937 *or_mt_new(uint32_t s);
void     or_mt_free(or_mt19937 *g);
void     or_mt_fill_u32(or_mt19937 *g, int64_t n, uint32_t *out);

/* ---- Philox4x32-10 (Salmon et al., SC'11) ---- */
void or_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);

/* ---- neighbour table (src/starrynight-montecarlo-core.c:38-72) ---- */
typedef struct {
    int32_t dx, dy, dz;
    float d;
} or_neighbour;
/* Returns the number of offsets written (dx, dy, dz ascending), -1 on overflow. */
int or_gen_neighbour(int rc, int Z, or_neighbour *out, int cap);

/* ---- model parameters (the reference's globals, src/starrynight-config.c:12-93) ---- */
typedef struct {
    int32_t X, Y, Z, rc;
    double J;       /* CageStrain */
    double E[3];    /* Efield */
    double K;       /* in-plane anisotropy, applied only when K > 0 */
    int32_t constrain_to_x;
    double beta;    /* 300/T (T=0 -> +inf) */
} or_model;

/* ---- float-faithful restatement; lattice = float[N][4] (x, y, z, length) ---- */
double or_site_energy_f32(const or_model *m, const or_neighbour *nb, int nnb, const float *lat,
                          int x, int y, int z, const float newp[4], const float oldp[4]);
void or_random_sphere_point_f32(or_mt19937 *g, float p[3]);
void or_random_X_point_f32(or_mt19937 *g, float p[3]);
void or_init_lattice_random_f32(or_mt19937 *g, int X, int Y, int Z, float *lat);
int  or_solid_solution_f32(or_mt19937 *g, int X, int Y, int Z, int n, const float *len,
                           const float *prev, float *lat);
int  or_species_walk_f32(float sample, int n, const float *prev);

typedef struct {
    int32_t site;      /* (x*Y + y)*Z + z */
    int32_t vacancy;   /* 1: early return, nothing else drawn */
    float prop[3];     /* proposal (float bits as the reference holds them) */
    float dE;          /* float-rounded site_energy */
    double u;          /* genrand_real2 value if drawn */
    int32_t u_drawn;
    int32_t accepted;
} or_move_rec;
/* MC_moves(moves).  trace may be NULL.  Returns number of attempted moves. */
int64_t or_mc_moves_f32(const or_model *m, const or_neighbour *nb, int nnb, float *lat,
                        or_mt19937 *g, int64_t moves, uint64_t *acc, uint64_t *rej,
                        or_move_rec *trace);

/* ---- FP64 restatement; lattice = double[N][4] ---- */
/* dE for site (x,y,z) changing to newp (old value read from the lattice).
 * *scale receives sum of |term| (the rounding scale used for 1e-12 checks). */
double or_site_energy_f64(const or_model *m, const or_neighbour *nb, int nnb, const double *lat,
                          int x, int y, int z, const double newp[3], double *scale);
void or_site_energy_batch_f64(const or_model *m, const or_neighbour *nb, int nnb,
                              const double *lat, int64_t n, const int32_t *site,
                              const double *prop, double *dE, double *scale);
/* Reference algorithm (random site with replacement, MT19937) in FP64. */
int64_t or_mc_moves_f64(const or_model *m, const or_neighbour *nb, int nnb, double *lat,
                        or_mt19937 *g, int64_t moves, uint64_t *acc, uint64_t *rej);

/* ---- counter-RNG ordered sweep (the GPU colour-sweep semantics) ---- */
void or_counter_draw(uint64_t seed, uint32_t replica, uint32_t sweep, uint32_t site,
                     int constrain_to_x, double p[3], double *u);
void or_sweep_ordered_f64(const or_model *m, const or_neighbour *nb, int nnb, double *lat,
                          const int32_t *order, int64_t n, uint64_t seed, uint32_t replica,
                          uint32_t sweep, uint64_t *acc, uint64_t *rej);

/* ---- Hamiltonian implied by site_energy: energy and gradient ---- */
/* parts = {E_dd, E_cage, E_field, E_K}; compensated summation. */
double or_energy_f64(const or_model *m, const or_neighbour *nb, int nnb, const double *lat,
                     double parts[4]);
/* grad: N x 3, dE/dp_i.  dparams = {dE/dJ, dE/dEx, dE/dEy, dE/dEz, dE/ds, dE/dK}. */
void or_gradient_f64(const or_model *m, const or_neighbour *nb, int nnb, const double *lat,
                     double *grad, double dparams[6]);
/* local energy of one site as a function of its own dipole: the part of E that
 * depends on p_i (linear in p_i when K == 0). */
double or_local_energy_f64(const or_model *m, const or_neighbour *nb, int nnb, const double *lat,
                           int x, int y, int z, const double p[3]);

/* ---- observables (src/starrynight-analysis.c:48-62, 506-598) ---- */
void or_moment_sum(const double *lat, int64_t N, double M[3]);
/* radial_order_parameter: bins r^2 = 0..80 (bin 81 never reported). */
void or_radial_order_f64(int X, int Y, int Z, const double *lat, double fe[81], double afe[81],
                         int64_t count[81]);
/* dipole_potential (src/starrynight-analysis.c:65-94) at every site, FP64:
 * pot_i = sum_{0 < |r| <= 6} L_j (p_j . r) / |r|^3, j = i + r (the dz range
 * runs when Z == 1, onto the same plane). */
void or_dipole_potential_f64(int X, int Y, int Z, const double *lat, double *pot);
/* dipole_electricfield (src/starrynight-analysis.c:393-457) at every site,
 * FP64, as a vector: E_i = sum_{0 < |r| <= cutoff} (3 n (n . p_j) - p_j) / |r|^3
 * - p_i / 3 (lengths are not applied, as in the reference); the reference
 * returns |E_i|. */
void or_dipole_efield_f64(int X, int Y, int Z, int cutoff, const double *lat, double *E);
/* recombination_calculator (src/starrynight-analysis.c:97-164), FP64, from
 * the dipole potential: out = {ZBe, ZBh, ZFDe, ZFDh, R_Boltz, R_FD,
 * FD-Total-electron, FD-Total-hole}. */
void or_recombination_f64(int X, int Y, int Z, const double *lat, double out[8]);

#ifdef __cplusplus
}
#endif
#endif
