/*
 * starrynight_amd.h -- C ABI of the MI355X-native StarryNight Metropolis path.
 *
 * A drop-in for the reference's hot path (src/starrynight-montecarlo-core.c):
 * the reference binds MC_moves / site_energy as file-static functions that talk
 * through globals inside one translation unit; this library replaces them with
 * plain-pointer entry points over a context that owns device memory, a HIP
 * stream and the colour schedule.  All arithmetic is IEEE double.
 *
 * Conventions
 *  - Every entry point returns an int status (SN_OK = 0); nothing calls exit().
 *    sn_last_error(ctx) gives a message for the last failure on that context.
 *  - Lattice exchange format: double[N][4] = (x, y, z, length) per site,
 *    site index (x*Y + y)*Z + z, as the reference's lattice[x][y][z]
 *    (src/starrynight-main.c:155-161).  length == 0 marks a vacancy.
 *  - Host buffers are owned by the caller; the library copies to / from HBM.
 *  - One host thread per context; contexts are independent.  Kernels run on a
 *    context-owned stream; calls that return data synchronise it.
 *  - A context holds R >= 1 replicas (independent lattices, possibly of
 *    different sizes and parameters) that share DipoleCutOff and 2D/3D-ness.
 */
#ifndef STARRYNIGHT_AMD_H
#define STARRYNIGHT_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SN_OK 0
#define SN_ERR_INVALID 1     /* bad argument or configuration */
#define SN_ERR_HIP 2         /* HIP runtime error (no device, launch failure, ...) */
#define SN_ERR_NOMEM 3       /* host or device allocation failed */
#define SN_ERR_RANGE 4       /* index or size out of range */
#define SN_ERR_UNSUPPORTED 5 /* configuration outside what the kernels handle */

typedef struct sn_ctx sn_ctx;

/* Model parameters of one replica.  Mirrors the reference globals
 * (src/starrynight-config.c:12-93) that site_energy / MC_move read. */
typedef struct sn_params {
    int32_t X, Y, Z;          /* lattice size; Z == 1 is the 2D lattice (config keys X, Y, Z) */
    int32_t dipole_cutoff;    /* DipoleCutOff (integer, lattice units) */
    double cage_strain;       /* CageStrain, the nearest-neighbour coupling J */
    double efield[3];         /* Efield.{x,y,z} */
    double K;                 /* in-plane anisotropy, applied only when K > 0 */
    int32_t constrain_to_x;   /* ConstrainToX: proposals restricted to +-x, +-y, +-z */
    double temperature;       /* T in kelvin; beta = 300 / T, T == 0 gives beta = +inf */
    uint64_t seed;            /* key of the counter-based RNG for this replica */
} sn_params;

/* Sweep-kernel tuning knobs.  Zero fields mean the built-in default;
 * sn_autotune picks them by measurement. */
typedef struct sn_tuning {
    int32_t tile[3];          /* target tile edge per axis (sites), >= DipoleCutOff */
    int32_t lanes_per_site;   /* lanes that share one site's neighbour sum: 2..64 */
    int32_t fine_colouring;   /* 0 auto, 1 cubic (rc+1)^3, 2 min-index sublattice, 3 BCC (rc=3) */
    int32_t use_graph;        /* 0 or 1: replay sweeps from a captured hipGraph; -1: direct launches */
    int32_t kernel;           /* 0 auto (kernel 3 for replicas of up to 4096 sites when it can
                                 colour them, else kernel 2; NULL tuning is the same),
                                 1 field-update tile kernel, 2 per-phase gather tile kernel
                                 (both visit sites in the same order with the same draws),
                                 3 whole-lattice sweeps of small replicas: one workgroup per replica
                                 holds it in LDS, colours of a periodic sublattice or (up to 16
                                 neighbour pairs) a greedy periodic colouring, whichever needs
                                 fewer steps; many sweeps per launch (refused when a replica does
                                 not fit or no periodic colouring exists; sn_schedule_order
                                 reports its order) */
    int32_t block_threads;    /* gather kernel's threads per tile: 0 or 256, or 512 */
} sn_tuning;

/* ---- lifecycle ---------------------------------------------------------- */

/* Create a context with n_replicas replicas (params is an array of that many).
 * Replaces the reference's load_config() + malloc of lattice + gen_neighbour()
 * (src/starrynight-main.c:137-180, src/starrynight-montecarlo-core.c:38-72).
 * tuning may be NULL.  Lattices start zeroed; load them with sn_set_lattice or
 * sn_init_lattice_random.  On failure *out may still hold the partial context
 * (for sn_last_error); release it with sn_destroy. */
int sn_create(const sn_params *params, int32_t n_replicas, const sn_tuning *tuning, sn_ctx **out);
int sn_destroy(sn_ctx *ctx);
const char *sn_last_error(const sn_ctx *ctx);
const char *sn_status_string(int status);

/* ---- lattice state ------------------------------------------------------ */

/* Copy replica r's lattice in / out (double[N][4], see conventions).
 * Replaces direct access to the global lattice[x][y][z] (src/starrynight-config.c:32-36). */
int sn_set_lattice(sn_ctx *ctx, int32_t replica, const double *xyzl);
int sn_get_lattice(sn_ctx *ctx, int32_t replica, double *xyzl);
/* sn_set_lattice from a device buffer (double4[N] on the context's device;
 * the caller's writes to it must be complete).  Same checks: non-finite
 * values are refused and the lengths decide uniform or mixed. */
int sn_set_lattice_device(sn_ctx *ctx, int32_t replica, const void *xyzl_dev);
/* Device address of replica r's lattice (double4[N] in HBM), for zero-copy
 * consumers on the same device.  Valid until sn_destroy. */
int sn_lattice_device_ptr(sn_ctx *ctx, int32_t replica, void **dev_ptr);

/* Random unit dipoles (Marsaglia on the counter RNG, stream of the replica's
 * seed) and lengths drawn from the species table (the Dipoles / Prevalence
 * solid solution), on the device.  Replaces initialise_lattice_random +
 * solid_solution (src/starrynight-lattice.c:25-37, 139-171) for runs that do
 * not need the MT19937 initial state. */
int sn_init_lattice_random(sn_ctx *ctx, int32_t replica, int32_t n_species,
                           const double *lengths, const double *prevalence);

/* Reference-compatible initial state, on the host (no context, no device):
 * init_genrand(seed), the InitialLattice pattern ("random", "ferroelectric",
 * "buckled", "antiferro_wall", "ferro_wall", "antiferro_slip", "spectrum",
 * "slab_delete"; any other name is "random"), then solid_solution with the
 * Dipoles / Prevalence table -- the float values the reference holds after
 * src/starrynight-main.c:172-205 (src/starrynight-lattice.c:25-171), written as
 * double[N][4].  The reference seeds with 0xDEADBEEF + T.  SN_ERR_INVALID when
 * Prevalence sums below the drawn sample (the reference reads past its table). */
int sn_reference_initial_lattice(int32_t X, int32_t Y, int32_t Z, const char *initial_lattice,
                                 uint32_t seed, int32_t n_species, const double *lengths,
                                 const double *prevalence, double *xyzl);

/* ---- parameters --------------------------------------------------------- */

/* Change T, Efield and CageStrain of replica r between sweeps (temperature
 * schedules, field ramps) without rebuilding the schedule or the graph.
 * Replaces assignments to the globals beta / Efield / CageStrain
 * (src/starrynight-main.c:215, 239). */
int sn_set_params(sn_ctx *ctx, int32_t replica, double temperature, const double efield[3],
                  double cage_strain);

/* ---- the hot path ------------------------------------------------------- */

/* site_energy drop-in (src/starrynight-montecarlo-core.c:76-141): for each of
 * n queries, dE of replacing the dipole at site[k] of replica r by prop[k]
 * (unit vector, 3 doubles), read against the current lattice.  Host buffers. */
int sn_site_energy_batch(sn_ctx *ctx, int32_t replica, int64_t n, const int32_t *site,
                         const double *prop, double *dE);
/* Same with device pointers (site: int32[n], prop: double[n][3], dE: double[n]),
 * enqueued on the context stream without synchronising.  The site indices are
 * not checked on the host: a query whose index lies outside 0..N-1 reads
 * nothing and gets dE = NaN. */
int sn_site_energy_batch_device(sn_ctx *ctx, int32_t replica, int64_t n, const int32_t *site_dev,
                                const double *prop_dev, double *dE_dev);

/* MC_moves drop-in (src/starrynight-montecarlo-core.c:143-191) in whole
 * sweeps: every replica advances nsweeps sweeps; one sweep attempts one
 * Metropolis update of every site (N attempts, vacancies counted as attempts
 * but neither accepted nor rejected, as in the reference).  Sites are visited
 * in the conflict-free colour schedule; proposals and the acceptance uniform
 * come from Philox4x32-10 keyed by (seed; site, sweep, replica).  Asynchronous:
 * returns once the work is enqueued; sn_sync waits. */
int sn_mc_sweeps(sn_ctx *ctx, int32_t nsweeps);
int sn_sync(sn_ctx *ctx);

/* Replay of a recorded reference move sequence on replica r
 * (src/starrynight-montecarlo-core.c:151-191), one move after another:
 * site[k], proposal prop[k] (3 doubles), u[k] = the genrand_real2 value the
 * reference drew (ignored when u_drawn[k] == 0), ref_accept[k] = the
 * reference's decision.  dE is evaluated in FP64; a move whose FP64 decision
 * is within the float rounding bound of the reference's threshold is a near
 * tie and takes the reference's decision.  decision_out[k] receives the
 * decision applied; *mismatches counts decisions outside that bound that
 * differ from the reference; *near_ties counts the ties.  Host buffers. */
int sn_mc_replay(sn_ctx *ctx, int32_t replica, int64_t n, const int32_t *site, const double *prop,
                 const double *u, const uint8_t *u_drawn, const uint8_t *ref_accept,
                 uint8_t *decision_out, uint64_t *mismatches, uint64_t *near_ties);

/* Accept / reject / attempted counters of replica r since creation or the last
 * reset (the reference's ACCEPT / REJECT globals, src/starrynight-config.c:26-27). */
int sn_counters(sn_ctx *ctx, int32_t replica, uint64_t *accept, uint64_t *reject,
                uint64_t *attempted);
int sn_reset_counters(sn_ctx *ctx);

/* Sweep index of replica r: the sweep counter that keys its counter-RNG draws
 * (the next sweep draws as sweep `index` of the stream; 32 bits, it wraps).
 * Setting it with the saved lattice resumes a run exactly where it stopped:
 * the reference keeps only its seed and restarts its generator
 * (src/starrynight-main.c:172-176), so it has no counterpart. */
int sn_get_sweep_index(sn_ctx *ctx, int32_t replica, uint64_t *index);
int sn_set_sweep_index(sn_ctx *ctx, int32_t replica, uint64_t index);

/* ---- energy and adjoint (no reference counterpart: the reference has no
 * working total-energy function, src/starrynight-main.c:63) ---------------- */

/* Total energy of replica r under the Hamiltonian implied by site_energy;
 * parts = {dipole-dipole, cage strain, field, K} (may be NULL). */
int sn_energy(sn_ctx *ctx, int32_t replica, double *energy, double parts[4]);
/* Adjoint: grad_p (double[N][3], may be NULL) = dE/dp_i on every site (the
 * local field is its negative), and dparams = {dE/dCageStrain, dE/dEx, dE/dEy,
 * dE/dEz, dE/ds, dE/dK} where s scales every dipole length (at s = 1).
 * energy (may be NULL) receives the total energy from the same pass. */
int sn_gradient(sn_ctx *ctx, int32_t replica, double *grad_p, double dparams[6], double *energy);
/* sn_gradient with grad_dev (double[N][3] on the context's device, may be
 * NULL) written in place: no host copy of the N x 3 field.  Returns after
 * grad_dev is written. */
int sn_gradient_device(sn_ctx *ctx, int32_t replica, void *grad_dev, double dparams[6], double *energy);
/* Every replica's adjoint in one pass (one launch over all replicas' tiles).
 * grad_dev (device, may be NULL) holds 3 doubles per site of all replicas in
 * lattice order: replica r's rows start after the sites of replicas 0..r-1.
 * dparams (double[R][6]) and energy (double[R]) may be NULL. */
int sn_gradient_all_device(sn_ctx *ctx, void *grad_dev, double *dparams, double *energy);

/* sn_gradient_all_device that allocates nothing, for callers that cannot call
 * hipMalloc (code under HIP graph capture, tight serving loops, harnesses that
 * forbid allocation).  It has no counterpart in the reference.  The per-
 * workgroup partial sums live in a fixed device array of SN_NOALLOC_PARTIALS
 * doubles (1 MiB) per GPU, held until the process ends and shared by every
 * context on that GPU; the energy sums go through buffers sn_create already
 * allocated.  It needs 10 doubles per tile of all replicas (or per 256 sites
 * of the largest replica when the lattices have no tiling) and returns
 * SN_ERR_RANGE beyond the array.  Calls that overlap on one GPU (contexts used
 * from several host threads) race on the array.  Outputs as
 * sn_gradient_all_device. */
#define SN_NOALLOC_PARTIALS 131072
int sn_gradient_all_device_noalloc(sn_ctx *ctx, void *grad_dev, double *dparams, double *energy);

/* Sum of dipole vectors of replica r (M = sum_i p_i, vacancies included as the
 * reference's landau_order / polarisation do, src/starrynight-analysis.c:48-62,
 * 506-526). */
int sn_moment(sn_ctx *ctx, int32_t replica, double M[3]);

/* radial_order_parameter (src/starrynight-analysis.c:528-598) of replica r:
 * for every r^2 = dx^2+dy^2+dz^2 in 0..80 over offsets in [-9,9]^3, the mean
 * over all (site, offset) pairs of p_i.p_j (fe) and p_i.p_j - 3(n.p_i)(n.p_j)
 * (afe), and the pair count.  Bins with count 0 hold 0.  Accumulates in FP64
 * (the reference accumulates in float); the file the reference appends these
 * rows to is written by the caller.  Axes shorter than 9 wrap periodically
 * (the reference indexes out of its array there). */
int sn_radial_order(sn_ctx *ctx, int32_t replica, double fe[81], double afe[81], int64_t count[81]);

/* dipole_potential (src/starrynight-analysis.c:65-94) at every site of replica
 * r: pot[i] = sum_{0 < |r| <= 6} L_j (p_j . r) / |r|^3, j = i + r, in FP64
 * (the reference sums float terms in double).  pot: double[N], site order as
 * the lattice.  The reference writes these values as XYZ maps and .cube files
 * (:264-308); the caller formats them. */
int sn_dipole_potential(sn_ctx *ctx, int32_t replica, double *pot);
/* dipole_electricfield (src/starrynight-analysis.c:393-457) at every site, as
 * a vector: E[i] = sum_{0 < |r| <= cutoff} (3 n (n . p_j) - p_j) / |r|^3 - p_i / 3
 * (lengths not applied, as in the reference, which reports |E[i]|).
 * E: double[N][3]; cutoff in 1..9. */
int sn_dipole_efield(sn_ctx *ctx, int32_t replica, int32_t cutoff, double *E);
/* recombination_calculator (src/starrynight-analysis.c:97-164) of replica r
 * from its dipole potential (0.165 / 5 eV per unit, kT = 0.025 eV), FP64:
 * out = {ZBe, ZBh, ZFDe, ZFDh, R_Boltz = N^2 / (ZBe ZBh), R_FD = N sum e_i h_i,
 * FD-Total-electron, FD-Total-hole} -- the numbers of its log line. */
int sn_recombination(sn_ctx *ctx, int32_t replica, double out[8]);

/* ---- schedule introspection (pure host logic, no device work) ----------- */

/* Number of sites N of replica r and number of sequential phases per sweep. */
int sn_schedule_info(sn_ctx *ctx, int32_t replica, int64_t *n_sites, int32_t *n_phases);
/* The order in which one sweep visits replica r's sites: order[N] is a
 * permutation of 0..N-1, grouped into phases (phase_start[n_phases + 1]);
 * sites within a phase are pairwise farther apart than DipoleCutOff (through
 * the periodic boundary), so visiting them in any order, or all at once, is
 * the same update. */
int sn_schedule_order(sn_ctx *ctx, int32_t replica, int32_t *order, int32_t *phase_start);

/* The same visit order without a context or a device: plans the schedule for
 * the given replicas / tuning on the host and reports replica r's order.
 * Call with order == NULL first to get *n_sites and *n_phases.  No counterpart
 * in the reference (its MC_move picks sites at random). */
int sn_plan_schedule(const sn_params *params, int32_t n_replicas, const sn_tuning *tuning,
                     int32_t replica, int64_t *n_sites, int32_t *n_phases, int32_t *order,
                     int32_t *phase_start);

/* Stream the context's kernels run on (a hipStream_t), for event timing. */
int sn_stream(sn_ctx *ctx, void **hip_stream);
/* Launch configuration the sweep uses: tiles per sweep, tile-sweep kernel
 * launches per sweep, lanes per site, fine colours (kernel 3: 0 tiles, one
 * launch, the periodic colours). */
int sn_sweep_config(sn_ctx *ctx, int32_t *tiles, int32_t *launches_per_sweep,
                    int32_t *lanes_per_site, int32_t *fine_colours);
/* The sweep kernel the context runs (no counterpart in the reference, which
 * has one code path): kernel 1 field update, 2 per-phase gather, 3 whole
 * lattice; its threads per workgroup, lanes per site and pairs per lane (0
 * for kernel 1, whose lanes split a site's stencil, not its pairs).  Any
 * output may be NULL. */
int sn_kernel_choice(sn_ctx *ctx, int32_t *kernel, int32_t *block_threads, int32_t *lanes_per_site,
                     int32_t *pairs_per_lane);
/* Time the tile-sweep kernel alone: runs nsweeps sweeps with a HIP event pair
 * around every tile-sweep launch and returns the mean launch duration (ms)
 * and the number of launches timed. */
int sn_time_sweep_kernel(sn_ctx *ctx, int32_t nsweeps, double *mean_ms, int32_t *launches);

/* Autotuner (no counterpart in the reference): times candidate schedules
 * (tile edge targets, fine colouring, field-update or gather kernel) for the
 * given replicas on the current device, each as a real context over a random
 * lattice drawn with the species table (n_species lengths / prevalences, as
 * sn_init_lattice_random; n_species = 0: unit lengths), and returns the
 * fastest in *best (site-updates/s in *best_rate).  Mixed lengths change the
 * kernels' LDS use, so the table matters.  budget_ms bounds the search (0: try
 * every candidate).  With use_cache != 0 a previous result for the same
 * device, DipoleCutOff, replica sizes and uniform / mixed lengths is returned
 * from the cache file ($SN_TUNE_CACHE, else
 * $HOME/.cache/starrynight_amd/tune.txt) and new results are appended. */
int sn_autotune(const sn_params *params, int32_t n_replicas, int32_t n_species, const double *lengths,
                const double *prevalence, int32_t budget_ms, int32_t use_cache, sn_tuning *best,
                double *best_rate);

#ifdef __cplusplus
}
#endif
#endif
