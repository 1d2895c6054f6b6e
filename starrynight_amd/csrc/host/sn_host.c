/*
 * sn_host.c -- the C driver: reads starrynight.cfg and runs the Metropolis
 * sweeps on the GPU through the C ABI.  It plays the role of the reference's
 * main() (src/starrynight-main.c:125-277): same config keys, same argv
 * overrides (argv[1] = T, argv[2] = CageStrain, :142-151), same seed
 * 0xDEADBEEF + T (:172), same equilibration / production loop with the
 * "MC Moves (per second): %f MHz" line (:219-259) and the final
 * ACCEPT / REJECT ratio (:273).  The analysis and image writers of the
 * reference are not part of this driver.
 *
 * Keys added for the GPU build (a libconfig reader ignores unknown keys):
 *   TemperatureSchedule = [ ... ]   one replica per temperature, run batched
 *   Replicas = n                    copies of each temperature (seeds differ)
 *   EfieldRamp = { axis = "z"; from = -0.1; to = 0.1; steps = 41; sweeps = 500; }
 *   Seed = n                        RNG key (default 0xDEADBEEF + T)
 *   Tile = [8, 8, 8]                tile edge the schedule aims for
 *   SaveLattice = "file"            final lattices, double[N][4] per replica (replica 0
 *                                   first), and "file.sweep": one line "replica
 *                                   sweep_index seed" per replica
 *   LoadLattice = "file"            start from a SaveLattice file instead of the
 *                                   initial pattern, each replica's draws resumed at
 *                                   its saved sweep index (an exact restart)
 *   Autotune = true                 pick the sweep schedule by measurement (sn_autotune)
 *
 * CalculateRadialOrderParameter writes the reference's RDF files
 * (src/starrynight-main.c:39, 58, 75-76): rows appended to "rdf.dat" after
 * initialisation and to "T_%04d_%d_%03d-RDF.dat" (T, (int) CageStrain,
 * megastep) after every megastep, in the format of
 * src/starrynight-analysis.c:582-594; with several replicas the file names
 * carry "_r<replica>" before the extension.
 *
 * usage: sn_host [-c config] [--dry-run lattice.bin] [T [CageStrain]]
 *
 * --dry-run reads the config and the argv overrides, prints the resolved
 * parameters of every replica (key=value lines on stdout), writes replica 0's
 * reference initial lattice (double[N][4]) to the named file and exits before
 * any device call.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "sn_config.h"
#include "starrynight_amd.h"

/* radial_order_parameter's file rows (src/starrynight-analysis.c:582-594):
 * header, then "r2 r FE AFE count T" for every occupied bin, then a blank line. */
static int append_rdf(sn_ctx *ctx, int replica, const char *name, int T)
{
    double fe[81], afe[81];
    int64_t cnt[81];
    int st = sn_radial_order(ctx, replica, fe, afe, cnt);
    if (st != SN_OK) return st;
    FILE *f = fopen(name, "a");
    if (!f) return SN_ERR_INVALID;
    fprintf(f, "# r^2 r orientational_FE_correlation[r^2] orientational_AFE_correlation[r^2] "
               "orientational_count[r^2] T\n");
    for (int i = 0; i < 81; i++)
        if (cnt[i] > 0)
            fprintf(f, "%d %f %f %f %lld %d\n", i, sqrt((double)i), fe[i], afe[i], (long long)cnt[i], T);
    fprintf(f, "\n");
    fclose(f);
    return SN_OK;
}

/* "x y z value" rows of lattice_potential_XYZ / lattice_Efield_XYZ
 * (src/starrynight-analysis.c:264-276, 466-479) */
static int write_xyz(const char *name, int X, int Y, int Z, const double *v)
{
    FILE *f = fopen(name, "w");
    if (!f) return SN_ERR_INVALID;
    for (int x = 0; x < X; x++)
        for (int y = 0; y < Y; y++)
            for (int z = 0; z < Z; z++) fprintf(f, "%d %d %d %f\n", x, y, z, v[((long long)x * Y + y) * Z + z]);
    fclose(f);
    return SN_OK;
}

/* lattice_potential_cube's Gaussian cube file (src/starrynight-analysis.c:280-308) */
static int write_cube(const char *name, int X, int Y, int Z, const double *v)
{
    FILE *f = fopen(name, "w");
    if (!f) return SN_ERR_INVALID;
    fprintf(f, "Starrynight Cube file: %d %d %d\n\n", X, Y, Z);
    fprintf(f, "1 0.0 0.0 0.0\n");
    fprintf(f, "%d 1.0 0.0 0.0\n", X);
    fprintf(f, "%d 0.0 1.0 0.0\n", Y);
    fprintf(f, "%d 0.0 0.0 1.0\n", Z);
    fprintf(f, "1 0.0 0.0 0.0 0.0 0.0\n");
    for (int x = 0; x < X; x++)
        for (int y = 0; y < Y; y++) {
            for (int z = 0; z < Z; z++) {
                fprintf(f, "%g ", v[((long long)x * Y + y) * Z + z]);
                if (z % 6 == 5) fprintf(f, "\n");
            }
            fprintf(f, "\n");
        }
    fclose(f);
    return SN_OK;
}

/* potential / field maps of one replica after a megastep (or equilibration):
 * prefix as the reference's analysis_midpoint (src/starrynight-main.c:85-92) */
static int write_maps(sn_ctx *ctx, int replica, int R, const char *prefix, int X, int Y, int Z, int potential,
                      int cube, int efield, double *buf)
{
    char base[160], name[256];
    int st;
    if (potential || cube) {
        if ((st = sn_dipole_potential(ctx, replica, buf))) return st;
        if (potential) {
            snprintf(base, sizeof(base), "%s_potential", prefix);
            if (R > 1) snprintf(name, sizeof(name), "%s_r%03d.xyz", base, replica);
            else snprintf(name, sizeof(name), "%s.xyz", base);
            if ((st = write_xyz(name, X, Y, Z, buf))) return st;
        }
        if (cube) {
            snprintf(base, sizeof(base), "%s_potential", prefix);
            if (R > 1) snprintf(name, sizeof(name), "%s_r%03d.cube", base, replica);
            else snprintf(name, sizeof(name), "%s.cube", base);
            if ((st = write_cube(name, X, Y, Z, buf))) return st;
        }
    }
    if (efield) {
        /* lattice_Efield_XYZ reports |E| with cutoff 4 (src/starrynight-analysis.c:477) */
        const long long N = (long long)X * Y * Z;
        double *E = (double *)malloc(sizeof(double) * 3 * N);
        if (!E) return SN_ERR_NOMEM;
        if ((st = sn_dipole_efield(ctx, replica, 4, E))) { free(E); return st; }
        for (long long i = 0; i < N; i++)
            buf[i] = sqrt(E[3 * i] * E[3 * i] + E[3 * i + 1] * E[3 * i + 1] + E[3 * i + 2] * E[3 * i + 2]);
        free(E);
        snprintf(base, sizeof(base), "%s_efield", prefix);
        if (R > 1) snprintf(name, sizeof(name), "%s_r%03d.xyz", base, replica);
        else snprintf(name, sizeof(name), "%s.xyz", base);
        if ((st = write_xyz(name, X, Y, Z, buf))) return st;
    }
    return SN_OK;
}

/* the reference's file name, with "_r<replica>" before the extension when
 * several replicas run */
static void rdf_name(char *out, size_t n, const char *base, const char *ext, int replica, int R)
{
    if (R > 1) snprintf(out, n, "%s_r%03d%s", base, replica, ext);
    else snprintf(out, n, "%s%s", base, ext);
}

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

#define CHECK(ctx, call)                                                                   \
    do {                                                                                   \
        int st_ = (call);                                                                  \
        if (st_ != SN_OK) {                                                                \
            fprintf(stderr, "%s failed: %s (%s)\n", #call, sn_status_string(st_),         \
                    (ctx) ? sn_last_error(ctx) : "");                                      \
            return 1;                                                                      \
        }                                                                                  \
    } while (0)

typedef struct {
    int T, X, Y, Z, rc, constrain_x, eqm, mega;
    double E[3], K, J, moves;
    int nspec;
    double len[10], prev[10];
    const char *initial;
} run_cfg;

/* load_config (src/starrynight-config.c:98-182): the reference's defaults
 * (:12-93), then whatever keys the file sets, with libconfig's typing. */
static int load(const char *path, snc_config *cf, run_cfg *r)
{
    r->T = 0; r->X = 20; r->Y = 20; r->Z = 20; r->rc = 3; r->constrain_x = 0;
    r->eqm = 10; r->mega = 400; r->moves = 1.0;
    r->E[0] = r->E[1] = r->E[2] = 0.0; r->K = 1.0; r->J = 1.0;
    r->nspec = 1; r->len[0] = 1.0; r->prev[0] = 1.0;
    r->initial = "random";
    if (snc_read_file(cf, path)) {
        fprintf(stderr, "%s:%d - %s\n", path, cf->error_line, cf->error);
        return -1;
    }
    double t;
    snc_lookup_int(cf, "T", &r->T);
    snc_lookup_int(cf, "X", &r->X);
    snc_lookup_int(cf, "Y", &r->Y);
    snc_lookup_int(cf, "Z", &r->Z);
    /* Efield is narrowed to float (src/starrynight-config.c:132-134) */
    if (snc_lookup_float(cf, "Efield.x", &t)) r->E[0] = (float)t;
    if (snc_lookup_float(cf, "Efield.y", &t)) r->E[1] = (float)t;
    if (snc_lookup_float(cf, "Efield.z", &t)) r->E[2] = (float)t;
    snc_lookup_float(cf, "K", &r->K);
    snc_lookup_float(cf, "CageStrain", &r->J);
    const snc_node *d = snc_lookup(cf, "Dipoles"), *p = snc_lookup(cf, "Prevalence");
    if (d && p) {
        int n = snc_length(p);
        if (n < 1 || n > 10 || snc_length(d) < n) {
            fprintf(stderr, "Dipoles / Prevalence must list 1..10 species\n");
            return -1;
        }
        r->nspec = n;
        for (int i = 0; i < n; i++) {
            r->len[i] = (float)snc_float_elem(d, i);
            r->prev[i] = (float)snc_float_elem(p, i);
        }
    }
    snc_lookup_bool(cf, "ConstrainToX", &r->constrain_x);
    snc_lookup_int(cf, "DipoleCutOff", &r->rc);
    snc_lookup_string(cf, "InitialLattice", &r->initial);
    snc_lookup_int(cf, "MCEqmSteps", &r->eqm);
    snc_lookup_int(cf, "MCMegaSteps", &r->mega);
    snc_lookup_float(cf, "MCMoves", &r->moves);
    return 0;
}

/* SaveLattice: every replica's lattice, then "<file>.sweep" with each
 * replica's sweep index and RNG key (the state a restart needs). */
static int checkpoint(sn_ctx *ctx, const char *path, int R, long long N, const sn_params *ps, double *lat)
{
    FILE *f = fopen(path, "wb");
    if (!f) { fprintf(stderr, "cannot write %s\n", path); return 1; }
    for (int i = 0; i < R; i++) {
        const int st = sn_get_lattice(ctx, i, lat);
        if (st != SN_OK || fwrite(lat, sizeof(double) * 4, N, f) != (size_t)N) {
            fprintf(stderr, "cannot write replica %d to %s (%s)\n", i, path, sn_status_string(st));
            fclose(f);
            return 1;
        }
    }
    fclose(f);
    char side[1024];
    snprintf(side, sizeof(side), "%s.sweep", path);
    FILE *g = fopen(side, "w");
    if (!g) { fprintf(stderr, "cannot write %s\n", side); return 1; }
    for (int i = 0; i < R; i++) {
        uint64_t idx = 0;
        const int st = sn_get_sweep_index(ctx, i, &idx);
        if (st != SN_OK) {
            fprintf(stderr, "sn_get_sweep_index(%d) failed: %s\n", i, sn_status_string(st));
            fclose(g);
            return 1;
        }
        fprintf(g, "%d %llu %llu\n", i, (unsigned long long)idx, (unsigned long long)ps[i].seed);
    }
    fclose(g);
    return 0;
}

/* LoadLattice: the lattices and sweep indices of a SaveLattice file.  The
 * replicas' RNG keys must match the saved ones, or the restart would not
 * continue the same draw stream. */
static int restore(sn_ctx *ctx, const char *path, int R, long long N, const sn_params *ps, double *lat)
{
    FILE *f = fopen(path, "rb");
    if (!f) { fprintf(stderr, "cannot read %s\n", path); return 1; }
    for (int i = 0; i < R; i++) {
        if (fread(lat, sizeof(double) * 4, N, f) != (size_t)N) {
            fprintf(stderr, "%s holds fewer than %d lattices of %lld sites\n", path, R, N);
            fclose(f);
            return 1;
        }
        const int st = sn_set_lattice(ctx, i, lat);
        if (st != SN_OK) {
            fprintf(stderr, "sn_set_lattice(%d) failed: %s (%s)\n", i, sn_status_string(st), sn_last_error(ctx));
            fclose(f);
            return 1;
        }
    }
    fclose(f);
    char side[1024];
    snprintf(side, sizeof(side), "%s.sweep", path);
    FILE *g = fopen(side, "r");
    if (!g) { fprintf(stderr, "cannot read %s\n", side); return 1; }
    for (int i = 0; i < R; i++) {
        int r = -1;
        unsigned long long idx = 0, seed = 0;
        if (fscanf(g, "%d %llu %llu", &r, &idx, &seed) != 3 || r != i) {
            fprintf(stderr, "%s: no sweep index for replica %d\n", side, i);
            fclose(g);
            return 1;
        }
        if (seed != (unsigned long long)ps[i].seed) {
            fprintf(stderr, "%s: replica %d was saved with RNG key %llu, this run uses %llu\n", side, i, seed,
                    (unsigned long long)ps[i].seed);
            fclose(g);
            return 1;
        }
        const int st = sn_set_sweep_index(ctx, i, idx);
        if (st != SN_OK) {
            fprintf(stderr, "sn_set_sweep_index(%d) failed: %s\n", i, sn_status_string(st));
            fclose(g);
            return 1;
        }
    }
    fclose(g);
    fprintf(stderr, "Restarted from %s\n", path);
    return 0;
}

int main(int argc, char **argv)
{
    const char *path = "starrynight.cfg", *dry_run = NULL;
    int ai = 1;
    while (argc > ai + 1 && (strcmp(argv[ai], "-c") == 0 || strcmp(argv[ai], "--dry-run") == 0)) {
        if (strcmp(argv[ai], "-c") == 0) path = argv[ai + 1];
        else dry_run = argv[ai + 1];
        ai += 2;
    }
    snc_config cf;
    run_cfg r;
    fprintf(stderr, "Starry Night (MI355X) - Monte Carlo brushstrokes.\n");
    if (load(path, &cf, &r)) return 1;
    if (argc > ai) { sscanf(argv[ai], "%d", &r.T); fprintf(stderr, "Command line temperature: T = %d\n", r.T); }
    if (argc > ai + 1) {
        sscanf(argv[ai + 1], "%lf", &r.J);
        fprintf(stderr, "Command Line CageStrain: CageStrain = %lf\n", r.J);
    }

    /* replicas: one per scheduled temperature (or the single T), times Replicas */
    int nT = 1, copies = 1;
    double Ts[1024];
    Ts[0] = r.T;
    const snc_node *sched = snc_lookup(&cf, "TemperatureSchedule");
    if (sched && snc_length(sched) > 0) {
        nT = snc_length(sched) < 1024 ? snc_length(sched) : 1024;
        for (int i = 0; i < nT; i++) {
            const snc_node *e = sched->kids[i];
            Ts[i] = e->type == SNC_INT ? (double)e->ival : e->fval;
        }
    }
    snc_lookup_int(&cf, "Replicas", &copies);
    if (copies < 1) copies = 1;
    int R = nT * copies;
    int64_t seed_base = (int64_t)(uint32_t)(0xDEADBEEFu + (uint32_t)r.T);
    int have_seed = snc_lookup_int64(&cf, "Seed", &seed_base);

    sn_params *ps = (sn_params *)calloc(R, sizeof(sn_params));
    if (!ps) return 1;
    for (int i = 0; i < R; i++) {
        double Ti = Ts[i % nT];
        ps[i].X = r.X; ps[i].Y = r.Y; ps[i].Z = r.Z; ps[i].dipole_cutoff = r.rc;
        ps[i].cage_strain = r.J;
        ps[i].efield[0] = r.E[0]; ps[i].efield[1] = r.E[1]; ps[i].efield[2] = r.E[2];
        ps[i].K = r.K; ps[i].constrain_to_x = r.constrain_x; ps[i].temperature = Ti;
        ps[i].seed = have_seed ? (uint64_t)seed_base + (uint64_t)i
                               : (uint64_t)(uint32_t)(0xDEADBEEFu + (uint32_t)(int)Ti) + ((uint64_t)(i / nT) << 32);
    }
    if (dry_run) {
        /* resolved parameters, then replica 0's initial lattice; no device call */
        printf("T=%d\nCageStrain=%.17g\nX=%d\nY=%d\nZ=%d\nDipoleCutOff=%d\nInitialLattice=%s\nreplicas=%d\n", r.T,
               r.J, r.X, r.Y, r.Z, r.rc, r.initial, R);
        for (int i = 0; i < R; i++)
            printf("replica=%d temperature=%.17g seed=%llu mt_seed=%u efield=%.9g,%.9g,%.9g\n", i, ps[i].temperature,
                   (unsigned long long)ps[i].seed, (unsigned)(0xDEADBEEFu + (uint32_t)(int)Ts[i % nT]),
                   ps[i].efield[0], ps[i].efield[1], ps[i].efield[2]);
        long long Nd = (long long)r.X * r.Y * r.Z;
        double *l0 = (double *)malloc(sizeof(double) * 4 * Nd);
        if (!l0) return 1;
        int st0 = sn_reference_initial_lattice(r.X, r.Y, r.Z, r.initial, (uint32_t)(0xDEADBEEFu + (uint32_t)(int)Ts[0]),
                                               r.nspec, r.len, r.prev, l0);
        if (st0 != SN_OK) { fprintf(stderr, "initial lattice: %s\n", sn_status_string(st0)); return 1; }
        FILE *f = fopen(dry_run, "wb");
        if (!f || fwrite(l0, sizeof(double) * 4, Nd, f) != (size_t)Nd) { fprintf(stderr, "cannot write %s\n", dry_run); return 1; }
        fclose(f);
        free(l0);
        free(ps);
        snc_destroy(&cf);
        return 0;
    }
    sn_tuning tun;
    memset(&tun, 0, sizeof(tun));
    const snc_node *tile = snc_lookup(&cf, "Tile");
    if (tile && snc_length(tile) == 3)
        for (int k = 0; k < 3; k++) tun.tile[k] = tile->kids[k]->type == SNC_INT ? (int)tile->kids[k]->ival : 0;
    int autotune = 0;
    snc_lookup_bool(&cf, "Autotune", &autotune);
    if (autotune) {
        double rate = 0.0;
        int ast = sn_autotune(ps, R, r.nspec, r.len, r.prev, 3000, 1, &tun, &rate);
        if (ast != SN_OK) { fprintf(stderr, "sn_autotune failed: %s\n", sn_status_string(ast)); return 1; }
        fprintf(stderr, "Autotune: tile %d %d %d, colouring %d, kernel %d, %d threads: %.1f MHz in trial\n",
                tun.tile[0], tun.tile[1], tun.tile[2], tun.fine_colouring, tun.kernel, tun.block_threads,
                1e-6 * rate);
    }
    int rdf = 0, want_pot = 0, want_cube = 0, want_efield = 0, want_recomb = 0;
    snc_lookup_bool(&cf, "CalculateRadialOrderParameter", &rdf);
    snc_lookup_bool(&cf, "CalculateRecombination", &want_recomb);
    snc_lookup_bool(&cf, "CalculatePotential", &want_pot);
    snc_lookup_bool(&cf, "SavePotentialCube", &want_cube);
    snc_lookup_bool(&cf, "CalculateEfield", &want_efield);

    fprintf(stderr, "Lattice %dx%dx%d, DipoleCutOff %d, CageStrain %f, Efield (%f %f %f), K %f, %d replica(s)\n",
            r.X, r.Y, r.Z, r.rc, r.J, r.E[0], r.E[1], r.E[2], r.K, R);
    sn_ctx *ctx = NULL;
    int st = sn_create(ps, R, &tun, &ctx);
    if (st != SN_OK) {
        fprintf(stderr, "sn_create failed: %s (%s)\n", sn_status_string(st), ctx ? sn_last_error(ctx) : "");
        return 1;
    }
    long long N = (long long)r.X * r.Y * r.Z;
    double *lat = (double *)malloc(sizeof(double) * 4 * N);
    if (!lat) return 1;
    for (int i = 0; i < R; i++) {
        /* Mersenne Twister seeded 0xDEADBEEF + T, as the reference (:172-176) */
        uint32_t mt_seed = (uint32_t)(0xDEADBEEFu + (uint32_t)(int)Ts[i % nT]);
        CHECK(ctx, sn_reference_initial_lattice(r.X, r.Y, r.Z, r.initial, mt_seed, r.nspec, r.len, r.prev, lat));
        CHECK(ctx, sn_set_lattice(ctx, i, lat));
    }
    const char *load = NULL;
    if (snc_lookup_string(&cf, "LoadLattice", &load) && restore(ctx, load, R, N, ps, lat)) return 1;
    /* the run log "Recombination_T_%04d.log" and its header (src/starrynight-main.c:164-178) */
    FILE **logs = (FILE **)calloc(R, sizeof(FILE *));
    if (!logs) return 1;
    for (int i = 0; i < R; i++) {
        char base[64], name[256];
        snprintf(base, sizeof(base), "Recombination_T_%04d", (int)Ts[i % nT]);
        rdf_name(name, sizeof(name), base, ".log", i, R);
        logs[i] = fopen(name, "w");
        if (!logs[i]) { fprintf(stderr, "cannot open %s\n", name); return 1; }
        fprintf(logs[i], "# Starrynight - simulation run on time(NULL)= %ld\n# Mersenne Twister Seed: %X\n",
                (long)time(NULL), (unsigned)(0xDEADBEEFu + (uint32_t)(int)Ts[i % nT]));
    }
    if (rdf)   /* analysis_initial (src/starrynight-main.c:39) */
        for (int i = 0; i < R; i++) {
            char name[256];
            rdf_name(name, sizeof(name), "rdf", ".dat", i, R);
            CHECK(ctx, append_rdf(ctx, i, name, (int)Ts[i % nT]));
        }
    int sweeps = (int)lround(r.moves);
    if (fabs(r.moves - sweeps) > 1e-9)
        fprintf(stderr, "MCMoves = %g: the GPU runs whole sweeps, using %d per megastep\n", r.moves, sweeps);
    fprintf(stderr, "'.' is %e MC moves attempted.\n", (double)N * sweeps);

    fprintf(stderr, "Equilibriation MC moves... %e\n", (double)N * sweeps * r.eqm);
    for (int i = 0; i < r.eqm; i++) CHECK(ctx, sn_mc_sweeps(ctx, sweeps));
    CHECK(ctx, sn_sync(ctx));
    if (want_efield)   /* "equilib_lattice_efield.xyz" (src/starrynight-main.c:225) */
        for (int i = 0; i < R; i++) CHECK(ctx, write_maps(ctx, i, R, "equilib_lattice", r.X, r.Y, r.Z, 0, 0, 1, lat));

    /* optional field ramp (hysteresis): `steps` field values from `from` to
     * `to`, `sweeps` sweeps at each; loop = true runs back from `to` to
     * `from` as well (the P(E) loop of figures/hysterisis/hysterisis.gpt).
     * Each step prints the mean over replicas of P = sum_i p_i / N along the
     * axis and its standard error over replicas: with sample = k > 0, P is
     * averaged over the step's second half, read every k sweeps; with
     * sample = 0, P is read once at the end of the step. */
    const snc_node *ramp = snc_lookup(&cf, "EfieldRamp");
    if (ramp) {
        const char *axis = "z";
        double from = -0.1, to = 0.1;
        int steps = 41, per = 500, loop = 0, sample = 0;
        snc_lookup_string(&cf, "EfieldRamp.axis", &axis);
        snc_lookup_float(&cf, "EfieldRamp.from", &from);
        snc_lookup_float(&cf, "EfieldRamp.to", &to);
        snc_lookup_int(&cf, "EfieldRamp.steps", &steps);
        snc_lookup_int(&cf, "EfieldRamp.sweeps", &per);
        snc_lookup_bool(&cf, "EfieldRamp.loop", &loop);
        snc_lookup_int(&cf, "EfieldRamp.sample", &sample);
        if (steps < 1 || per < 1 || sample < 0 || (sample > 0 && sample > per / 2)) {
            fprintf(stderr, "EfieldRamp: steps >= 1, sweeps >= 1 and 0 <= sample <= sweeps / 2\n");
            return 1;
        }
        int ax = axis[0] == 'x' ? 0 : axis[0] == 'y' ? 1 : 2;
        double *acc = (double *)calloc((size_t)R, sizeof(double));
        if (!acc) {
            fprintf(stderr, "EfieldRamp: out of host memory\n");
            return 1;
        }
        printf("# step branch E_%c mean_P_%c_over_replicas stderr_over_replicas\n", "xyz"[ax], "xyz"[ax]);
        for (int b = 0; b < (loop ? 2 : 1); b++) {
            for (int k = 0; k < steps; k++) {
                const int s = b == 0 ? k : steps - 1 - k;
                double Ev = steps > 1 ? from + (to - from) * s / (steps - 1) : from;
                for (int i = 0; i < R; i++) {
                    double E[3] = {r.E[0], r.E[1], r.E[2]};
                    E[ax] = Ev;
                    CHECK(ctx, sn_set_params(ctx, i, Ts[i % nT], E, r.J));
                }
                int nsamp = 0;
                for (int i = 0; i < R; i++) acc[i] = 0.0;
                if (sample > 0) {
                    const int first = per - (per / 2 / sample) * sample;   /* the second half, in reads of `sample` */
                    CHECK(ctx, sn_mc_sweeps(ctx, first));
                    for (int d = first; d < per; d += sample) {
                        CHECK(ctx, sn_mc_sweeps(ctx, sample));
                        for (int i = 0; i < R; i++) {
                            double M[3];
                            CHECK(ctx, sn_moment(ctx, i, M));
                            acc[i] += M[ax] / (double)N;
                        }
                        nsamp++;
                    }
                } else {
                    CHECK(ctx, sn_mc_sweeps(ctx, per));
                    for (int i = 0; i < R; i++) {
                        double M[3];
                        CHECK(ctx, sn_moment(ctx, i, M));
                        acc[i] = M[ax] / (double)N;
                    }
                    nsamp = 1;
                }
                double m = 0.0, v = 0.0;
                for (int i = 0; i < R; i++) m += acc[i] / nsamp;
                m /= R;
                for (int i = 0; i < R; i++) v += (acc[i] / nsamp - m) * (acc[i] / nsamp - m);
                const double se = R > 1 ? sqrt(v / (R - 1) / R) : 0.0;
                printf("%d %s %f %f %f\n", s, b == 0 ? "up" : "down", Ev, m, se);
                fflush(stdout);
            }
        }
        free(acc);
    }

    /* production: MCMegaSteps x MC_moves, timed around the moves only */
    printf("# megastep replica T energy_per_site |P| P_x\n");
    for (int m = 0; m < r.mega; m++) {
        double t0 = now_s();
        CHECK(ctx, sn_mc_sweeps(ctx, sweeps));
        CHECK(ctx, sn_sync(ctx));
        double t1 = now_s();
        for (int i = 0; i < R; i++) {
            double M[3], E;
            CHECK(ctx, sn_moment(ctx, i, M));
            CHECK(ctx, sn_energy(ctx, i, &E, NULL));
            printf("%d %d %g %f %f %f\n", m, i, Ts[i % nT], E / N,
                   sqrt(M[0] * M[0] + M[1] * M[1] + M[2] * M[2]) / N, M[0] / N);
            if (rdf) {   /* analysis_midpoint (src/starrynight-main.c:58, 75-76) */
                char base[128], name[256];
                snprintf(base, sizeof(base), "T_%04d_%d_%03d-RDF", (int)Ts[i % nT], (int)r.J, m);
                rdf_name(name, sizeof(name), base, ".dat", i, R);
                CHECK(ctx, append_rdf(ctx, i, name, (int)Ts[i % nT]));
            }
            if (want_recomb) {   /* recombination_calculator's log line (src/starrynight-analysis.c:127-162) */
                double rc_[8];
                CHECK(ctx, sn_recombination(ctx, i, rc_));
                fprintf(logs[i], "T: %d ZBe: %e ZBh: %e ZFDe: %e ZFDh: %e R_Boltz: %e ", (int)Ts[i % nT], rc_[0],
                        rc_[1], rc_[2], rc_[3], rc_[4]);
                fprintf(logs[i], "R_FD: %e FD-Total-electron: %e FD-Total-hole: %e\n", rc_[5], rc_[6], rc_[7]);
                fflush(logs[i]);
            }
            if (want_pot || want_cube || want_efield) {   /* src/starrynight-main.c:85-92 */
                char prefix[128];
                snprintf(prefix, sizeof(prefix), "T_%04d_%d_%03d", (int)Ts[i % nT], (int)r.J, m);
                CHECK(ctx, write_maps(ctx, i, R, prefix, r.X, r.Y, r.Z, want_pot, want_cube, want_efield, lat));
            }
        }
        fflush(stdout);
        fprintf(stderr, "MC Moves (per second): %f MHz\n", 1e-6 * (double)N * sweeps * R / (t1 - t0));
    }
    uint64_t A = 0, Rj = 0;
    for (int i = 0; i < R; i++) {
        uint64_t a, rj;
        CHECK(ctx, sn_counters(ctx, i, &a, &rj, NULL));
        A += a; Rj += rj;
    }
    fprintf(stderr, "Monte Carlo moves - ACCEPT: %llu REJECT: %llu ratio: %f\n", (unsigned long long)A,
            (unsigned long long)Rj, (float)A / (float)(A + Rj));
    const char *save = NULL;
    if (snc_lookup_string(&cf, "SaveLattice", &save) && checkpoint(ctx, save, R, N, ps, lat)) return 1;
    for (int i = 0; i < R; i++) fclose(logs[i]);
    free(logs);
    free(lat);
    free(ps);
    sn_destroy(ctx);
    snc_destroy(&cf);
    return 0;
}
