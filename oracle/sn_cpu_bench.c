/*
 * sn_cpu_bench.c -- timed CPU baseline (test/bench infrastructure only).
 *
 * One single-threaded process running the float-faithful restatement of the
 * reference hot path exactly as src/starrynight-main.c drives it: MT19937
 * seeded 0xDEADBEEF + T (:172-176), initialise_lattice_random + solid_solution
 * with Dipoles = [1.0] (:203-205), beta = 1/((float)T/300.0) (:215), then
 * MC_moves in the timed region only (:247-249).
 *
 * usage: sn_cpu_bench X Y Z rc T moves [f32|f64]
 * prints one JSON line: {"updates": n, "seconds": t, "rate": n/t, ...}
 */
#include "sn_oracle.h"
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

static double now(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

int main(int argc, char **argv)
{
    if (argc < 7) {
        fprintf(stderr, "usage: %s X Y Z rc T moves [f32|f64]\n", argv[0]);
        return 2;
    }
    int X = atoi(argv[1]), Y = atoi(argv[2]), Z = atoi(argv[3]), rc = atoi(argv[4]);
    int T = atoi(argv[5]);
    long long moves = atoll(argv[6]);
    int f64 = argc > 7 && strcmp(argv[7], "f64") == 0;
    or_neighbour nb[10000];
    int nnb = or_gen_neighbour(rc, Z, nb, 10000);
    if (nnb < 0 || X < 1 || Y < 1 || Z < 1) return 2;
    or_model m = {X, Y, Z, rc, 1.0, {0, 0, 0}, 0.0, 0, 1.0 / ((float)T / 300.0)};
    or_mt19937 g;
    or_mt_seed(&g, (uint32_t)(0xDEADBEEFu + (uint32_t)T));
    long long N = (long long)X * Y * Z;
    float *lat = (float *)malloc(sizeof(float) * 4 * N);
    if (!lat) return 3;
    or_init_lattice_random_f32(&g, X, Y, Z, lat);
    float len[1] = {1.0f}, prev[1] = {1.0f};
    or_solid_solution_f32(&g, X, Y, Z, 1, len, prev, lat);
    uint64_t acc = 0, rej = 0;
    double t0, t1;
    if (!f64) {
        t0 = now();
        or_mc_moves_f32(&m, nb, nnb, lat, &g, moves, &acc, &rej, NULL);
        t1 = now();
    } else {
        double *ld = (double *)malloc(sizeof(double) * 4 * N);
        if (!ld) return 3;
        for (long long i = 0; i < 4 * N; i++) ld[i] = lat[i];
        t0 = now();
        or_mc_moves_f64(&m, nb, nnb, ld, &g, moves, &acc, &rej);
        t1 = now();
        free(ld);
    }
    printf("{\"updates\": %lld, \"seconds\": %.6f, \"rate\": %.6e, \"accept\": %llu, "
           "\"reject\": %llu, \"mode\": \"%s\"}\n",
           moves, t1 - t0, moves / (t1 - t0), (unsigned long long)acc, (unsigned long long)rej,
           f64 ? "f64" : "f32");
    free(lat);
    return 0;
}
