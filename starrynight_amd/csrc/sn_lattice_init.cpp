// sn_lattice_init.cpp -- reference-compatible initial lattices (host only).
//
// Reproduces the state the reference builds before its first MC move:
// init_genrand(seed) (src/starrynight-main.c:172-176), the InitialLattice
// pattern (src/starrynight-lattice.c:25-137, chosen as in
// src/starrynight-main.c:184-203), then solid_solution
// (src/starrynight-lattice.c:139-171).  The lattice is held in float as the
// reference holds it, so the returned doubles are exactly its values.
#include <cmath>
#include <cstdint>

// every float operation rounds on its own, as in the reference's SSE build
#pragma STDC FP_CONTRACT OFF
#include <cstring>
#include <vector>

#include "../../include/starrynight_amd.h"

namespace {

// MT19937 (Matsumoto & Nishimura 1998): init_genrand + tempered output,
// the stream of src/mt19937ar-cok.c:63-197.
struct Mt {
    uint32_t s[624];
    int i = 624;
    explicit Mt(uint32_t seed)
    {
        s[0] = seed;
        for (int j = 1; j < 624; j++) s[j] = 1812433253u * (s[j - 1] ^ (s[j - 1] >> 30)) + (uint32_t)j;
    }
    uint32_t next()
    {
        if (i >= 624) {
            for (int k = 0; k < 624; k++) {
                uint32_t y = (s[k] & 0x80000000u) | (s[(k + 1) % 624] & 0x7fffffffu);
                s[k] = s[(k + 397) % 624] ^ (y >> 1) ^ ((s[(k + 1) % 624] & 1u) ? 0x9908b0dfu : 0u);
            }
            i = 0;
        }
        uint32_t y = s[i++];
        y ^= y >> 11;
        y ^= (y << 7) & 0x9d2c5680u;
        y ^= (y << 15) & 0xefc60000u;
        y ^= y >> 18;
        return y;
    }
    double real1() { return (double)next() * (1.0 / 4294967295.0); }
};

struct Dip {
    float x, y, z, len;
};

// random_sphere_point, DIM == 3 (src/starrynight-config.c:203-227): the
// float / double mix of the reference, operation by operation.
void sphere_point(Mt &g, Dip &p)
{
    float x1, x2;
    do {
        x1 = 2.0 * g.real1() - 1.0;
        x2 = 2.0 * g.real1() - 1.0;
    } while (x1 * x1 + x2 * x2 > 1.0);
    // C's sqrt takes a double: the float argument is widened, as in the reference
    p.x = 2 * x1 * std::sqrt((double)(1 - x1 * x1 - x2 * x2));
    p.y = 2 * x2 * std::sqrt((double)(1 - x1 * x1 - x2 * x2));
    p.z = 1.0 - 2.0 * (x1 * x1 + x2 * x2);
}

}  // namespace

extern "C" int sn_reference_initial_lattice(int32_t X, int32_t Y, int32_t Z, const char *initial_lattice,
                                            uint32_t seed, int32_t n_species, const double *lengths,
                                            const double *prevalence, double *xyzl)
{
    if (X < 1 || Y < 1 || Z < 1 || !xyzl) return SN_ERR_INVALID;
    if (n_species < 1 || n_species > 10 || !lengths || !prevalence) return SN_ERR_INVALID;
    const long long N = (long long)X * Y * Z;
    std::vector<Dip> lat(N, Dip{0.f, 0.f, 0.f, 0.f});
    auto at = [&](int x, int y, int z) -> Dip & { return lat[((long long)x * Y + y) * Z + z]; };
    Mt g(seed);
    const char *k = initial_lattice ? initial_lattice : "random";
    if (!std::strcmp(k, "ferroelectric")) {
        for (Dip &d : lat) { d.x = 1.0f; d.y = 0.0f; d.z = 0.0f; }
    } else if (!std::strcmp(k, "buckled")) {
        for (int x = 0; x < X; x++)
            for (int y = 0; y < Y; y++)
                for (int z = 0; z < Z; z++) { Dip &d = at(x, y, z); d.x = x % 2; d.y = y % 2; d.z = z % 2; }
    } else if (!std::strcmp(k, "antiferro_wall")) {
        for (int x = 0; x < X; x++)
            for (int y = 0; y < Y; y++)
                for (int z = 0; z < Z; z++) {
                    Dip &d = at(x, y, z);
                    if ((y < Y / 2) ^ (x > X / 2)) { d.x = (2. * ((z + y) % 2)) - 1.0; d.y = 0.0f; }
                    else { d.x = 0.0f; d.y = (2. * ((x + z) % 2)) - 1.0; }
                    d.z = 0.0f;
                }
    } else if (!std::strcmp(k, "ferro_wall")) {
        for (int x = 0; x < X; x++)
            for (int y = 0; y < Y; y++)
                for (int z = 0; z < Z; z++) {
                    Dip &d = at(x, y, z);
                    d.x = 0.0f; d.y = (x < X / 2) ? -1.0f : 1.0f; d.z = 0.0f;
                }
    } else if (!std::strcmp(k, "antiferro_slip")) {
        for (int x = 0; x < X; x++)
            for (int y = 0; y < Y; y++)
                for (int z = 0; z < Z; z++) {
                    Dip &d = at(x, y, z);
                    d.x = (x < X / 2) ? (2. * ((z + y) % 2)) - 1.0 : (2. * ((z + y + 1) % 2)) - 1.0;
                    d.y = 0.0f; d.z = 0.0f;
                }
    } else if (!std::strcmp(k, "spectrum")) {
        for (int x = 0; x < X; x++)
            for (int y = 0; y < Y; y++)
                for (int z = 0; z < Z; z++) {
                    float angle = 2 * M_PI * (x * X + y) / ((float)X * Y);
                    Dip &d = at(x, y, z);
                    d.x = std::sin((double)angle); d.y = std::cos((double)angle); d.z = 0.0f;
                }
    } else if (!std::strcmp(k, "slab_delete")) {
        // The reference zeroes the planes x < 6 and leaves every other site as
        // uninitialised memory; here the other sites are zero vectors too.
        for (Dip &d : lat) { d.x = 0.0f; d.y = 0.0f; d.z = 0.0f; }
    } else {
        // "random", and the reference's default for any other name
        for (Dip &d : lat) sphere_point(g, d);
    }
    // solid_solution: one float sample per site walked down the prevalence list
    std::vector<float> len(n_species), prev(n_species);
    for (int i = 0; i < n_species; i++) { len[i] = (float)lengths[i]; prev[i] = (float)prevalence[i]; }
    for (Dip &d : lat) {
        float sample = g.real1();
        int i = 0;
        for (; sample > prev[i]; sample -= prev[i], i++)
            if (i + 1 >= n_species) return SN_ERR_INVALID;   // the reference would read past Dipoles[]
        d.len = len[i];
    }
    for (long long s = 0; s < N; s++) {
        xyzl[4 * s] = lat[s].x; xyzl[4 * s + 1] = lat[s].y; xyzl[4 * s + 2] = lat[s].z;
        xyzl[4 * s + 3] = lat[s].len;
    }
    return SN_OK;
}
