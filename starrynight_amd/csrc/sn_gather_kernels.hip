// sn_gather_kernels.hip -- instances of the per-phase gather tile kernel.
#include <hip/hip_runtime.h>

#include "sn_dispatch.h"

namespace sn {

SweepFn pick_sweep(int G, int NPL, bool mixed, int blk, int max_region, int prop_chunk)
{
#define SN_FIXED(g, n, b)                                                                            \
    if (max_region == kFixedRegion && prop_chunk == kBlock && G == g && NPL == n && blk == b)       \
        return mixed ? (SweepFn)tile_sweep_kernel<g, n, true, b, kFixedRegion>                        \
                     : (SweepFn)tile_sweep_kernel<g, n, false, b, kFixedRegion>;
    SN_FIXED(16, 4, 256) SN_FIXED(16, 2, 256) SN_FIXED(8, 8, 256)
#undef SN_FIXED
#define SN_CASE(g, n, b)                                                                        \
    if (G == g && NPL == n && blk == b) return mixed ? (SweepFn)tile_sweep_kernel<g, n, true, b>  \
                                                     : (SweepFn)tile_sweep_kernel<g, n, false, b>;
    SN_CASE(4, 4, 256) SN_CASE(4, 8, 256) SN_CASE(4, 16, 256) SN_CASE(8, 2, 256)
    SN_CASE(8, 4, 256) SN_CASE(8, 8, 256) SN_CASE(8, 16, 256) SN_CASE(16, 2, 256)
    SN_CASE(16, 4, 256) SN_CASE(16, 8, 256) SN_CASE(16, 16, 256) SN_CASE(32, 2, 256)
    SN_CASE(32, 4, 256) SN_CASE(32, 8, 256)
    SN_CASE(16, 4, 512) SN_CASE(16, 8, 512) SN_CASE(16, 16, 512) SN_CASE(32, 2, 512)
    SN_CASE(32, 4, 512) SN_CASE(32, 8, 512) SN_CASE(64, 2, 512) SN_CASE(64, 4, 512)
    // DipoleCutOff 1 (3 pairs, colour classes of 100+ sites) and 5 (257 pairs),
    // DipoleCutOff 2 at 512 threads (16 pairs, classes of up to 64 sites)
    SN_CASE(4, 2, 256) SN_CASE(4, 2, 512) SN_CASE(32, 16, 256) SN_CASE(64, 8, 512) SN_CASE(8, 2, 512)
    SN_CASE(4, 1, 256) SN_CASE(4, 1, 512) SN_CASE(64, 5, 256) SN_CASE(16, 1, 256)
#undef SN_CASE
    return nullptr;
}

TorusFn pick_torus(int G, int NPL, bool mixed, int blk)
{
#define SN_CASE(g, n, b)                                                                          \
    if (G == g && NPL == n && blk == b) return mixed ? (TorusFn)torus_sweep_kernel<g, n, true, b>   \
                                                     : (TorusFn)torus_sweep_kernel<g, n, false, b>;
    SN_CASE(4, 4, 256) SN_CASE(8, 4, 256) SN_CASE(16, 4, 256) SN_CASE(32, 2, 1024) SN_CASE(64, 2, 1024)
    SN_CASE(8, 2, 256) SN_CASE(16, 1, 256) SN_CASE(16, 1, 512) SN_CASE(16, 2, 512) SN_CASE(32, 1, 512)
    SN_CASE(16, 1, 1024) SN_CASE(8, 2, 512)
#undef SN_CASE
    return nullptr;
}

EnergyFn pick_energy(int rc, bool two_d, bool mixed)
{
    switch (rc) {
    case 1: return pick_energy_rc1(two_d, mixed);
    case 2: return pick_energy_rc2(two_d, mixed);
    case 3: return pick_energy_rc3(two_d, mixed);
    case 4: return pick_energy_rc4(two_d, mixed);
    case 5: return pick_energy_rc5(two_d, mixed);
    }
    return nullptr;
}

FieldFn pick_field(int rc, bool two_d, bool mixed)
{
    switch (rc) {
    case 1: return pick_field_rc1(two_d, mixed);
    case 2: return pick_field_rc2(two_d, mixed);
    case 3: return pick_field_rc3(two_d, mixed);
    case 4: return pick_field_rc4(two_d, mixed);
    case 5: return pick_field_rc5(two_d, mixed);
    }
    return nullptr;
}

}  // namespace sn
