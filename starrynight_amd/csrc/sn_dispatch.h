// sn_dispatch.h -- launch-table lookups for the tile-sweep kernels, whose
// template instances are compiled in their own translation units
// (sn_gather_kernels.hip, sn_field_kernels_rc*.hip) so the build runs in parallel.
#pragma once
#include "sn_energy_tiled.h"
#include "sn_torus.h"
#include "sn_sweep_field.h"

namespace sn {

typedef void (*SweepFn)(SweepArgs, int, int);
typedef void (*FieldFn)(FieldArgs, int, int);

// 16 lanes share the scatter of one accepted site in the field-update kernel.
constexpr int kFieldG = 16;

// Staged-region size with compile-time-stride gather instances: the 14^3
// region of an 8^3 tile at DipoleCutOff 3 (10^3 at 2, 6^3 at 4, 12^3 at 1).
constexpr int kFixedRegion = 14 * 14 * 14;

// per-phase gather kernel for G lanes per site, NPL pairs per lane and blk
// threads per workgroup (or null); max_region == kFixedRegion with proposal
// chunks of kBlock selects the compile-time-stride instance where one is built
SweepFn pick_sweep(int G, int NPL, bool mixed, int blk, int max_region = 0, int prop_chunk = 0);
typedef void (*TorusFn)(TorusArgs, int);
// whole-lattice sweep kernel of small replicas (or null)
TorusFn pick_torus(int G, int NPL, bool mixed, int blk);
// field-update kernel for DipoleCutOff rc (1..5) or null
FieldFn pick_field(int rc, bool two_d, bool mixed);
FieldFn pick_field_rc1(bool two_d, bool mixed);
FieldFn pick_field_rc2(bool two_d, bool mixed);
FieldFn pick_field_rc3(bool two_d, bool mixed);
FieldFn pick_field_rc4(bool two_d, bool mixed);
FieldFn pick_field_rc5(bool two_d, bool mixed);

typedef void (*EnergyFn)(EnergyArgs);
// tiled energy / adjoint kernel for DipoleCutOff rc (1..5) or null
EnergyFn pick_energy(int rc, bool two_d, bool mixed);
EnergyFn pick_energy_rc1(bool two_d, bool mixed);
EnergyFn pick_energy_rc2(bool two_d, bool mixed);
EnergyFn pick_energy_rc3(bool two_d, bool mixed);
EnergyFn pick_energy_rc4(bool two_d, bool mixed);
EnergyFn pick_energy_rc5(bool two_d, bool mixed);

}  // namespace sn
