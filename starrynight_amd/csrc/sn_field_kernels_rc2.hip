// sn_field_kernels_rc2.hip -- field-update sweep and tiled energy kernel instances for
// DipoleCutOff = 2.
#include <hip/hip_runtime.h>

#include "sn_dispatch.h"

namespace sn {

FieldFn pick_field_rc2(bool two_d, bool mixed)
{
    if (two_d) return mixed ? (FieldFn)field_sweep_kernel<2, true, true, kFieldG>
                            : (FieldFn)field_sweep_kernel<2, true, false, kFieldG>;
    return mixed ? (FieldFn)field_sweep_kernel<2, false, true, kFieldG>
                 : (FieldFn)field_sweep_kernel<2, false, false, kFieldG>;
}

EnergyFn pick_energy_rc2(bool two_d, bool mixed)
{
    if (two_d) return mixed ? (EnergyFn)tile_energy_kernel<2, true, true>
                            : (EnergyFn)tile_energy_kernel<2, true, false>;
    return mixed ? (EnergyFn)tile_energy_kernel<2, false, true>
                 : (EnergyFn)tile_energy_kernel<2, false, false>;
}

}  // namespace sn
