// sn_field_kernels_rc1.hip -- field-update sweep and tiled energy kernel instances for
// DipoleCutOff = 1.
#include <hip/hip_runtime.h>

#include "sn_dispatch.h"

namespace sn {

FieldFn pick_field_rc1(bool two_d, bool mixed)
{
    if (two_d) return mixed ? (FieldFn)field_sweep_kernel<1, true, true, kFieldG>
                            : (FieldFn)field_sweep_kernel<1, true, false, kFieldG>;
    return mixed ? (FieldFn)field_sweep_kernel<1, false, true, kFieldG>
                 : (FieldFn)field_sweep_kernel<1, false, false, kFieldG>;
}

EnergyFn pick_energy_rc1(bool two_d, bool mixed)
{
    if (two_d) return mixed ? (EnergyFn)tile_energy_kernel<1, true, true>
                            : (EnergyFn)tile_energy_kernel<1, true, false>;
    return mixed ? (EnergyFn)tile_energy_kernel<1, false, true>
                 : (EnergyFn)tile_energy_kernel<1, false, false>;
}

}  // namespace sn
