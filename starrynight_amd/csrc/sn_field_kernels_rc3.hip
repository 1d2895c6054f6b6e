// sn_field_kernels_rc3.hip -- field-update sweep and tiled energy kernel instances for
// DipoleCutOff = 3.
#include <hip/hip_runtime.h>

#include "sn_dispatch.h"

namespace sn {

FieldFn pick_field_rc3(bool two_d, bool mixed)
{
    if (two_d) return mixed ? (FieldFn)field_sweep_kernel<3, true, true, kFieldG>
                            : (FieldFn)field_sweep_kernel<3, true, false, kFieldG>;
    return mixed ? (FieldFn)field_sweep_kernel<3, false, true, kFieldG>
                 : (FieldFn)field_sweep_kernel<3, false, false, kFieldG>;
}

EnergyFn pick_energy_rc3(bool two_d, bool mixed)
{
    if (two_d) return mixed ? (EnergyFn)tile_energy_kernel<3, true, true>
                            : (EnergyFn)tile_energy_kernel<3, true, false>;
    return mixed ? (EnergyFn)tile_energy_kernel<3, false, true>
                 : (EnergyFn)tile_energy_kernel<3, false, false>;
}

}  // namespace sn
