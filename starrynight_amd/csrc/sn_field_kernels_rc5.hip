// sn_field_kernels_rc5.hip -- field-update sweep and tiled energy kernel instances for
// DipoleCutOff = 5.
#include <hip/hip_runtime.h>

#include "sn_dispatch.h"

namespace sn {

FieldFn pick_field_rc5(bool two_d, bool mixed)
{
    if (two_d) return mixed ? (FieldFn)field_sweep_kernel<5, true, true, kFieldG>
                            : (FieldFn)field_sweep_kernel<5, true, false, kFieldG>;
    return mixed ? (FieldFn)field_sweep_kernel<5, false, true, kFieldG>
                 : (FieldFn)field_sweep_kernel<5, false, false, kFieldG>;
}

EnergyFn pick_energy_rc5(bool two_d, bool mixed)
{
    if (two_d) return mixed ? (EnergyFn)tile_energy_kernel<5, true, true>
                            : (EnergyFn)tile_energy_kernel<5, true, false>;
    return mixed ? (EnergyFn)tile_energy_kernel<5, false, true>
                 : (EnergyFn)tile_energy_kernel<5, false, false>;
}

}  // namespace sn
