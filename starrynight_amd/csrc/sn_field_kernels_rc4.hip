// sn_field_kernels_rc4.hip -- field-update sweep and tiled energy kernel instances for
// DipoleCutOff = 4.
#include <hip/hip_runtime.h>

#include "sn_dispatch.h"

namespace sn {

FieldFn pick_field_rc4(bool two_d, bool mixed)
{
    if (two_d) return mixed ? (FieldFn)field_sweep_kernel<4, true, true, kFieldG>
                            : (FieldFn)field_sweep_kernel<4, true, false, kFieldG>;
    return mixed ? (FieldFn)field_sweep_kernel<4, false, true, kFieldG>
                 : (FieldFn)field_sweep_kernel<4, false, false, kFieldG>;
}

EnergyFn pick_energy_rc4(bool two_d, bool mixed)
{
    if (two_d) return mixed ? (EnergyFn)tile_energy_kernel<4, true, true>
                            : (EnergyFn)tile_energy_kernel<4, true, false>;
    return mixed ? (EnergyFn)tile_energy_kernel<4, false, true>
                 : (EnergyFn)tile_energy_kernel<4, false, false>;
}

}  // namespace sn
