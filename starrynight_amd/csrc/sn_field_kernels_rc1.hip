// sn_field_kernels_rc1.hip -- field-update tile kernel instances for DipoleCutOff = 1.
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

}  // namespace sn
