// sn_field_kernels_rc5.hip -- field-update tile kernel instances for DipoleCutOff = 5.
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

}  // namespace sn
