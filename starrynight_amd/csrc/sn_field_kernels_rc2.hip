// sn_field_kernels_rc2.hip -- field-update tile kernel instances for DipoleCutOff = 2.
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

}  // namespace sn
