// FP32 element kernels (mixed-precision Jacobians for the FP32 Krylov solve).
#include "elements_impl.h"

namespace pyhip {
int launch_elements_f32(int kind, int form, const ElemArgs<float> &a, hipStream_t s) {
    return launch_elements_impl<float>(kind, form, a, s);
}
}  // namespace pyhip
