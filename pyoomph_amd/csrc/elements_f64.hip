// FP64 element kernels. Built with -ffp-contract=off so the operation order written in
// element_kernels.h is the executed order (bit-for-bit equal to the CPU restatement of
// the generated element code for +, -, *, /, sqrt).
#include "elements_impl.h"

namespace pyhip {
int launch_elements_f64(int kind, int form, const ElemArgs<double> &a, hipStream_t s) {
    return launch_elements_impl<double>(kind, form, a, s);
}
}  // namespace pyhip
