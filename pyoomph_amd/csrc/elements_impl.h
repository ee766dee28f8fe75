// Launch dispatch for the element kernels of one precision (included once per TU).
#pragma once
#include "element_kernels.h"

namespace pyhip {

template <typename T, int KIND, int FORM>
static int launch_scalar(const ElemArgs<T> &a, hipStream_t s) {
    constexpr int EPB = scalar_epb<KIND>();
    const int blocks = (a.e1 - a.e0 + EPB - 1) / EPB;
    if (blocks == 0) return PYHIP_OK;
    if (a.hvec)
        hipLaunchKernelGGL((scalar_element_kernel<T, KIND, FORM, false, true>), dim3(blocks), dim3(256), 0, s, a);
    else if (a.transient)
        hipLaunchKernelGGL((scalar_element_kernel<T, KIND, FORM, true, false>), dim3(blocks), dim3(256), 0, s, a);
    else
        hipLaunchKernelGGL((scalar_element_kernel<T, KIND, FORM, false, false>), dim3(blocks), dim3(256), 0, s, a);
    PYHIP_CHECK_HIP(hipGetLastError());
    return PYHIP_OK;
}

template <typename T, int KIND>
static int launch_kind(int form, const ElemArgs<T> &a, hipStream_t s) {
    switch (form) {
    case PYHIP_FORM_POISSON: return launch_scalar<T, KIND, PYHIP_FORM_POISSON>(a, s);
    case PYHIP_FORM_BRATU: return launch_scalar<T, KIND, PYHIP_FORM_BRATU>(a, s);
    case PYHIP_FORM_REACTION_DIFFUSION:
        return launch_scalar<T, KIND, PYHIP_FORM_REACTION_DIFFUSION>(a, s);
    }
    set_error("weak form not available on this element family");
    return PYHIP_ERR_INVALID;
}

template <typename T>
static int launch_elements_impl(int kind, int form, const ElemArgs<T> &a, hipStream_t s) {
    if (form == PYHIP_FORM_NAVIER_STOKES_TH) {
        if (kind != ELEM_QUAD_Q2) {
            set_error("Taylor-Hood Navier-Stokes needs Q2 quadrilaterals");
            return PYHIP_ERR_INVALID;
        }
        constexpr int EPB = 256 / 22;
        const int blocks = (a.e1 - a.e0 + EPB - 1) / EPB;
        if (blocks == 0) return PYHIP_OK;
        if (a.hvec) {
            // the product of the steady Jacobian (the time term is linear)
            hipLaunchKernelGGL((th_element_kernel<T, false, true>), dim3(blocks), dim3(256), 0, s, a);
        } else if (a.transient)
            hipLaunchKernelGGL((th_element_kernel<T, true>), dim3(blocks), dim3(256), 0, s, a);
        else
            hipLaunchKernelGGL((th_element_kernel<T, false>), dim3(blocks), dim3(256), 0, s, a);
        PYHIP_CHECK_HIP(hipGetLastError());
        return PYHIP_OK;
    }
    switch (kind) {
    case ELEM_TRI_P1: return launch_kind<T, ELEM_TRI_P1>(form, a, s);
    case ELEM_TRI_P2: return launch_kind<T, ELEM_TRI_P2>(form, a, s);
    case ELEM_QUAD_Q2: return launch_kind<T, ELEM_QUAD_Q2>(form, a, s);
    case ELEM_HEX_Q2: return launch_kind<T, ELEM_HEX_Q2>(form, a, s);
    }
    set_error("unknown element family");
    return PYHIP_ERR_INVALID;
}

}  // namespace pyhip
