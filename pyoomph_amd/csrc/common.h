// Shared host/device definitions of the pyoomph HIP library.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "../../include/pyoomph_hip.h"
#include "element_common.h"

namespace pyhip {

// Frees device memory on clean-up paths (errors there are not actionable).
inline void dfree(void *p) {
    if (p) (void)hipFree(p);
}

// Thread-local last error, reported through pyhip_last_error().
void set_error(const std::string &msg);

#define PYHIP_CHECK_HIP(expr)                                                         \
    do {                                                                              \
        hipError_t _e = (expr);                                                       \
        if (_e != hipSuccess) {                                                       \
            ::pyhip::set_error(std::string(#expr) + ": " + hipGetErrorString(_e));    \
            return PYHIP_ERR_HIP;                                                     \
        }                                                                             \
    } while (0)

#define PYHIP_REQUIRE(cond, msg)                                                      \
    do {                                                                              \
        if (!(cond)) {                                                                \
            ::pyhip::set_error(msg);                                                  \
            return PYHIP_ERR_INVALID;                                                 \
        }                                                                             \
    } while (0)

// Element families on the hot path (a row of the reference's BulkElement* classes):
// ElemKind in element_common.h, with the values of PYHIP_ELEM_*.
static_assert(ELEM_TRI_P1 == PYHIP_ELEM_TRI_P1 && ELEM_TRI_P2 == PYHIP_ELEM_TRI_P2 &&
                  ELEM_QUAD_Q2 == PYHIP_ELEM_QUAD_Q2 && ELEM_HEX_Q2 == PYHIP_ELEM_HEX_Q2,
              "element family ids");

// Reference-element tables at the integration points of the default scheme.
struct RefTables {
    int dim, nnode, nqp, nnode1;
    double w[27];
    double psi[27 * 27];
    double dpsi[27 * 27 * 3];
    double psi1[27 * 4];
};

int fill_ref_tables(int kind, RefTables *t);

inline int elem_nnode(int kind) {
    switch (kind) {
    case ELEM_TRI_P1: return 3;
    case ELEM_TRI_P2: return 6;
    case ELEM_QUAD_Q2: return 9;
    case ELEM_HEX_Q2: return 27;
    }
    return -1;
}
inline int elem_dim(int kind) { return kind == ELEM_HEX_Q2 ? 3 : 2; }

// Device allocation counted in `counter` (bytes); n == 0 allocates one element.
template <typename T>
int dev_alloc(T **p, size_t n, int64_t &counter) {
    *p = nullptr;
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e != hipSuccess) {
        set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
        *p = nullptr;
        return PYHIP_ERR_NOMEM;
    }
    counter += (int64_t)(n * sizeof(T));
    return PYHIP_OK;
}

}  // namespace pyhip
