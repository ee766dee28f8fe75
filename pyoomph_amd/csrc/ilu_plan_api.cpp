// Host entry points to the ILU(0) plan builder for tests on the CPU (no device work):
// they expose the row order, the permuted pattern and the level sets.
#include <cstring>
#include <string>

#include "ilu_plan.h"

using namespace pyhip;

namespace {
thread_local std::string g_ilu_err;
}

extern "C" {

const char *pyhip_ilu_plan_error(void) { return g_ilu_err.c_str(); }

// is_p may be null (one group per sub-problem); order as ilu_plan_host
void *pyhip_ilu_plan_new(int64_t n, const int32_t *rs, const int32_t *col, int nsub, const int64_t *sub_off,
                         const int8_t *is_p, int order, int auto_levels) {
    if (n < 0 || !rs || !col || nsub < 1 || !sub_off || sub_off[0] != 0 || sub_off[nsub] != n || order < 0 ||
        order > 2) {
        g_ilu_err = "invalid arguments";
        return nullptr;
    }
    std::vector<int64_t> off(sub_off, sub_off + nsub + 1);
    std::vector<char> g(n, 0);
    if (is_p)
        for (int64_t i = 0; i < n; i++) g[i] = is_p[i] ? 1 : 0;
    IluHostPlan *P = new IluHostPlan();
    if (ilu_plan_host(n, rs, col, off, g, order, auto_levels, *P, g_ilu_err)) {
        delete P;
        return nullptr;
    }
    return P;
}

// sizes[0..3] = lower levels, upper levels, nnz, multicolour flag
int pyhip_ilu_plan_sizes(const void *h, int64_t *sizes) {
    const IluHostPlan *P = (const IluHostPlan *)h;
    if (!P || !sizes) return -1;
    sizes[0] = (int64_t)P->lo_ptr.size() - 1;
    sizes[1] = (int64_t)P->up_ptr.size() - 1;
    sizes[2] = (int64_t)P->pcol.size();
    sizes[3] = P->multicolor ? 1 : 0;
    return 0;
}

int pyhip_ilu_plan_copy(const void *h, int32_t *perm, int32_t *prs, int32_t *pcol, int32_t *psrc, int32_t *pdiag,
                        int32_t *lo_rows, int64_t *lo_ptr, int32_t *up_rows, int64_t *up_ptr) {
    const IluHostPlan *P = (const IluHostPlan *)h;
    if (!P) return -1;
    auto cp = [](auto *dst, const auto &v) {
        if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(v[0]));
    };
    cp(perm, P->perm);
    cp(prs, P->prs);
    cp(pcol, P->pcol);
    cp(psrc, P->psrc);
    cp(pdiag, P->pdiag);
    cp(lo_rows, P->lo_rows);
    cp(lo_ptr, P->lo_ptr);
    cp(up_rows, P->up_rows);
    cp(up_ptr, P->up_ptr);
    return 0;
}

void pyhip_ilu_plan_free(void *h) { delete (IluHostPlan *)h; }

}  // extern "C"
