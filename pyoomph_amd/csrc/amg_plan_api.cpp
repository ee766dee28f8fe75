// Host entry points to the multigrid plan builder for tests on the CPU (no device work):
// they expose the plan's patterns level by level.
#include <cstring>
#include <string>

#include "amg_plan.h"

using namespace pyhip;

namespace {
thread_local std::string g_plan_err;
}

extern "C" {

const char *pyhip_amg_plan_error(void) { return g_plan_err.c_str(); }

void *pyhip_amg_plan_new(int64_t n, const int32_t *rs, const int32_t *col, int nsub, const int64_t *sub_off,
                         int cap, int bottom_rows, int max_levels) {
    if (n < 0 || !rs || !col || nsub < 1 || !sub_off) {
        g_plan_err = "invalid arguments";
        return nullptr;
    }
    AmgPlan *P = new AmgPlan();
    std::vector<int64_t> off(sub_off, sub_off + nsub + 1);
    if (amg_plan_build(n, rs, col, off, cap, bottom_rows, max_levels, *P, g_plan_err)) {
        delete P;
        return nullptr;
    }
    return P;
}

// the same with the element level: on[nsub], nc[nsub], rows frs[n+1] / fcol / fw over the
// level-0 rows (AmgFixedInterp)
void *pyhip_amg_plan_new_fixed(int64_t n, const int32_t *rs, const int32_t *col, int nsub, const int64_t *sub_off,
                               int cap, int bottom_rows, int max_levels, const int32_t *on, const int64_t *nc,
                               const int32_t *frs, const int32_t *fcol, const double *fw) {
    if (n < 0 || !rs || !col || nsub < 1 || !sub_off || !on || !nc || !frs || (frs[n] > 0 && (!fcol || !fw))) {
        g_plan_err = "invalid arguments";
        return nullptr;
    }
    AmgFixedInterp F;
    F.on.assign(on, on + nsub);
    F.nc.assign(nc, nc + nsub);
    F.rs.assign(frs, frs + n + 1);
    F.col.assign(fcol, fcol + frs[n]);
    F.w.assign(fw, fw + frs[n]);
    AmgPlan *P = new AmgPlan();
    std::vector<int64_t> off(sub_off, sub_off + nsub + 1);
    if (amg_plan_build(n, rs, col, off, cap, bottom_rows, max_levels, *P, g_plan_err, &F)) {
        delete P;
        return nullptr;
    }
    return P;
}

// values of level l's P entries before the smoothing [nnz(P)]
int pyhip_amg_plan_level_pfix(const void *h, int l, double *pfix) {
    const AmgPlan *P = (const AmgPlan *)h;
    if (!P || l < 0 || l >= (int)P->lv.size() || !pfix) return -1;
    const AmgLevelPlan &L = P->lv[l];
    if (!L.pfix.empty()) std::memcpy(pfix, L.pfix.data(), L.pfix.size() * sizeof(double));
    return 0;
}

int pyhip_amg_plan_levels(const void *h) { return h ? (int)((const AmgPlan *)h)->lv.size() : -1; }

// sizes[0..5] = n, nc, nnz(A_l), nnz(P), nnz(Q), contributions of P
int pyhip_amg_plan_level_sizes(const void *h, int l, int64_t *sizes) {
    const AmgPlan *P = (const AmgPlan *)h;
    if (!P || l < 0 || l >= (int)P->lv.size() || !sizes) return -1;
    const AmgLevelPlan &L = P->lv[l];
    sizes[0] = L.n;
    sizes[1] = L.nc;
    sizes[2] = L.rs.back();
    sizes[3] = L.prs.back();
    sizes[4] = L.qrs.back();
    sizes[5] = L.pcptr.back();
    return 0;
}

int pyhip_amg_plan_level_copy(const void *h, int l, int32_t *rs, int32_t *col, int32_t *agg, int32_t *prs,
                              int32_t *pcol, int32_t *pcptr, int32_t *pcidx, int32_t *trs, int32_t *trow,
                              int32_t *tsrc, int32_t *qrs, int32_t *qcol) {
    const AmgPlan *P = (const AmgPlan *)h;
    if (!P || l < 0 || l >= (int)P->lv.size()) return -1;
    const AmgLevelPlan &L = P->lv[l];
    auto cp = [](int32_t *dst, const std::vector<int32_t> &v) {
        if (dst && !v.empty()) std::memcpy(dst, v.data(), v.size() * sizeof(int32_t));
    };
    cp(rs, L.rs);
    cp(col, L.col);
    cp(agg, L.agg);
    cp(prs, L.prs);
    cp(pcol, L.pcol);
    cp(pcptr, L.pcptr);
    cp(pcidx, L.pcidx);
    cp(trs, L.trs);
    cp(trow, L.trow);
    cp(tsrc, L.tsrc);
    cp(qrs, L.qrs);
    cp(qcol, L.qcol);
    return 0;
}

// coarsest level: sizes[0] = rows, sizes[1] = nnz, sizes[2] = largest sub-problem
int pyhip_amg_plan_bottom(const void *h, int64_t *sizes, int32_t *rs, int32_t *col, int64_t *sub_off) {
    const AmgPlan *P = (const AmgPlan *)h;
    if (!P || !sizes) return -1;
    sizes[0] = P->nbot;
    sizes[1] = P->bot_rs.back();
    sizes[2] = P->max_bot_rows;
    if (rs) std::memcpy(rs, P->bot_rs.data(), P->bot_rs.size() * sizeof(int32_t));
    if (col && !P->bot_col.empty()) std::memcpy(col, P->bot_col.data(), P->bot_col.size() * sizeof(int32_t));
    if (sub_off) std::memcpy(sub_off, P->bot_off.data(), P->bot_off.size() * sizeof(int64_t));
    return 0;
}

void pyhip_amg_plan_free(void *h) { delete (AmgPlan *)h; }

}  // extern "C"
