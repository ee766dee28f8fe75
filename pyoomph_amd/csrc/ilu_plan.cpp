// Row order, permuted pattern and level sets of the ILU(0) preconditioner (see ilu_plan.h).
//
// Order within a sub-problem: the rows of group 0 (all rows, or Taylor-Hood velocities),
// then those of group 1 (Taylor-Hood pressures), which keeps ILU(0) pivots non-zero for
// the zero pressure-pressure block. Within a group either the natural order, or a
// multicolour order: a greedy colouring of the group's rows (in natural order, each row
// takes the smallest colour none of its already coloured neighbours in the group has),
// rows sorted by (colour, index). Rows of one colour of one group are independent, so the
// triangular solves have at most as many levels as there are colours. Levels: a row's
// lower level is one more than the largest level of the columns left of its diagonal,
// its upper level likewise from the right.
#include "ilu_plan.h"

#include <algorithm>
#include <utility>

namespace pyhip {

int ilu_plan_host(int64_t n, const int32_t *rs, const int32_t *col, const std::vector<int64_t> &sub_off,
                  const std::vector<char> &is_p, int order, int auto_levels, IluHostPlan &out, std::string &err) {
    out = IluHostPlan();
    const int nsub = (int)sub_off.size() - 1;
    const int64_t nnz = rs[n];
    if ((int64_t)is_p.size() != n || nsub < 0) {
        err = "ILU(0) plan: inconsistent arguments";
        return -1;
    }
    std::vector<int32_t> perm(n), iperm(n);
    // multicolour order for the sub-problems with mc[s] (colours are computed for all
    // rows; the neighbours of a row are in its own sub-problem)
    auto make_perm = [&](const std::vector<char> &mc) {
        std::vector<int32_t> color(n, 0);
        bool any = false;
        for (char c : mc) any = any || c;
        if (any) {
            std::vector<int32_t> mark;   // mark[c] == row: colour c taken by a neighbour
            for (int64_t i = 0; i < n; i++) {
                for (int32_t k = rs[i]; k < rs[i + 1]; k++) {
                    const int32_t c = col[k];
                    if (c < i && is_p[c] == is_p[i]) {
                        if ((size_t)color[c] >= mark.size()) mark.resize(color[c] + 1, -1);
                        mark[color[c]] = (int32_t)i;
                    }
                }
                int32_t c = 0;
                while ((size_t)c < mark.size() && mark[c] == i) c++;
                color[i] = c;
            }
        }
        int64_t w = 0;
        std::vector<int32_t> rows;
        for (int s = 0; s < nsub; s++)
            for (char g = 0; g < 2; g++) {
                rows.clear();
                for (int64_t i = sub_off[s]; i < sub_off[s + 1]; i++)
                    if (is_p[i] == g) rows.push_back((int32_t)i);
                if (mc[s])
                    std::stable_sort(rows.begin(), rows.end(),
                                     [&](int32_t x, int32_t y) { return color[x] < color[y]; });
                for (int32_t r : rows) perm[w++] = r;
            }
        for (int64_t i = 0; i < n; i++) iperm[perm[i]] = (int32_t)i;
    };
    std::vector<int32_t> prs(n + 1, 0), pcol(nnz), psrc(nnz), pdiag(n, -1), lev(n, 0);
    auto make_pattern = [&]() -> int {
        std::vector<std::pair<int32_t, int32_t>> tmp;
        for (int64_t i = 0; i < n; i++) {
            const int32_t o = perm[i];
            tmp.clear();
            for (int32_t k = rs[o]; k < rs[o + 1]; k++) tmp.push_back({iperm[col[k]], k});
            std::sort(tmp.begin(), tmp.end());
            prs[i + 1] = prs[i] + (int32_t)tmp.size();
            pdiag[i] = -1;
            for (size_t t = 0; t < tmp.size(); t++) {
                pcol[prs[i] + t] = tmp[t].first;
                psrc[prs[i] + t] = tmp[t].second;
                if (tmp[t].first == i) pdiag[i] = prs[i] + (int32_t)t;
            }
            if (pdiag[i] < 0) {
                err = "ILU(0): a row has no structural diagonal entry";
                return -1;
            }
        }
        return 0;
    };
    // lower levels (depends on smaller columns)
    auto lower_levels = [&]() {
        int32_t nl = 0;
        for (int64_t i = 0; i < n; i++) {
            int32_t l = 0;
            for (int32_t k = prs[i]; k < pdiag[i]; k++) l = std::max(l, lev[pcol[k]] + 1);
            lev[i] = l;
            nl = std::max(nl, l + 1);
        }
        return nl;
    };
    int rc0;
    std::vector<char> mc(nsub, order == 2 ? 1 : 0);
    make_perm(mc);
    if ((rc0 = make_pattern())) return rc0;
    int32_t nlo = lower_levels();
    if (order == 0 && nlo > auto_levels) {
        // per sub-problem, so that its order does not depend on the batch
        bool any = false;
        for (int s = 0; s < nsub; s++) {
            int32_t nl = 0;
            for (int64_t i = sub_off[s]; i < sub_off[s + 1]; i++) nl = std::max(nl, lev[i] + 1);
            mc[s] = nl > auto_levels;
            any = any || mc[s];
        }
        if (any) {
            make_perm(mc);
            if ((rc0 = make_pattern())) return rc0;
            nlo = lower_levels();
        }
    }
    auto bucket = [&](int32_t nl, std::vector<int32_t> &rows, std::vector<int64_t> &ptr) {
        ptr.assign(nl + 1, 0);
        for (int64_t i = 0; i < n; i++) ptr[lev[i] + 1]++;
        for (int32_t l = 0; l < nl; l++) ptr[l + 1] += ptr[l];
        rows.resize(n);
        std::vector<int64_t> f(ptr.begin(), ptr.end() - 1);
        for (int64_t i = 0; i < n; i++) rows[f[lev[i]]++] = (int32_t)i;
    };
    bucket(nlo, out.lo_rows, out.lo_ptr);
    // upper levels (depends on larger columns), computed in decreasing row order
    int32_t nup = 0;
    for (int64_t i = n - 1; i >= 0; i--) {
        int32_t l = 0;
        for (int32_t k = pdiag[i] + 1; k < prs[i + 1]; k++) l = std::max(l, lev[pcol[k]] + 1);
        lev[i] = l;
        nup = std::max(nup, l + 1);
    }
    bucket(nup, out.up_rows, out.up_ptr);
    bool mc_any = false;
    for (char c : mc) mc_any = mc_any || c;
    out.perm = std::move(perm);
    out.prs = std::move(prs);
    out.pcol = std::move(pcol);
    out.psrc = std::move(psrc);
    out.pdiag = std::move(pdiag);
    out.multicolor = mc_any;
    return 0;
}

}  // namespace pyhip
