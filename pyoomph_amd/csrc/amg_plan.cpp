// Symbolic set-up of the smoothed-aggregation multigrid preconditioner (see amg_plan.h).
#include "amg_plan.h"

#include <algorithm>
#include <thread>

#include "../../include/pyoomph_hip.h"

namespace pyhip {

namespace {

template <typename F>
void parallel_for(int64_t n, F &&fn) {
    unsigned nt = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    if (n < 256) nt = 1;   // coarse levels have few rows but long ones: thread from 256 rows
    const int64_t chunk = (n + nt - 1) / nt;
    std::vector<std::thread> th;
    for (unsigned t = 0; t < nt; t++) {
        const int64_t b = (int64_t)t * chunk, e = std::min<int64_t>(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([&fn, b, e]() { fn(b, e); });
    }
    for (auto &x : th) x.join();
}

// Sorted distinct values of buf (all in [0, ncols)), using a per-thread marker array
// (mark[x] == stamp: x already taken), so long product lists are not sorted whole.
static void distinct_sorted(std::vector<int32_t> &buf, std::vector<int64_t> &mark, int64_t stamp) {
    size_t w = 0;
    for (size_t t = 0; t < buf.size(); t++) {
        const int32_t x = buf[t];
        if (mark[x] != stamp) {
            mark[x] = stamp;
            buf[w++] = x;
        }
    }
    buf.resize(w);
    std::sort(buf.begin(), buf.end());
}

// CSR whose row r holds the sorted distinct values in [0, ncols) produced by gen(r, push)
template <typename G>
bool build_rows(int64_t n, int64_t ncols, G &&gen, std::vector<int32_t> &rs, std::vector<int32_t> &col,
                std::string &err) {
    std::vector<int64_t> len(n + 1, 0);
    parallel_for(n, [&](int64_t b, int64_t e) {
        std::vector<int32_t> buf;
        std::vector<int64_t> mark(std::max<int64_t>(1, ncols), -1);
        for (int64_t r = b; r < e; r++) {
            buf.clear();
            gen(r, buf);
            distinct_sorted(buf, mark, r);
            len[r + 1] = (int64_t)buf.size();
        }
    });
    rs.assign(n + 1, 0);
    int64_t acc = 0;
    for (int64_t r = 0; r < n; r++) {
        acc += len[r + 1];
        if (acc > INT32_MAX) {
            err = "multigrid pattern exceeds int32 range";
            return false;
        }
        rs[r + 1] = (int32_t)acc;
    }
    col.resize(acc);
    parallel_for(n, [&](int64_t b, int64_t e) {
        std::vector<int32_t> buf;
        std::vector<int64_t> mark(std::max<int64_t>(1, ncols), -1);
        for (int64_t r = b; r < e; r++) {
            buf.clear();
            gen(r, buf);
            distinct_sorted(buf, mark, r);
            std::copy(buf.begin(), buf.end(), col.begin() + rs[r]);
        }
    });
    return true;
}

bool diag_positions(int64_t n, const std::vector<int32_t> &rs, const std::vector<int32_t> &col,
                    std::vector<int32_t> &diag, std::string &err) {
    diag.assign(n, -1);
    bool ok = true;
    parallel_for(n, [&](int64_t b, int64_t e) {
        for (int64_t r = b; r < e; r++) {
            const int32_t *c0 = col.data() + rs[r], *c1 = col.data() + rs[r + 1];
            const int32_t *d = std::lower_bound(c0, c1, (int32_t)r);
            if (d == c1 || *d != r) ok = false;
            else diag[r] = (int32_t)(d - col.data());
        }
    });
    if (!ok) err = "multigrid: a row has no diagonal entry";
    return ok;
}

}  // namespace

int amg_plan_build(int64_t n0, const int32_t *rs0, const int32_t *col0, const std::vector<int64_t> &sub_off0,
                   int cap, int bottom_rows, int max_levels, AmgPlan &plan, std::string &err,
                   const AmgFixedInterp *fix) {
    plan = AmgPlan();
    const int nsub = (int)sub_off0.size() - 1;
    std::vector<int32_t> rs(rs0, rs0 + n0 + 1), col(col0, col0 + rs0[n0]);
    std::vector<int64_t> sub_off = sub_off0;
    int64_t n = n0;
    auto max_sub = [&](const std::vector<int64_t> &off) {
        int64_t m = 0;
        for (int s = 0; s < nsub; s++) m = std::max(m, off[s + 1] - off[s]);
        return m;
    };
    while ((int)plan.lv.size() < max_levels) {
        std::vector<int32_t> cont(nsub, 0);
        bool any = false;
        for (int s = 0; s < nsub; s++) {
            cont[s] = (sub_off[s + 1] - sub_off[s]) > bottom_rows;
            any = any || cont[s];
        }
        if (!any) break;
        AmgLevelPlan L;
        L.n = n;
        L.sub_off = sub_off;
        L.cont = cont;
        for (int s = 0; s < nsub; s++)
            if (!cont[s] && sub_off[s + 1] > sub_off[s]) L.ends.push_back(s);
        if (!diag_positions(n, rs, col, L.diag, err)) return -1;
        // sub-problems whose first coarsening is the element level (level 0 only; a fixed
        // interpolation that does not coarsen falls back to aggregation)
        std::vector<char> fixed_sub(nsub, 0);
        std::vector<int32_t> rsub;
        if (plan.lv.empty() && fix) {
            if ((int64_t)fix->rs.size() != n + 1 || (int)fix->on.size() != nsub || (int)fix->nc.size() != nsub) {
                err = "multigrid: element-level interpolation of another pattern";
                return -1;
            }
            rsub.assign(n, 0);
            for (int s = 0; s < nsub; s++) {
                const int64_t rows = sub_off[s + 1] - sub_off[s];
                fixed_sub[s] = cont[s] && fix->on[s] && fix->nc[s] > 0 && fix->nc[s] < rows;
                for (int64_t r = sub_off[s]; r < sub_off[s + 1]; r++) {
                    rsub[r] = s;
                    if (!fixed_sub[s]) continue;
                    for (int32_t t = fix->rs[r]; t < fix->rs[r + 1]; t++)
                        if (fix->col[t] < 0 || fix->col[t] >= fix->nc[s] || (t > fix->rs[r] && fix->col[t] <= fix->col[t - 1])) {
                            err = "multigrid: element-level interpolation columns out of range or unsorted";
                            return -1;
                        }
                }
            }
        }
        auto fixed_row = [&](int64_t r) { return !rsub.empty() && fixed_sub[rsub[r]]; };
        // aggregation of the continuing sub-problems (independent), numbered in row order
        L.agg.assign(n, -1);
        std::vector<int64_t> nagg(nsub + 1, 0);
        bool stalled = false;
        parallel_for(nsub, [&](int64_t b, int64_t e) {
            for (int64_t s = b; s < e; s++) {
                if (!cont[s]) continue;
                if (fixed_sub[s]) {
                    // the row's first coarse column marks it as continuing (P comes from fix)
                    for (int64_t r = sub_off[s]; r < sub_off[s + 1]; r++)
                        L.agg[r] = fix->rs[r] < fix->rs[r + 1] ? fix->col[fix->rs[r]] : 0;
                    nagg[s + 1] = fix->nc[s];
                    continue;
                }
                int32_t na = 0;
                for (int64_t r = sub_off[s]; r < sub_off[s + 1]; r++) {
                    if (L.agg[r] >= 0) continue;
                    L.agg[r] = na;
                    int taken = 0;
                    for (int32_t k = rs[r]; k < rs[r + 1] && taken < cap; k++) {
                        const int32_t c = col[k];
                        // the pattern is block-diagonal over sub-problems; a column of another
                        // sub-problem is never taken, so aggregates stay inside their own
                        if (c != r && c >= sub_off[s] && c < sub_off[s + 1] && L.agg[c] < 0) {
                            L.agg[c] = na;
                            taken++;
                        }
                    }
                    na++;
                }
                nagg[s + 1] = na;
                if (na >= sub_off[s + 1] - sub_off[s]) stalled = true;
            }
        });
        if (stalled) {
            err = "multigrid: a sub-problem's rows do not coarsen (no couplings)";
            return -1;
        }
        std::vector<int64_t> coff(nsub + 1, 0);
        for (int s = 0; s < nsub; s++) coff[s + 1] = coff[s] + nagg[s + 1];
        parallel_for(nsub, [&](int64_t b, int64_t e) {
            for (int64_t s = b; s < e; s++)
                if (cont[s])
                    for (int64_t r = sub_off[s]; r < sub_off[s + 1]; r++) L.agg[r] += (int32_t)coff[s];
        });
        const int64_t nc = coff[nsub];
        L.nc = nc;
        // P pattern: aggregates of the row's neighbours (element level: the fixed columns)
        if (!build_rows(n, nc, [&](int64_t r, std::vector<int32_t> &out) {
                if (L.agg[r] < 0) return;   // the sub-problem ends on this level
                if (fixed_row(r)) {
                    for (int32_t t = fix->rs[r]; t < fix->rs[r + 1]; t++)
                        out.push_back((int32_t)(coff[rsub[r]] + fix->col[t]));
                    return;
                }
                for (int32_t k = rs[r]; k < rs[r + 1]; k++) out.push_back(L.agg[col[k]]);
            }, L.prs, L.pcol, err))
            return -1;
        // values before the smoothing: 1 on the row's own aggregate; the element level's
        // weights (its columns ascending and distinct, so in the same order as the pattern)
        L.pfix.assign(L.prs[n], 0.0);
        parallel_for(n, [&](int64_t b, int64_t e) {
            for (int64_t r = b; r < e; r++) {
                if (L.agg[r] < 0) continue;
                if (fixed_row(r)) {
                    for (int32_t q = L.prs[r], t = fix->rs[r]; q < L.prs[r + 1]; q++, t++) L.pfix[q] = fix->w[t];
                } else {
                    for (int32_t q = L.prs[r]; q < L.prs[r + 1]; q++) L.pfix[q] = L.pcol[q] == L.agg[r] ? 1.0 : 0.0;
                }
            }
        });
        // P contributions: for entry (r, J) the A entries k of row r with agg(col k) == J
        const int64_t nnzp = L.prs[n];
        L.pcptr.assign(nnzp + 1, 0);
        {
            // entry of P row r holding aggregate agg(col k) (P rows are sorted and distinct)
            auto pentry = [&](int64_t r, int32_t k) {
                const int32_t *b0 = L.pcol.data() + L.prs[r], *b1 = L.pcol.data() + L.prs[r + 1];
                return (int32_t)(std::lower_bound(b0, b1, L.agg[col[k]]) - L.pcol.data());
            };
            // (the element level's entries are not smoothed: no A entries)
            std::vector<int32_t> cnt(nnzp, 0);
            parallel_for(n, [&](int64_t b, int64_t e) {
                for (int64_t r = b; r < e; r++)
                    if (L.agg[r] >= 0 && !fixed_row(r))
                        for (int32_t k = rs[r]; k < rs[r + 1]; k++) cnt[pentry(r, k)]++;
            });
            for (int64_t q = 0; q < nnzp; q++) L.pcptr[q + 1] = L.pcptr[q] + cnt[q];
            L.pcidx.resize(L.pcptr[nnzp]);
            // each entry's list in increasing k: a row's entries are walked in k order, and
            // the entries of P row r belong to row r only
            std::vector<int32_t> fill(L.pcptr.begin(), L.pcptr.end() - 1);
            parallel_for(n, [&](int64_t b, int64_t e) {
                for (int64_t r = b; r < e; r++)
                    if (L.agg[r] >= 0 && !fixed_row(r))
                        for (int32_t k = rs[r]; k < rs[r + 1]; k++) L.pcidx[fill[pentry(r, k)]++] = k;
            });
        }
        // P^T (coarse row J: fine rows in increasing order)
        L.trs.assign(nc + 1, 0);
        for (int64_t q = 0; q < nnzp; q++) L.trs[L.pcol[q] + 1]++;
        for (int64_t J = 0; J < nc; J++) L.trs[J + 1] += L.trs[J];
        L.trow.resize(nnzp);
        L.tsrc.resize(nnzp);
        {
            std::vector<int32_t> fill(L.trs.begin(), L.trs.end() - 1);
            for (int64_t r = 0; r < n; r++)
                for (int32_t q = L.prs[r]; q < L.prs[r + 1]; q++) {
                    const int32_t o = fill[L.pcol[q]]++;
                    L.trow[o] = (int32_t)r;
                    L.tsrc[o] = q;
                }
        }
        // Q = A P pattern
        if (!build_rows(n, nc, [&](int64_t r, std::vector<int32_t> &out) {
                for (int32_t k = rs[r]; k < rs[r + 1]; k++) {
                    const int32_t c = col[k];
                    for (int32_t q = L.prs[c]; q < L.prs[c + 1]; q++) out.push_back(L.pcol[q]);
                }
            }, L.qrs, L.qcol, err))
            return -1;
        // A_{l+1} = P^T Q pattern
        std::vector<int32_t> crs, ccol;
        if (!build_rows(nc, nc, [&](int64_t J, std::vector<int32_t> &out) {
                for (int32_t t = L.trs[J]; t < L.trs[J + 1]; t++) {
                    const int32_t r = L.trow[t];
                    for (int32_t q = L.qrs[r]; q < L.qrs[r + 1]; q++) out.push_back(L.qcol[q]);
                }
            }, crs, ccol, err))
            return -1;
        L.rs = std::move(rs);
        L.col = std::move(col);
        plan.lv.push_back(std::move(L));
        rs = std::move(crs);
        col = std::move(ccol);
        sub_off = coff;
        n = nc;
    }
    plan.nbot = n;
    plan.bot_off = sub_off;
    for (int s = 0; s < nsub; s++)
        if (sub_off[s + 1] > sub_off[s]) plan.bot_subs.push_back(s);
    if (!diag_positions(n, rs, col, plan.bot_diag, err)) return -1;
    plan.bot_rs = std::move(rs);
    plan.bot_col = std::move(col);
    int64_t mb = max_sub(plan.bot_off);
    for (const auto &L : plan.lv)
        for (int s : L.ends) mb = std::max(mb, L.sub_off[s + 1] - L.sub_off[s]);
    plan.max_bot_rows = (int)mb;
    return 0;
}

void amg_element_level_rows(int kind, int64_t nnode, int64_t nelem, const int32_t *conn, const int32_t *eqn,
                            int nvalue, int ncomp, int64_t ndof, bool applies, AmgFixedInterp &F) {
    const bool quad = kind == PYHIP_ELEM_TRI_P2 || kind == PYHIP_ELEM_QUAD_Q2 || kind == PYHIP_ELEM_HEX_Q2;
    const int64_t base = (int64_t)F.col.size();
    if (!quad || !applies || ndof == 0) {
        F.on.push_back(0);
        F.nc.push_back(0);
        for (int64_t q = 0; q < ndof; q++) F.rs.push_back((int32_t)base);
        return;
    }
    const int nn = kind == PYHIP_ELEM_TRI_P2 ? 6 : (kind == PYHIP_ELEM_QUAD_Q2 ? 9 : 27);
    // vertex local nodes and the weight of vertex v at local node l
    int nv = 0, vloc[8];
    double wt[27][8] = {};
    if (kind == PYHIP_ELEM_TRI_P2) {
        nv = 3;
        for (int v = 0; v < 3; v++) vloc[v] = v, wt[v][v] = 1.0;
        const int mid[3][2] = {{0, 1}, {1, 2}, {0, 2}};   // nodes 3, 4, 5
        for (int m = 0; m < 3; m++) wt[3 + m][mid[m][0]] = wt[3 + m][mid[m][1]] = 0.5;
    } else {
        // tensor nodes l = i + 3 j (+ 9 k), i, j, k in 0..2; vertices at 0 and 2
        const int dim = kind == PYHIP_ELEM_HEX_Q2 ? 3 : 2;
        nv = 1 << dim;
        for (int v = 0; v < nv; v++) {
            int l = 0, st = 1;
            for (int a = 0; a < dim; a++, st *= 3) l += ((v >> a) & 1) * 2 * st;
            vloc[v] = l;
        }
        for (int l = 0; l < nn; l++)
            for (int v = 0; v < nv; v++) {
                double w = 1.0;
                for (int a = 0, t = l; a < dim; a++, t /= 3) {
                    const int i = t % 3, hi = (v >> a) & 1;
                    w *= i == 1 ? 0.5 : ((i == 2) == (hi == 1) ? 1.0 : 0.0);
                }
                wt[l][v] = w;
            }
    }
    std::vector<char> isv(nnode, 0);
    for (int64_t e = 0; e < nelem; e++)
        for (int v = 0; v < nv; v++) isv[conn[e * nn + vloc[v]]] = 1;
    // coarse unknowns: (vertex, component) pairs that are unknowns, node-major
    std::vector<int32_t> cdof((size_t)nnode * ncomp, -1);
    int64_t nc = 0;
    for (int64_t n = 0; n < nnode; n++)
        for (int c = 0; c < ncomp; c++)
            if (isv[n] && eqn[n * nvalue + c] >= 0) cdof[(size_t)n * ncomp + c] = (int32_t)nc++;
    // each unknown's entries from the first element holding it (at most 8, ascending; the
    // vertices' unknowns of its own component)
    std::vector<int8_t> cnt(ndof, -1);
    std::vector<int32_t> ecol((size_t)ndof * nv);
    std::vector<double> ew((size_t)ndof * nv);
    for (int64_t e = 0; e < nelem; e++)
        for (int l = 0; l < nn; l++)
            for (int comp = 0; comp < ncomp; comp++) {
                const int32_t n = conn[e * nn + l], q = eqn[(int64_t)n * nvalue + comp];
                if (q < 0 || cnt[q] >= 0) continue;
                int c = 0;
                for (int v = 0; v < nv; v++) {
                    const int32_t cv = cdof[(size_t)conn[e * nn + vloc[v]] * ncomp + comp];
                    if (wt[l][v] == 0.0 || cv < 0) continue;
                    int k = c++;
                    while (k > 0 && ecol[(size_t)q * nv + k - 1] > cv) {
                        ecol[(size_t)q * nv + k] = ecol[(size_t)q * nv + k - 1];
                        ew[(size_t)q * nv + k] = ew[(size_t)q * nv + k - 1];
                        k--;
                    }
                    ecol[(size_t)q * nv + k] = cv;
                    ew[(size_t)q * nv + k] = wt[l][v];
                }
                cnt[q] = (int8_t)c;
            }
    F.on.push_back(nc > 0 && nc < ndof ? 1 : 0);
    F.nc.push_back(nc);
    for (int64_t q = 0; q < ndof; q++) {
        F.rs.push_back((int32_t)F.col.size());
        for (int k = 0; k < std::max<int>(0, cnt[q]); k++) {
            F.col.push_back(ecol[(size_t)q * nv + k]);
            F.w.push_back(ew[(size_t)q * nv + k]);
        }
    }
}

}  // namespace pyhip
