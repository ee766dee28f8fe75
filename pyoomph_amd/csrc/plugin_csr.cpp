// Host-side CSR preparation for the linear-solver plugin entry (see plugin_csr.h).
#include "plugin_csr.h"

#include <algorithm>

namespace pyhip {

int plugin_build_csr(int n, int nnz, const double *values, const int32_t *rowind,
                     const int32_t *colptr, int transpose, std::vector<int32_t> &rs,
                     std::vector<int32_t> &cl, std::vector<int64_t> &src,
                     std::vector<char> &last, std::string &err) {
    if (n <= 0 || nnz < 0 || !values || !rowind || !colptr) {
        err = "invalid matrix";
        return -1;
    }
    if (colptr[0] != 0 || colptr[n] != nnz) {
        err = "inconsistent column pointers";
        return -1;
    }
    for (int i = 0; i < n; i++)
        if (colptr[i] > colptr[i + 1]) {
            err = "inconsistent column pointers";
            return -1;
        }
    for (int k = 0; k < nnz; k++)
        if (rowind[k] < 0 || rowind[k] >= n) {
            err = "index out of range";
            return -1;
        }
    std::vector<std::vector<std::pair<int32_t, int64_t>>> rows(n);
    if (transpose) {
        for (int i = 0; i < n; i++)
            for (int32_t k = colptr[i]; k < colptr[i + 1]; k++) rows[i].push_back({rowind[k], k});
    } else {
        for (int c = 0; c < n; c++)
            for (int32_t k = colptr[c]; k < colptr[c + 1]; k++) rows[rowind[k]].push_back({c, k});
    }
    std::vector<char> zero_diag(n, 1);
    for (int i = 0; i < n; i++)
        for (auto &e : rows[i])
            if (e.first == i && values[e.second] != 0.0) zero_diag[i] = 0;
    {
        std::vector<std::vector<int32_t>> zcols(n);  // c -> zero-diagonal rows p with (p, c)
        for (int p = 0; p < n; p++)
            if (zero_diag[p])
                for (auto &e : rows[p])
                    if (!zero_diag[e.first]) zcols[e.first].push_back(p);
        for (int p = 0; p < n; p++) {
            if (!zero_diag[p]) continue;
            std::vector<int32_t> add;
            for (auto &e : rows[p])
                if (!zero_diag[e.first])
                    for (int32_t q : zcols[e.first]) add.push_back(q);
            add.push_back(p);
            std::sort(add.begin(), add.end());
            add.erase(std::unique(add.begin(), add.end()), add.end());
            std::vector<int32_t> have;
            for (auto &e : rows[p]) have.push_back(e.first);
            std::sort(have.begin(), have.end());
            for (int32_t q : add)
                if (!std::binary_search(have.begin(), have.end(), q)) rows[p].push_back({q, -1});
        }
    }
    rs.assign(n + 1, 0);
    for (int i = 0; i < n; i++) {
        std::sort(rows[i].begin(), rows[i].end());
        rs[i + 1] = rs[i] + (int32_t)rows[i].size();
    }
    const int64_t total = rs[n];
    cl.assign(total, 0);
    src.assign(total, -1);
    for (int i = 0; i < n; i++)
        for (size_t t = 0; t < rows[i].size(); t++) {
            cl[rs[i] + t] = rows[i][t].first;
            src[rs[i] + t] = rows[i][t].second;
        }
    last.assign(zero_diag.begin(), zero_diag.end());
    return 0;
}

}  // namespace pyhip

// C entry used by the CPU tests of this preparation step (no device involved).
extern "C" int pyhip_plugin_build_csr_host(int n, int nnz, const double *values,
                                           const int32_t *rowind, const int32_t *colptr,
                                           int transpose, int32_t *rs_out, int32_t *cl_out,
                                           int64_t *src_out, int64_t cap) {
    std::vector<int32_t> rs, cl;
    std::vector<int64_t> src;
    std::vector<char> last;
    std::string err;
    if (pyhip::plugin_build_csr(n, nnz, values, rowind, colptr, transpose, rs, cl, src, last, err))
        return -1;
    if ((int64_t)cl.size() > cap) return -2;
    std::copy(rs.begin(), rs.end(), rs_out);
    std::copy(cl.begin(), cl.end(), cl_out);
    std::copy(src.begin(), src.end(), src_out);
    return (int)cl.size();
}
