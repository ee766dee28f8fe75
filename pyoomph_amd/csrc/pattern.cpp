// Host-side sparsity pattern and element-to-nnz map (see pattern.h).
#include "pattern.h"

#include <algorithm>
#include <string>
#include <thread>

namespace pyhip {

static unsigned pattern_threads() {
    return std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
}

template <typename F>
static void parallel_rows(int64_t n, F &&fn) {
    unsigned nt = pattern_threads();
    if (n < 20000) nt = 1;
    std::vector<std::thread> th;
    const int64_t chunk = (n + nt - 1) / nt;
    for (unsigned t = 0; t < nt; t++) {
        const int64_t b = (int64_t)t * chunk, e = std::min<int64_t>(n, b + chunk);
        if (b >= e) break;
        th.emplace_back([&fn, b, e, t]() { fn(b, e, t); });
    }
    for (auto &x : th) x.join();
}

int build_pattern(int64_t ndof, const std::vector<ElementBlock> &blocks, HostPattern &P,
                  std::string &err) {
    P = HostPattern();
    P.ndof = ndof;
    if (ndof >= INT32_MAX) {
        err = "too many equations";
        return -1;
    }
    // 1. occurrences per row
    std::vector<int64_t> cnt(ndof + 1, 0);
    for (const auto &B : blocks) {
        if (B.ns > kMaxSlot) {
            err = "too many dofs per element";
            return -1;
        }
        if (B.ej_off + B.nel * B.ns * B.ns > (int64_t)UINT32_MAX) {
            err = "element matrices exceed 2^32 entries";
            return -1;
        }
        for (int64_t e = 0; e < B.nel; e++)
            for (int i = 0; i < B.ns; i++) {
                const int32_t r = B.seqn[e * B.ns + i];
                if (r >= ndof) {
                    err = "equation number out of range";
                    return -1;
                }
                if (r >= 0) cnt[r + 1]++;
            }
    }
    P.occ_ptr.assign(ndof + 1, 0);
    for (int64_t r = 0; r < ndof; r++) P.occ_ptr[r + 1] = P.occ_ptr[r] + cnt[r + 1];
    P.nocc = P.occ_ptr[ndof];
    P.occ_roff.resize(P.nocc);
    std::vector<int64_t> occ_eoff(P.nocc);
    std::vector<const int32_t *> srow(P.nocc);
    std::vector<int8_t> sns(P.nocc);
    // 2. fill in element order (so every row's occurrences are in element order)
    std::vector<int64_t> fill(P.occ_ptr.begin(), P.occ_ptr.end() - 1);
    for (const auto &B : blocks)
        for (int64_t e = 0; e < B.nel; e++)
            for (int i = 0; i < B.ns; i++) {
                const int32_t r = B.seqn[e * B.ns + i];
                if (r < 0) continue;
                const int64_t o = fill[r]++;
                P.occ_roff[o] = B.er_off + e * B.ns + i;
                occ_eoff[o] = B.ej_off + (e * B.ns + i) * B.ns;
                srow[o] = B.seqn + e * B.ns;
                sns[o] = (int8_t)B.ns;
            }
    // 3a. unique columns per row
    std::vector<int64_t> rlen(ndof + 1, 0);
    const unsigned nthr = pattern_threads();
    std::vector<std::vector<int32_t>> marks(nthr);
    bool toolong = false;
    parallel_rows(ndof, [&](int64_t b, int64_t e, unsigned t) {
        auto &mark = marks[t];
        mark.assign(ndof, -1);
        for (int64_t r = b; r < e; r++) {
            int64_t n = 0;
            for (int64_t o = P.occ_ptr[r]; o < P.occ_ptr[r + 1]; o++)
                for (int j = 0; j < sns[o]; j++) {
                    const int32_t c = srow[o][j];
                    if (c >= 0 && mark[c] != r) {
                        mark[c] = (int32_t)r;
                        n++;
                    }
                }
            rlen[r + 1] = n;
            if (n > kMaxRowLen) toolong = true;
        }
    });
    if (toolong) {
        err = "a row has more than 255 entries";
        return -1;
    }
    P.row_start.assign(ndof + 1, 0);
    int64_t acc = 0;
    for (int64_t r = 0; r < ndof; r++) {
        acc += rlen[r + 1];
        if (acc > INT32_MAX) {
            err = "nnz exceeds int32 range";
            return -1;
        }
        P.row_start[r + 1] = (int32_t)acc;
    }
    P.nnz = acc;
    P.col.resize(P.nnz);
    P.diag_pos.assign(ndof, -1);
    std::vector<int32_t> ccount(P.nnz, 0);
    // 3b. sorted columns, diagonal, contributions per entry (marks become column -> position)
    parallel_rows(ndof, [&](int64_t b, int64_t e, unsigned t) {
        auto &pos = marks[t];
        for (int64_t r = b; r < e; r++) {
            int32_t *cs = P.col.data() + P.row_start[r];
            const int64_t n = rlen[r + 1];
            int64_t k = 0;
            for (int64_t o = P.occ_ptr[r]; o < P.occ_ptr[r + 1]; o++)
                for (int j = 0; j < sns[o]; j++) {
                    const int32_t c = srow[o][j];
                    if (c >= 0 && pos[c] != -2 - (int32_t)r) {
                        pos[c] = -2 - (int32_t)r;
                        cs[k++] = c;
                    }
                }
            std::sort(cs, cs + n);
            for (int64_t q = 0; q < n; q++) pos[cs[q]] = (int32_t)q;
            for (int64_t o = P.occ_ptr[r]; o < P.occ_ptr[r + 1]; o++)
                for (int j = 0; j < sns[o]; j++) {
                    const int32_t c = srow[o][j];
                    if (c >= 0) ccount[P.row_start[r] + pos[c]]++;
                }
            const int32_t *d = std::lower_bound(cs, cs + n, (int32_t)r);
            if (d != cs + n && *d == r) P.diag_pos[r] = (int32_t)(d - P.col.data());
        }
    });
    P.cptr.assign(P.nnz + 1, 0);
    int64_t nc = 0;
    for (int64_t q = 0; q < P.nnz; q++) {
        nc += ccount[q];
        if (nc > INT32_MAX) {
            err = "element contributions exceed int32 range";
            return -1;
        }
        P.cptr[q + 1] = (int32_t)nc;
    }
    P.ncontrib = nc;
    P.cidx.resize(nc);
    // 3c. contribution lists, element order within every entry
    parallel_rows(ndof, [&](int64_t b, int64_t e, unsigned t) {
        auto &pos = marks[t];
        std::vector<int32_t> next;
        for (int64_t r = b; r < e; r++) {
            const int32_t s = P.row_start[r], n = P.row_start[r + 1] - s;
            for (int32_t q = 0; q < n; q++) pos[P.col[s + q]] = q;
            next.assign(P.cptr.begin() + s, P.cptr.begin() + s + n);
            for (int64_t o = P.occ_ptr[r]; o < P.occ_ptr[r + 1]; o++)
                for (int j = 0; j < sns[o]; j++) {
                    const int32_t c = srow[o][j];
                    if (c >= 0) P.cidx[next[pos[c]]++] = (uint32_t)(occ_eoff[o] + j);
                }
        }
        pos.clear();
        pos.shrink_to_fit();
    });
    return 0;
}

}  // namespace pyhip
