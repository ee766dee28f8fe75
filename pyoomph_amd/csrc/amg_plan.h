// Symbolic set-up of the smoothed-aggregation multigrid preconditioner (host, once per
// sparsity pattern). No reference counterpart: pyoomph solves its linear systems with
// SuperLU (pyoomph/solvers/scipy.py:47-104); this preconditioner serves the batched
// Krylov solvers that replace that direct solve.
//
// Level l has the matrix A_l (level 0: the problem's column-sorted pattern). A
// sub-problem with more than `bottom_rows` rows on level l continues: its rows are grouped
// into aggregates (in row order, a row not yet aggregated starts an aggregate and takes up
// to `cap` of its not-yet-aggregated neighbours, column order). A sub-problem with fewer
// rows ends on level l: it gets its coarsest-level solve there and has no rows below.
// So every sub-problem gets the hierarchy it would get alone, whatever the batch. The
// pattern is block-diagonal over the sub-problems, so aggregates never cross them, and the
// coarse rows of a sub-problem are contiguous. For the continuing rows:
//   P   = (I - W A_l) P_t, P_t[i][agg(i)] = 1, W = diag(omega / a_ii)   (n_l x n_{l+1})
//   Q   = A_l P                                                         (n_l x n_{l+1})
//   A_{l+1} = P^T Q
// The plan holds every pattern and, for each entry of P, the list of A_l entries whose
// sum it needs; values are computed on the device for every new matrix.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace pyhip {

struct AmgLevelPlan {
    int64_t n = 0, nc = 0;                 // rows of this level and of the next
    std::vector<int64_t> sub_off;          // [nsub+1] row offsets per sub-problem
    std::vector<int32_t> cont;             // [nsub] 1 if the sub-problem continues below
    std::vector<int32_t> ends;             // sub-problems whose coarsest solve is on this level
    std::vector<int32_t> rs, col, diag;    // A_l pattern (column-sorted) and diagonal positions
    std::vector<int32_t> agg;              // [n] coarse row of each row (-1: sub-problem ends)
    std::vector<int32_t> prs, pcol;        // P pattern [n+1], [nnzP]
    std::vector<int32_t> pcptr, pcidx;     // per P entry: A_l entries summed (column order)
    std::vector<int32_t> trs, trow, tsrc;  // P^T [nc+1], fine row, P entry index
    std::vector<int32_t> qrs, qcol;        // Q = A_l P pattern
};

struct AmgPlan {
    std::vector<AmgLevelPlan> lv;          // levels with a coarser level below them
    // coarsest level: pattern and sub-problem offsets
    int64_t nbot = 0;
    std::vector<int64_t> bot_off;
    std::vector<int32_t> bot_subs;         // sub-problems with rows on the coarsest level
    std::vector<int32_t> bot_rs, bot_col, bot_diag;
    int max_bot_rows = 0;                  // largest coarsest-level solve, over all levels
};

// Builds the plan from the level-0 pattern. A sub-problem stops coarsening once it has at
// most `bottom_rows` rows; the hierarchy has at most `max_levels` coarsenings (the
// sub-problems still larger then end on the last level). Returns 0 or -1 (err).
int amg_plan_build(int64_t n, const int32_t *rs, const int32_t *col, const std::vector<int64_t> &sub_off,
                   int cap, int bottom_rows, int max_levels, AmgPlan &plan, std::string &err);

}  // namespace pyhip
