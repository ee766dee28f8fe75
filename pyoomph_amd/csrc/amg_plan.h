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
//
// Element level: a sub-problem on quadratic elements (P2 triangles, Q2 quadrilaterals and
// bricks) may instead take, as its first coarse level, the linear functions on the element
// vertices: P is the fixed interpolation of those functions into the element nodes (1 at
// a vertex, the linear shape functions' values at the other nodes), so A_1 = P^T A_0 P is
// the operator on the P1 / Q1 subspace, which the aggregation then coarsens further.
// Aggregation coarsens a quadratic element's stencil poorly (a conjugate-gradient solve
// preconditioned by its V-cycle needs about twice the iterations it needs on linear
// elements); this level restores the linear elements' rate.
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
    std::vector<double> pfix;              // per P entry: its value before the smoothing,
                                           // 1 or 0 (aggregates), the weight (element level)
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

// Fixed first-level interpolation of the element level, over the level-0 rows: the rows of
// a sub-problem with on[s] != 0 hold their P entries (columns numbered within the
// sub-problem's nc[s] coarse rows, ascending; weights w), the other sub-problems' rows are
// empty and coarsen by aggregation.
struct AmgFixedInterp {
    std::vector<int32_t> on;               // [nsub]
    std::vector<int64_t> nc;               // [nsub]
    std::vector<int32_t> rs, col;          // [n+1], entries
    std::vector<double> w;
};

// Appends one sub-problem's rows of the element level to F (its unknowns' rows; F.rs gets
// no closing entry): on P2 triangles, Q2 quadrilaterals and Q2 bricks (PYHIP_ELEM_*) when
// `applies`, the interpolation of the linear (P1 / Q1) functions on the element vertices
// into the element nodes, for each of the nodal values 0..ncomp-1 (a scalar field, or the
// velocity components of the Taylor-Hood layout) separately; the rows of other values
// stay empty. conn [nelem][nodes per element], eqn [nnode][nvalue] (the unknown or -1).
// Coarse unknowns: the (vertex, value) pairs that are unknowns, node-major. A vertex row
// holds itself with weight 1; a node between vertices (edge midpoints, face and cell
// centres) holds the vertices' linear shape functions at its position (1/2, 1/4, 1/8),
// pinned vertices left out.
void amg_element_level_rows(int elem_kind, int64_t nnode, int64_t nelem, const int32_t *conn, const int32_t *eqn,
                            int nvalue, int ncomp, int64_t ndof, bool applies, AmgFixedInterp &F);

// Builds the plan from the level-0 pattern. A sub-problem stops coarsening once it has at
// most `bottom_rows` rows; the hierarchy has at most `max_levels` coarsenings (the
// sub-problems still larger then end on the last level). Returns 0 or -1 (err).
// With `fix`, its sub-problems take the element level as their first coarsening.
int amg_plan_build(int64_t n, const int32_t *rs, const int32_t *col, const std::vector<int64_t> &sub_off,
                   int cap, int bottom_rows, int max_levels, AmgPlan &plan, std::string &err,
                   const AmgFixedInterp *fix = nullptr);

}  // namespace pyhip
