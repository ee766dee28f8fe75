// Batched Krylov solver workspace (GMRES + Jacobi / ILU(0)), see krylov.hip.
#pragma once
#include <vector>

#include "amg.h"
#include "problem.h"

namespace pyhip {

constexpr int kChunk = 1024;  // rows per reduction chunk (chunks never straddle sub-problems)
constexpr int kMaxRestart = 1000;   // longest GMRES restart cycle
constexpr int kWlRows = 256;         // rows per block of the work-list kernels
constexpr int kWlBlocksPerChunk = kChunk / kWlRows;

struct IluPlan {
    bool built = false;
    int requested = -1;           // PYHIP_ILU_ORDER_* the plan was built for
    bool multicolor = false;      // some sub-problem takes the multicolour order
    int64_t n = 0, nnz = 0;
    int32_t *perm = nullptr;      // [n] new -> old
    int32_t *rs = nullptr;        // [n+1] permuted, column-sorted
    int32_t *col = nullptr;       // [nnz]
    int32_t *src = nullptr;       // [nnz] index of the entry in the problem's CSR values
    int32_t *diag = nullptr;      // [n]
    int32_t *lo_rows = nullptr;   // rows grouped by lower-solve / factorisation level
    int32_t *up_rows = nullptr;   // rows grouped by upper-solve level
    std::vector<int64_t> lo_ptr, up_ptr;   // level boundaries (host)
    void *lu = nullptr;           // [nnz] factor values (T)
    int lu_prec = -1;
    double *tmp64 = nullptr;      // [n] permuted work vectors
    double *tmp64b = nullptr;
    float *tmp32 = nullptr;
    float *tmp32b = nullptr;
};

// Block-triangular preconditioner of Taylor-Hood Jacobians [A G; D 0] (block_precond.hip):
//   z_p = (mu / m_p) r_p per pressure unknown (the lumped Q1 pressure mass scaled by the
//   viscosity approximating the Schur complement), z_u = V-cycle(A_d; r_u - G z_p), with A_d
//   the velocity block without the coupling of different velocity components (rows of
//   scalar forms: their whole matrix), on its own multigrid hierarchy.
struct BlockPlan {
    bool built = false;
    bool element_level = false;    // A_d's multigrid built with the element level
    int64_t nv = 0, np = 0, nnzv = 0;
    int32_t *vrows = nullptr;      // [nv] problem row of each velocity-block row
    int32_t *vsrc = nullptr;       // [nnzv] index of each A_d entry in the problem's CSR values
    int32_t *prow = nullptr;       // [np] problem row of each pressure unknown
    int32_t *psub = nullptr;       // [np] its sub-problem
    int32_t *vsub = nullptr;       // [nv] sub-problem of each velocity-block row
    double *pmass = nullptr;       // [np] lumped pressure mass
    double *sinv = nullptr;        // [np] mu / pmass (set-up)
    double *vals = nullptr;        // [nnzv] A_d values
    double *tv = nullptr, *zv = nullptr;   // [nv]
    double *pz = nullptr, *t = nullptr;    // [n]: z_p in pressure rows (0 elsewhere); G z_p
    const double *jvals = nullptr;         // the Jacobian values of the last set-up
    std::vector<int32_t> rs_h, col_h;      // A_d pattern
    std::vector<int64_t> off_h;            // [nsub+1] A_d rows per sub-problem
    AmgDev amg;
};

struct KrylovWork {
    int64_t n = 0;
    int nsub = 0;
    int m = 0;                    // restart length the buffers are sized for
    int64_t nchunk = 0;
    int32_t *chunk_sub = nullptr;     // [nchunk]
    int64_t *chunk_begin = nullptr;   // [nchunk+1] (row ranges; chunk c = [begin[c], end[c]))
    int64_t *chunk_end = nullptr;
    int32_t *sub_chunk_ptr = nullptr; // [nsub+1]
    double *partial = nullptr;        // [nchunk][m+2]
    double *dots = nullptr;           // [m+2][nsub]
    // GMRES state per sub-problem
    double *H = nullptr;              // [nsub][m+1][m]
    double *cs = nullptr, *sn = nullptr;  // [nsub][m]
    double *g = nullptr;              // [nsub][m+1]
    double *beta0 = nullptr;          // [nsub]
    double *bnorm = nullptr;          // [nsub]
    int32_t *kcols = nullptr;         // [nsub] columns in use
    int32_t *active = nullptr;        // [nsub] 1 while iterating
    int32_t *nactive_h = nullptr;     // pinned host counter
    int32_t *nactive_d = nullptr;
    double *y = nullptr;              // [nsub][m]
    // vectors
    void *V = nullptr;                // [(m+1)][n] basis (T)
    int vprec = -1;
    int vcount = 0;                   // vectors allocated in V
    double *r64 = nullptr, *x64 = nullptr, *d64 = nullptr, *b64 = nullptr;  // [n]
    float *r32 = nullptr, *d32 = nullptr, *w32 = nullptr, *t32 = nullptr;
    double *w64 = nullptr, *t64 = nullptr;
    float *vals32 = nullptr;          // FP32 copy of the Jacobian values (when needed)
    double *vals64 = nullptr;         // FP64 Jacobian values (Newton)
    double *res64 = nullptr;          // Newton residual
    double *maxres = nullptr;         // [nsub]
    double *maxres_h = nullptr;       // pinned [nsub]
    double *sub_scalar = nullptr;     // [nsub] scratch
    double *jdiag64 = nullptr;        // [n] inverse diagonal (Jacobi)
    float *jdiag32 = nullptr;
    int32_t *dof_mask = nullptr;      // [nsub] Newton: 1 if the sub-problem takes a step
    int32_t *need = nullptr;          // [nsub] refinement: sub-problems still refining
    double *cg_rz = nullptr, *cg_pq = nullptr, *cg_beta = nullptr;   // [nsub]
    int32_t *nfail_d = nullptr;       // breakdown / non-convergence counter
    // single-reduction CG (work list over active chunks)
    int32_t *achunk = nullptr;        // [nchunk] chunks of the sub-problems iterating
    int32_t *achunk_h = nullptr;      // pinned [nchunk]
    double *gpart = nullptr;          // [3][nchunk*kWlBlocksPerChunk] block partials
    int32_t *cg_iters = nullptr;      // [nsub] iterations done in the current solve
    int32_t *cg_fail = nullptr;       // [nsub] breakdown / non-finite residual
    int32_t *cg_iters2 = nullptr;     // [nsub] preconditioned CG: preconditioner applies
    int32_t *flags_h = nullptr;       // pinned [3][nsub]: active, fail, iterations
    std::vector<int64_t> sub_nnz_h;   // [nsub] pattern entries per sub-problem
    std::vector<int32_t> chunk_ptr_h; // [nsub+1] host copy of sub_chunk_ptr
    int lpr = 0;                      // work-list spmv lanes per row (1, 2, 4, 8), 0 = not tuned
    int smode = 0;                    // spmv variant (spmv.h: lanes per row or LDS staging)
    int32_t *act_h = nullptr;         // [nsub] host copy of the active flags (pinned)
    // arclength continuation: Dof_current, Dof_derivative, y, z, scratch [n]; scalars [nsub]
    double *ac_cur = nullptr, *ac_der = nullptr, *ac_y = nullptr, *ac_z = nullptr, *ac_t = nullptr;
    double *ac_zs = nullptr;          // z of each sub-problem's last continuation Newton step
    double *ac_alpha = nullptr;
    int32_t *ac_mask = nullptr;
    IluPlan ilu;
    int ilu_order = 0;                // PYHIP_ILU_ORDER_* of the current solve
    AmgDev amg;
    BlockPlan blk;
    bool amg_stale = true;            // multigrid values to be rebuilt at the next set-up
    std::vector<int32_t> need_h;      // host copy of the sub-problems a preconditioner set-up serves
    std::vector<uint8_t> step_res_ok; // [nsub] res64 / maxres_h hold the sub-problem's residual at its dofs
    const double *amg_src64 = nullptr;   // FP64 matrix the FP32 V-cycle's set-up reads (or null)
    double *amg_conv64 = nullptr;     // [nnz] FP32 matrix widened for that set-up
    // per sub-problem: 1 if the last Krylov solve (or refinement) ended above its tolerance
    // or broke down; written by every solver before it returns
    std::vector<int32_t> lfail_h;
    int64_t bytes = 0;
};

}  // namespace pyhip
