// Smoothed-aggregation multigrid preconditioner: device data and entry points (amg.hip).
#pragma once
#include <vector>

#include "amg_plan.h"
#include "common.h"

namespace pyhip {

constexpr int kAggregateCap = 16;      // neighbours an aggregate seed takes at most
constexpr int kBottomRows = 48;        // a sub-problem stops coarsening at this many rows
constexpr int kBottomMaxRows = 64;     // dense Cholesky size of the coarsest solve (LDS)
constexpr int kMaxLevels = 16;

struct AmgDevLevel {
    int64_t n = 0, nc = 0, nnz = 0, nnzp = 0, nnzq = 0;
    int32_t *rs = nullptr, *col = nullptr, *diag = nullptr, *agg = nullptr;
    int32_t *prs = nullptr, *pcol = nullptr, *pcptr = nullptr, *pcidx = nullptr;
    int32_t *trs = nullptr, *trow = nullptr, *tsrc = nullptr, *qrs = nullptr, *qcol = nullptr;
    int32_t *prow = nullptr, *tinv = nullptr;   // row of each P entry; its position in P^T
    int32_t *row_sub = nullptr;
    int64_t *sub_off = nullptr;
    int32_t *cont = nullptr;           // [nsub] 1: the sub-problem continues below this level
    int32_t *ends = nullptr;           // sub-problems solved on this level (coarsest solve)
    int nends = 0;
    const double *vals = nullptr;      // A_l values (level 0: the problem's)
    double *own_vals = nullptr;        // A_l values computed here (levels > 0)
    double *d = nullptr, *w = nullptr, *pv = nullptr, *qv = nullptr, *ptv = nullptr;
    double *pfix = nullptr;            // per P entry: its value before the smoothing
    double *b = nullptr, *x = nullptr, *x2 = nullptr, *r = nullptr, *pw = nullptr, *py = nullptr;
    // host: per sub-problem rows, A entries and P entries (for counting the bytes of the
    // rows a filtered V-cycle visits)
    std::vector<int64_t> h_rows, h_nnz, h_nnzp;
    // spmv variant (spmv.h) of the level's launches (A_l, restriction P^T, prolongation P):
    // 0 until tuned (all variants give the same bits)
    int lpr_a = 0, lpr_t = 0, lpr_p = 0;
    // FP32 V-cycle (amg_make_f32): values rounded from the FP64 set-up, FP32 work vectors
    const float *vals32 = nullptr;     // A_l values (level 0: the caller's FP32 matrix)
    float *own_vals32 = nullptr, *w32 = nullptr, *pv32 = nullptr, *ptv32 = nullptr;
    float *b32 = nullptr, *x32 = nullptr, *x2_32 = nullptr, *r32 = nullptr;
};

struct AmgDev {
    bool built = false;
    bool element_level = false;        // built with the element level (an AmgFixedInterp)
    // non-symmetric operators (a convection-diffusion velocity block): the coarsest solve
    // uses a Gauss-Jordan inverse instead of the Cholesky factor (set before amg_build)
    bool nonsym = false;
    int nsub = 0;
    int max_bot_rows = 0;
    std::vector<AmgDevLevel> lv;       // levels with a coarser level
    AmgDevLevel bot;                   // coarsest level
    double *nrm = nullptr;             // [2*nsub] power-iteration norms
    double *chol = nullptr;            // [nsub][kBottomMaxRows^2] coarsest-level Cholesky factors
    int32_t *chol_fail = nullptr;      // [nsub] 1 if a pivot was not positive
    float *chol32 = nullptr;           // FP32 copy of the coarsest inverses
};

// plan (host) and device buffers, once per pattern; sizes counted in `bytes`; with fix,
// its sub-problems coarsen first to the element level
int amg_build(AmgDev &A, int64_t n, const int32_t *rs_h, const int32_t *col_h,
              const std::vector<int64_t> &sub_off, int64_t &bytes, const AmgFixedInterp *fix = nullptr);
struct KernelTimer;
// numeric set-up for the matrix values `vals` (level-0 pattern); with a timer, the
// power-iteration spmv launches are timed and counted as "spmv" launches
// need_h: host flags [nsub] of the sub-problems whose hierarchy is used (null: all); a
// coarsest matrix that is not positive definite is reported only for those
int amg_setup(AmgDev &A, hipStream_t st, const double *vals, KernelTimer *tm = nullptr,
              const int32_t *need_h = nullptr);
// the finest level's values (the current matrix; the rest of the hierarchy is kept)
int amg_set_fine(AmgDev &A, const double *vals);
// z = M^-1 b. With a timer, every level spmv is timed and counted as an "spmv" launch.
// With act (device [nsub], host copy act_h) only the rows of sub-problems with act[s] != 0
// are computed; the other rows of z keep their values.
// x0_ready: the caller has already written the finest level's first smoothing x = W b
// into amg_fine_x (see amg_fine_w), e.g. fused into the update that produced b.
int amg_apply(AmgDev &A, hipStream_t st, const double *b, double *z, KernelTimer *tm = nullptr,
              const int32_t *act = nullptr, const int32_t *act_h = nullptr, bool x0_ready = false);
// the finest level's smoothing weights W and first-sweep vector (nullptr without levels)
inline const double *amg_fine_w(const AmgDev &A) { return A.lv.empty() ? nullptr : A.lv[0].w; }
inline double *amg_fine_x(const AmgDev &A) { return A.lv.empty() ? nullptr : A.lv[0].x; }
// FP32 V-cycle: after amg_setup (FP64), FP32 copies of every level's values, weights, P,
// P^T and the coarsest inverses; then the finest FP32 matrix and z = M^-1 b in FP32 (the
// same cycle as amg_apply, each operation rounded to FP32).
int amg_make_f32(AmgDev &A, hipStream_t st, int64_t &bytes);
int amg_set_fine32(AmgDev &A, const float *vals);
int amg_apply32(AmgDev &A, hipStream_t st, const float *b, float *z, KernelTimer *tm = nullptr,
                const int32_t *act = nullptr, const int32_t *act_h = nullptr);
void amg_free(AmgDev &A);
// row-dependent spmv-family bytes of one V-cycle per sub-problem (element size esize)
std::vector<int64_t> amg_vcycle_bytes(const AmgDev &A, int esize);

}  // namespace pyhip
