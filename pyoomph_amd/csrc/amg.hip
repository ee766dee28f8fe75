// Smoothed-aggregation multigrid preconditioner on the device (FP64).
//
// Symbolic plan: amg_plan.h (host, once per pattern). For every new matrix (each Newton
// step) the device computes, level by level:
//   d = diag(A_l); lambda_s ~ lambda_max(D^-1 A_l) of every sub-problem s by six power
//   iterations from a fixed oscillating start; w_i = 4 / (3 * 1.1 * lambda_s) / d_i;
//   P = (I - W A_l) P_t; Q = A_l P; A_{l+1} = P^T Q   (fixed summation orders).
// One application z = M^-1 b is a V-cycle: on each level one weighted-Jacobi sweep from
// zero (x = W b), the residual restricted by P^T, the coarse correction prolonged by P,
// one weighted-Jacobi sweep; a sub-problem's coarsest level (the level where it stops
// coarsening, at most kBottomRows rows) is solved exactly with its dense Cholesky factor,
// computed in the set-up, in one workgroup.
// Every operation is row-local or per sub-problem, and every sub-problem has its own
// hierarchy, so its preconditioner is the same alone and in a batch; the V-cycle is a
// symmetric operator (usable inside conjugate gradients).
#include <algorithm>
#include <cstdlib>
#include <cstring>

#include "amg.h"
#include "problem.h"
#include "spmv.h"

namespace pyhip {

constexpr int kPowerIters = 6;
constexpr double kLambdaSafety = 1.1;
constexpr int kCholThreads = 64;

static inline unsigned ablk(int64_t n, int per = 256) { return (unsigned)std::max<int64_t>(1, (n + per - 1) / per); }

// ---- numeric set-up kernels --------------------------------------------------------------

__global__ __launch_bounds__(256) void amg_diag_kernel(int64_t n, const int32_t *__restrict__ diag,
                                                       const double *__restrict__ vals, double *__restrict__ d) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) d[i] = vals[diag[i]];
}

// fixed oscillating start vector of the power iteration (depends on the row's index
// within its sub-problem only)
__global__ __launch_bounds__(256) void amg_power_start_kernel(int64_t n, const int32_t *__restrict__ row_sub,
                                                              const int64_t *__restrict__ sub_off,
                                                              double *__restrict__ v) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const uint32_t k = (uint32_t)(i - sub_off[row_sub[i]]);
    uint32_t h = k * 2654435761u;
    h ^= h >> 15;
    v[i] = ((h & 1u) ? 1.0 : -1.0) * (1.0 + (double)((h >> 8) & 7u) * 0.125);
}

// per sub-problem, with y = A v: out[2s] = (D^-1 y, D^-1 y), out[2s+1] = (v, v); one
// block per sub-problem, strided rows then a fixed tree
__global__ __launch_bounds__(256) void amg_sub_norms_kernel(const int64_t *__restrict__ sub_off,
                                                            const double *__restrict__ d,
                                                            const double *__restrict__ y,
                                                            const double *__restrict__ v,
                                                            double *__restrict__ out) {
    __shared__ double sy[256], sv[256];
    const int s = blockIdx.x, t = threadIdx.x;
    double ay = 0.0, av = 0.0;
    for (int64_t i = sub_off[s] + t; i < sub_off[s + 1]; i += 256) {
        const double yi = y[i] / d[i];
        ay += yi * yi;
        av += v[i] * v[i];
    }
    sy[t] = ay;
    sv[t] = av;
    __syncthreads();
    for (int off = 128; off > 0; off >>= 1) {
        if (t < off) {
            sy[t] += sy[t + off];
            sv[t] += sv[t + off];
        }
        __syncthreads();
    }
    if (t == 0) {
        out[2 * s] = sy[0];
        out[2 * s + 1] = sv[0];
    }
}

// v = D^-1 y / |D^-1 y|_s
__global__ __launch_bounds__(256) void amg_power_scale_kernel(int64_t n, const int32_t *__restrict__ row_sub,
                                                              const double *__restrict__ nrm,
                                                              const double *__restrict__ d,
                                                              const double *__restrict__ y, double *__restrict__ v) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double yy = nrm[2 * row_sub[i]];
    v[i] = yy > 0.0 ? (y[i] / d[i]) / sqrt(yy) : 0.0;
}

// w_i = 4 / (3 * safety * lambda_s) / d_i, lambda_s = sqrt((y,y) / (v,v)) of the last step
__global__ __launch_bounds__(256) void amg_weights_kernel(int64_t n, const int32_t *__restrict__ row_sub,
                                                          const double *__restrict__ nrm,
                                                          const double *__restrict__ d, double *__restrict__ w) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int s = row_sub[i];
    const double vv = nrm[2 * s + 1];
    const double lam = vv > 0.0 ? sqrt(nrm[2 * s] / vv) : 1.0;
    const double om = (lam > 0.0 && isfinite(lam)) ? 4.0 / (3.0 * kLambdaSafety * lam) : 0.0;
    w[i] = (d[i] != 0.0 && isfinite(d[i])) ? om / d[i] : 0.0;
}

// P[q] = [pcol(q) == agg(i)] - w_i * (sum of the A entries listed for q), i = prow(q): one
// thread per entry of P, which also writes the value to its place in P^T (tinv)
__global__ __launch_bounds__(256) void amg_pvals_kernel(int64_t nnzp, const int32_t *__restrict__ prow,
                                                        const int32_t *__restrict__ pcptr,
                                                        const int32_t *__restrict__ pcidx,
                                                        const double *__restrict__ pfix,
                                                        const double *__restrict__ vals,
                                                        const double *__restrict__ w,
                                                        const int32_t *__restrict__ tinv, double *__restrict__ pv,
                                                        double *__restrict__ ptv) {
    const int64_t q = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (q >= nnzp) return;
    const int32_t i = prow[q];
    double s = 0.0;
    for (int32_t c = pcptr[q]; c < pcptr[q + 1]; c++) s += vals[pcidx[c]];
    // P = P_t - W A P_t on aggregates (pfix 1 or 0); the element level's entries have no A
    // entries (s = 0): their fixed weight
    const double v = pfix[q] - w[i] * s;
    pv[q] = v;
    ptv[tinv[q]] = v;
}

// position of column j in the sorted row [b, e) of cols, or -1
__device__ __forceinline__ int32_t find_col(const int32_t *__restrict__ cols, int32_t b, int32_t e, int32_t j) {
    while (b < e) {
        const int32_t m = (b + e) >> 1;
        const int32_t c = cols[m];
        if (c == j) return m;
        if (c < j) b = m + 1;
        else e = m;
    }
    return -1;
}

// Sparse row products with fixed patterns, for Q = A P and A_{l+1} = P^T Q:
//   out[I][J] = sum over the terms t of row I (increasing t) of tval[t] * src[trow[t]][J]
// (Q: terms = the entries k of A's row i, tval = A[i][k], source rows = P's rows col[k];
// Galerkin: terms = the entries of P^T's row I, tval = P[i][I], source rows = Q's rows i).
// Every row pattern is column-sorted, and the output pattern holds every column a term's
// source row has. G lanes own one output row: its columns and sums sit in LDS; the terms
// are fetched G at a time (row, value, source row bounds, one per lane) and walked in
// order, the lanes spreading over the source row's entries, each finding its column in
// the output row by binary search in LDS and adding its product with a fused multiply-add
// (a source row has distinct columns, so no two lanes touch one sum in a step). A sum thus
// receives its terms in increasing t whatever G, alone or in a batch. Output rows longer
// than CAPR take the same sums entry by entry (binary search in the source rows).
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int G, int CAPR>
__global__ __launch_bounds__(256) void amg_rowprod_kernel(int64_t nout, const int32_t *__restrict__ ors,
                                                          const int32_t *__restrict__ ocol, double *__restrict__ oval,
                                                          const int32_t *__restrict__ trs,
                                                          const int32_t *__restrict__ trow,
                                                          const double *__restrict__ tval,
                                                          const int32_t *__restrict__ srs,
                                                          const int32_t *__restrict__ scol,
                                                          const double *__restrict__ sval) {
    static_assert(G == 4 || G == 8 || G == 16 || G == 32 || G == 64, "lanes per row");
    constexpr int NG = 256 / G;
    __shared__ int32_t lc[NG][CAPR];
    __shared__ double la[NG][CAPR];
    const int g = threadIdx.x / G, lane = threadIdx.x % G;
    const int64_t I = (int64_t)blockIdx.x * NG + g;
    if (I >= nout) return;   // a whole group; no workgroup barrier below
    const int32_t o0 = ors[I], R = ors[I + 1] - o0, t0 = trs[I], t1 = trs[I + 1];
    if (R > CAPR) {
        for (int32_t q = o0 + lane; q < o0 + R; q += G) {
            const int32_t J = ocol[q];
            double s = 0.0;
            for (int32_t t = t0; t < t1; t++) {
                const int32_t r = trow[t];
                const int32_t e = find_col(scol, srs[r], srs[r + 1], J);
                if (e >= 0) s = __builtin_fma(tval[t], sval[e], s);
            }
            oval[q] = s;
        }
        return;
    }
    int32_t *mc = lc[g];
    double *ma = la[g];
    for (int32_t k = lane; k < R; k += G) {
        mc[k] = ocol[o0 + k];
        ma[k] = 0.0;
    }
    wave_lds_sync();
    for (int32_t tb = t0; tb < t1; tb += G) {
        const int32_t t = tb + lane;
        int32_t sb = 0, se = 0;
        double tv = 0.0;
        if (t < t1) {
            const int32_t r = trow[t];
            tv = tval[t];
            sb = srs[r];
            se = srs[r + 1];
        }
        const int nt = min(G, t1 - tb);
        for (int u = 0; u < nt; u++) {
            const double p = __shfl(tv, u, G);
            const int32_t b = __shfl(sb, u, G), e = __shfl(se, u, G);
            for (int32_t x = b + lane; x < e; x += G) {
                const int32_t J = scol[x];
                int32_t lo = 0, hi = R;
                while (lo < hi) {
                    const int32_t m = (lo + hi) >> 1;
                    if (mc[m] < J) lo = m + 1;
                    else hi = m;
                }
                if (lo < R && mc[lo] == J) ma[lo] = __builtin_fma(p, sval[x], ma[lo]);
            }
            wave_lds_sync();
        }
    }
    for (int32_t k = lane; k < R; k += G) oval[o0 + k] = ma[k];
}
constexpr int kQCap = 64, kCCap = 128;

// One row-product launch with G lanes per output row, G chosen from the mean source-row
// length (the lanes spread over a source row's entries: the element level's prolongator
// rows hold 1-8 entries, an aggregation level's 10-40). Every G gives the same sums.
template <int CAPR>
static void launch_rowprod(double mean_src, hipStream_t st, int64_t nout, const int32_t *ors, const int32_t *ocol,
                           double *oval, const int32_t *trs, const int32_t *trow, const double *tval,
                           const int32_t *srs, const int32_t *scol, const double *sval) {
#define PYHIP_ROWPROD(G)                                                                                          \
    hipLaunchKernelGGL((amg_rowprod_kernel<G, CAPR>), dim3(ablk(nout, 256 / G)), dim3(256), 0, st, nout, ors, ocol, \
                       oval, trs, trow, tval, srs, scol, sval)
    if constexpr (CAPR <= 64) {   // 4 lanes: 64 rows per workgroup, their LDS rows CAPR long
        if (mean_src <= 3.0) {
            PYHIP_ROWPROD(4);
            return;
        }
    }
    if (mean_src <= 6.0) PYHIP_ROWPROD(8);
    else if (mean_src <= 12.0) PYHIP_ROWPROD(16);
    else PYHIP_ROWPROD(32);
#undef PYHIP_ROWPROD
}

// ---- V-cycle kernels -----------------------------------------------------------------------

// x = W b
// (rows of sub-problems with act[row_sub] == 0 are skipped when act is given)
template <typename T>
__global__ __launch_bounds__(256) void amg_wb_kernel(int64_t n, const T *__restrict__ w,
                                                     const T *__restrict__ b, T *__restrict__ x,
                                                     const int32_t *__restrict__ rsub,
                                                     const int32_t *__restrict__ act) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n || (act && !act[rsub[i]])) return;
    x[i] = w[i] * b[i];
}

// Dense Cholesky factor of the listed sub-problems' matrices (<= kBottomMaxRows rows),
// one workgroup per sub-problem, right-looking, fixed order. A non-positive pivot is
// replaced by 1 and flagged. The workgroup then writes the explicit inverse
// X = L^-T L^-1 (thread c solves L y = e_c, L^T x = y for column c, 32 columns per pass)
// as a symmetric matrix (entry (i, j) from the column max(i, j)), so the coarsest solve is
// one dense symmetric product and the V-cycle stays a symmetric operator.
static_assert(kCholThreads == kBottomMaxRows, "one thread per coarsest row");
__global__ __launch_bounds__(kCholThreads) void amg_chol_factor_kernel(const int32_t *__restrict__ subs,
                                                                       const int64_t *__restrict__ sub_off,
                                                                       const int32_t *__restrict__ rs,
                                                                       const int32_t *__restrict__ col,
                                                                       const double *__restrict__ vals,
                                                                       double *__restrict__ cinv,
                                                                       int32_t *__restrict__ fail) {
    __shared__ double a[kBottomMaxRows][kBottomMaxRows + 1];
    __shared__ double w[kBottomMaxRows][kCholThreads / 2 + 1];
    const int s = subs[blockIdx.x], t = threadIdx.x;
    const int64_t r0 = sub_off[s];
    const int m = (int)(sub_off[s + 1] - r0);
    for (int i = t; i < kBottomMaxRows; i += kCholThreads)
        for (int j = 0; j < kBottomMaxRows; j++) a[i][j] = 0.0;
    __syncthreads();
    for (int i = t; i < m; i += kCholThreads)
        for (int32_t e = rs[r0 + i]; e < rs[r0 + i + 1]; e++) a[i][col[e] - r0] = vals[e];
    __syncthreads();
    for (int k = 0; k < m; k++) {
        if (t == 0) {
            double pv = a[k][k];
            if (!(pv > 0.0) || !isfinite(pv)) {
                fail[s] = 1;
                pv = 1.0;
            }
            a[k][k] = sqrt(pv);
        }
        __syncthreads();
        for (int i = k + 1 + t; i < m; i += kCholThreads) a[i][k] = a[i][k] / a[k][k];
        __syncthreads();
        for (int i = k + 1 + t; i < m; i += kCholThreads)
            for (int j = k + 1; j <= i; j++) a[i][j] = a[i][j] - a[i][k] * a[j][k];
        __syncthreads();
    }
    double *X = cinv + (int64_t)s * kBottomMaxRows * kBottomMaxRows;
    constexpr int H = kCholThreads / 2;
    for (int c0 = 0; c0 < m; c0 += H) {
        const int c = c0 + t;
        if (t < H && c < m) {
            for (int i = 0; i < m; i++) {          // L y = e_c (y_i = 0 for i < c)
                if (i < c) {
                    w[i][t] = 0.0;
                    continue;
                }
                double v = i == c ? 1.0 : 0.0;
                for (int k = c; k < i; k++) v = v - a[i][k] * w[k][t];
                w[i][t] = v / a[i][i];
            }
            for (int i = m - 1; i >= 0; i--) {     // L^T x = y
                double v = w[i][t];
                for (int k = i + 1; k < m; k++) v = v - a[k][i] * w[k][t];
                w[i][t] = v / a[i][i];
            }
        }
        __syncthreads();
        // columns c0 .. c0+H-1: entries (i, c) with i <= c, mirrored
        for (int q = t; q < H * m; q += kCholThreads) {
            const int cc = q / m, i = q % m, c = c0 + cc;
            if (c < m && i <= c) {
                const double v = w[i][cc];
                X[i * kBottomMaxRows + c] = v;
                X[c * kBottomMaxRows + i] = v;
            }
        }
        __syncthreads();
    }
}

// Non-symmetric coarsest matrices: in-place Gauss-Jordan inverse without pivoting (the
// operators are diagonally dominant velocity blocks), thread t owning row t; a zero or
// non-finite pivot is replaced by 1 and flagged. Written in the layout the solve below
// reads: X[j][i] = (A^-1)_{ij}.
__global__ __launch_bounds__(kCholThreads) void amg_gj_inverse_kernel(const int32_t *__restrict__ subs,
                                                                      const int64_t *__restrict__ sub_off,
                                                                      const int32_t *__restrict__ rs,
                                                                      const int32_t *__restrict__ col,
                                                                      const double *__restrict__ vals,
                                                                      double *__restrict__ cinv,
                                                                      int32_t *__restrict__ fail) {
    __shared__ double a[kBottomMaxRows][kBottomMaxRows + 1];
    const int s = subs[blockIdx.x], t = threadIdx.x;
    const int64_t r0 = sub_off[s];
    const int m = (int)(sub_off[s + 1] - r0);
    for (int j = 0; j < kBottomMaxRows; j++) a[t][j] = 0.0;
    __syncthreads();
    if (t < m)
        for (int32_t e = rs[r0 + t]; e < rs[r0 + t + 1]; e++) a[t][col[e] - r0] = vals[e];
    __syncthreads();
    for (int k = 0; k < m; k++) {
        double pv = a[k][k];
        if (!(fabs(pv) > 0.0) || !isfinite(pv)) {
            if (t == 0) fail[s] = 1;
            pv = 1.0;
        }
        const double inv = 1.0 / pv;
        __syncthreads();   // every thread has read the pivot
        if (t < m && t != k) a[k][t] = a[k][t] * inv;
        __syncthreads();   // row k scaled
        if (t < m && t != k) {
            const double f = a[t][k];
            for (int j = 0; j < m; j++)
                if (j != k) a[t][j] = a[t][j] - f * a[k][j];
            a[t][k] = -f * inv;
        }
        if (t == k) a[k][k] = inv;
        __syncthreads();
    }
    double *X = cinv + (int64_t)s * kBottomMaxRows * kBottomMaxRows;
    if (t < m)
        for (int j = 0; j < m; j++) X[j * kBottomMaxRows + t] = a[t][j];
}

// x = X b on the listed sub-problems' coarsest rows (X the symmetric inverse above): thread
// i sums X[j][i] b[j] over j in increasing order (X[j][i] = X[i][j]; the reads of one j are
// contiguous over the threads).
template <typename T>
__global__ __launch_bounds__(kCholThreads) void amg_chol_solve_kernel(const int32_t *__restrict__ subs,
                                                                      const int64_t *__restrict__ sub_off,
                                                                      const T *__restrict__ cinv,
                                                                      const T *b, T *x,
                                                                      const int32_t *__restrict__ act) {
    __shared__ T sb[kBottomMaxRows];
    const int s = subs[blockIdx.x], t = threadIdx.x;
    if (act && !act[s]) return;   // the whole workgroup leaves together
    const int64_t r0 = sub_off[s];
    const int m = (int)(sub_off[s + 1] - r0);
    const T *X = cinv + (int64_t)s * kBottomMaxRows * kBottomMaxRows;
    if (t < m) sb[t] = b[r0 + t];
    __syncthreads();
    if (t < m) {
        T acc = T(0);
        for (int j = 0; j < m; j++) acc = fmadd(X[j * kBottomMaxRows + t], sb[j], acc);
        x[r0 + t] = acc;
    }
}

// FP32 copy of FP64 values (round to nearest)
__global__ __launch_bounds__(256) void amg_to_f32_kernel(int64_t n, const double *__restrict__ a,
                                                         float *__restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

// ---- host ------------------------------------------------------------------------------------

#define ACHECK() PYHIP_CHECK_HIP(hipGetLastError())

template <typename T>
static int aup(T **dst, const std::vector<T> &v, int64_t &bytes) {
    int rc = dev_alloc(dst, v.size(), bytes);
    if (rc) return rc;
    if (!v.empty()) PYHIP_CHECK_HIP(hipMemcpy(*dst, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return PYHIP_OK;
}

static std::vector<int32_t> row_sub_of(const std::vector<int64_t> &off) {
    std::vector<int32_t> rsub(off.back());
    for (size_t s = 0; s + 1 < off.size(); s++)
        for (int64_t i = off[s]; i < off[s + 1]; i++) rsub[i] = (int32_t)s;
    return rsub;
}

void amg_free(AmgDev &A) {
    auto fl = [](AmgDevLevel &L) {
        void *ptrs[] = {L.rs, L.col, L.diag, L.agg, L.pfix, L.prs, L.pcol, L.pcptr, L.pcidx, L.trs, L.trow, L.tsrc,
                        L.qrs, L.qcol, L.row_sub, L.sub_off, L.own_vals, L.d, L.w, L.pv, L.qv, L.b, L.x,
                        L.x2, L.r, L.pw, L.py, L.cont, L.ends, L.ptv, L.prow, L.tinv, L.own_vals32, L.w32,
                        L.pv32, L.ptv32, L.b32, L.x32, L.x2_32, L.r32};
        for (void *q : ptrs) dfree(q);
        L = AmgDevLevel();
    };
    for (auto &L : A.lv) fl(L);
    fl(A.bot);
    dfree(A.nrm);
    dfree(A.chol);
    dfree(A.chol_fail);
    dfree(A.chol32);
    A = AmgDev();
}

static int upload_level(AmgDevLevel &D, const std::vector<int32_t> &rs, const std::vector<int32_t> &col,
                        const std::vector<int32_t> &diag, const std::vector<int64_t> &sub_off, bool own_vals,
                        int64_t &B) {
    int rc;
    D.n = (int64_t)rs.size() - 1;
    D.nnz = rs.back();
    if ((rc = aup(&D.rs, rs, B)) || (rc = aup(&D.col, col, B)) || (rc = aup(&D.diag, diag, B)) ||
        (rc = aup(&D.sub_off, sub_off, B)) || (rc = aup(&D.row_sub, row_sub_of(sub_off), B)))
        return rc;
    const size_t ns = sub_off.size() - 1;
    D.h_rows.assign(ns, 0);
    D.h_nnz.assign(ns, 0);
    D.h_nnzp.assign(ns, 0);
    for (size_t s = 0; s < ns; s++) {
        D.h_rows[s] = sub_off[s + 1] - sub_off[s];
        D.h_nnz[s] = (int64_t)rs[sub_off[s + 1]] - rs[sub_off[s]];
    }
    if (own_vals && (rc = dev_alloc(&D.own_vals, D.nnz, B))) return rc;
    if ((rc = dev_alloc(&D.d, D.n, B)) || (rc = dev_alloc(&D.w, D.n, B)) || (rc = dev_alloc(&D.b, D.n, B)) ||
        (rc = dev_alloc(&D.x, D.n, B)) || (rc = dev_alloc(&D.x2, D.n, B)) || (rc = dev_alloc(&D.r, D.n, B)) ||
        (rc = dev_alloc(&D.pw, D.n, B)) || (rc = dev_alloc(&D.py, D.n, B)))
        return rc;
    return PYHIP_OK;
}

int amg_build(AmgDev &A, int64_t n, const int32_t *rs_h, const int32_t *col_h, const std::vector<int64_t> &sub_off,
              int64_t &B, const AmgFixedInterp *fix) {
    if (A.built) return PYHIP_OK;
    AmgPlan plan;
    std::string err;
    A.element_level = fix != nullptr;
    if (amg_plan_build(n, rs_h, col_h, sub_off, kAggregateCap, kBottomRows, kMaxLevels, plan, err, fix)) {
        set_error(err);
        return PYHIP_ERR_INVALID;
    }
    if (plan.max_bot_rows > kBottomMaxRows) {
        set_error("multigrid: coarsest sub-problem larger than the dense coarsest solve");
        return PYHIP_ERR_INVALID;
    }
    A.nsub = (int)sub_off.size() - 1;
    int rc;
    A.lv.resize(plan.lv.size());
    for (size_t l = 0; l < plan.lv.size(); l++) {
        const AmgLevelPlan &P = plan.lv[l];
        AmgDevLevel &D = A.lv[l];
        if ((rc = upload_level(D, P.rs, P.col, P.diag, P.sub_off, l > 0, B))) return rc;
        D.nc = P.nc;
        D.nends = (int)P.ends.size();
        if ((rc = aup(&D.cont, P.cont, B)) || (rc = aup(&D.ends, P.ends, B))) return rc;
        D.nnzp = P.prs.back();
        for (size_t q = 0; q + 1 < P.sub_off.size(); q++)
            D.h_nnzp[q] = (int64_t)P.prs[P.sub_off[q + 1]] - P.prs[P.sub_off[q]];
        D.nnzq = P.qrs.back();
        // row of every P entry, and the place of every P entry in P^T
        std::vector<int32_t> prow(D.nnzp), tinv(D.nnzp);
        for (int64_t i = 0; i + 1 < (int64_t)P.prs.size(); i++)
            for (int32_t q = P.prs[i]; q < P.prs[i + 1]; q++) prow[q] = (int32_t)i;
        for (int64_t t = 0; t < D.nnzp; t++) tinv[P.tsrc[t]] = (int32_t)t;
        if ((rc = aup(&D.prow, prow, B)) || (rc = aup(&D.tinv, tinv, B))) return rc;
        if ((rc = aup(&D.pfix, P.pfix, B)) || (rc = aup(&D.prs, P.prs, B)) || (rc = aup(&D.pcol, P.pcol, B)) ||
            (rc = aup(&D.pcptr, P.pcptr, B)) || (rc = aup(&D.pcidx, P.pcidx, B)) || (rc = aup(&D.trs, P.trs, B)) ||
            (rc = aup(&D.trow, P.trow, B)) || (rc = aup(&D.tsrc, P.tsrc, B)) || (rc = aup(&D.qrs, P.qrs, B)) ||
            (rc = aup(&D.qcol, P.qcol, B)) || (rc = dev_alloc(&D.pv, D.nnzp, B)) ||
            (rc = dev_alloc(&D.ptv, D.nnzp, B)) ||
            (rc = dev_alloc(&D.qv, D.nnzq, B)))
            return rc;
    }
    if ((rc = upload_level(A.bot, plan.bot_rs, plan.bot_col, plan.bot_diag, plan.bot_off, !plan.lv.empty(), B)))
        return rc;
    A.bot.nends = (int)plan.bot_subs.size();
    if ((rc = aup(&A.bot.ends, plan.bot_subs, B))) return rc;
    if ((rc = dev_alloc(&A.nrm, 2 * std::max(1, A.nsub), B)) ||
        (rc = dev_alloc(&A.chol, (size_t)std::max(1, A.nsub) * kBottomMaxRows * kBottomMaxRows, B)) ||
        (rc = dev_alloc(&A.chol_fail, std::max(1, A.nsub), B)))
        return rc;
    A.max_bot_rows = plan.max_bot_rows;
    A.built = true;
    return PYHIP_OK;
}

// lanes per row of the canonical spmv for a matrix with this mean row length
static inline int lanes_for(int64_t nnz, int64_t n) { return (n > 0 && nnz <= 12 * n) ? 4 : 8; }

// One level spmv, timed as an "spmv" launch with its algorithmic bytes (values and column
// indices, row starts, the input entries and every row-indexed vector read or written, once
// each). Modes: y = A x (PLAIN), y = b - A x (SUB), y = b + A x (ADD), the weighted-Jacobi
// step out = x + W (b - A x) (JACOBI), the restriction y = A x with the next level's
// out = W y (SCALE).
enum { kSpmvPlain, kSpmvSub, kSpmvAdd, kSpmvJacobi, kSpmvScale };
// Rows filtered by (rsub, act): the bytes count the rows, entries and input entries of the
// visited sub-problems (nnz, rows, ncols are those counts).
template <typename T>
static void level_spmv(KernelTimer *tm, int mode, int lpr, hipStream_t st, int64_t n, int64_t nnz, int64_t rows,
                       int64_t ncols, const int32_t *rs, const int32_t *col, const T *vals, const T *x,
                       T *y, const T *b, const int32_t *rsub = nullptr, const int32_t *act = nullptr,
                       const T *w = nullptr, T *out = nullptr) {
    TimeScope ts(tm, "spmv", st);
    constexpr int64_t E = sizeof(T);
    int64_t vec = E * rows;   // y
    switch (mode) {
    case kSpmvAdd:
        launch_spmv_epi<T, T, T, kEpiAdd>(lpr, st, n, rs, col, vals, x, y, b, w, out, rsub, act);
        vec = 2 * E * rows;
        break;
    case kSpmvJacobi:
        launch_spmv_epi<T, T, T, kEpiJacobi>(lpr, st, n, rs, col, vals, x, y, b, w, out, rsub, act);
        vec = 4 * E * rows;   // b, w, own x, out
        break;
    case kSpmvScale:
        launch_spmv_epi<T, T, T, kEpiScale>(lpr, st, n, rs, col, vals, x, y, b, w, out, rsub, act);
        vec = 3 * E * rows;   // y, w, out
        break;
    default:
        launch_spmv_epi<T, T, T, kEpiPlain>(lpr, st, n, rs, col, vals, x, y, mode == kSpmvSub ? b : nullptr, w, out,
                                            rsub, act);
        if (mode == kSpmvSub) vec = 2 * E * rows;
    }
    if (tm) tm->add_work("spmv", (E + 4) * nnz + 4 * (rows + 1) + E * ncols + vec, 4);
}

// sum of v[s] over the sub-problems with act_h[s] != 0 (all without act_h)
static int64_t active_sum(const std::vector<int64_t> &v, const int32_t *act_h) {
    int64_t t = 0;
    for (size_t s = 0; s < v.size(); s++)
        if (!act_h || act_h[s]) t += v[s];
    return t;
}


// Variants of each level's three spmv launches: the fastest of the canonical spmv's
// variants (1, 2, 4, 8 lanes per row, LDS staging) over two timed launches each, once per
// hierarchy (every variant gives the same bits, so the choice changes no result), or
// PYHIP_SPMV_LPR for all of them. The outputs are scratch vectors of the level.
static int tune_levels(AmgDev &A, hipStream_t st) {
    if (const char *e = getenv("PYHIP_SPMV_LPR")) {
        const int v = atoi(e);
        if (spmv_variant_ok(v)) {
            for (auto &L : A.lv) L.lpr_a = L.lpr_t = L.lpr_p = v;
            return PYHIP_OK;
        }
    }
    hipEvent_t e0, e1;
    PYHIP_CHECK_HIP(hipEventCreate(&e0));
    PYHIP_CHECK_HIP(hipEventCreate(&e1));
    auto pick = [&](auto &&launch) {
        float best = 1e30f;
        int arg = 8;
        for (int lpr : kSpmvVariants) {
            launch(lpr);
            (void)hipEventRecord(e0, st);
            for (int k = 0; k < 2; k++) launch(lpr);
            (void)hipEventRecord(e1, st);
            float ms = 0.f;
            if (hipEventSynchronize(e1) == hipSuccess && hipEventElapsedTime(&ms, e0, e1) == hipSuccess &&
                ms < best) {
                best = ms;
                arg = lpr;
            }
        }
        return arg;
    };
    for (size_t l = 0; l < A.lv.size(); l++) {
        AmgDevLevel &L = A.lv[l];
        AmgDevLevel &C = (l + 1 < A.lv.size()) ? A.lv[l + 1] : A.bot;
        L.lpr_a = pick([&](int lpr) {
            launch_spmv<double, double, double>(lpr, st, L.n, L.rs, L.col, L.vals, L.b, L.r, nullptr);
        });
        if (L.nc) {
            L.lpr_t = pick([&](int lpr) {
                launch_spmv<double, double, double>(lpr, st, L.nc, L.trs, L.trow, L.ptv, L.r, C.b, nullptr);
            });
            L.lpr_p = pick([&](int lpr) {
                launch_spmv_add<double, double, double>(lpr, st, L.n, L.prs, L.pcol, L.pv, C.x, L.x2, L.x2);
            });
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    ACHECK();
    return PYHIP_OK;
}

static int level_weights(hipStream_t st, AmgDevLevel &L, const double *vals, double *nrm, int nsub,
                         KernelTimer *tm) {
    const int64_t n = L.n;
    const int lpr = L.lpr_a ? L.lpr_a : lanes_for(L.nnz, n);
    hipLaunchKernelGGL(amg_diag_kernel, dim3(ablk(n)), dim3(256), 0, st, n, L.diag, vals, L.d);
    hipLaunchKernelGGL(amg_power_start_kernel, dim3(ablk(n)), dim3(256), 0, st, n, L.row_sub, L.sub_off, L.pw);
    for (int it = 0; it < kPowerIters; it++) {
        level_spmv<double>(tm, kSpmvPlain, lpr, st, n, L.nnz, n, n, L.rs, L.col, vals, L.pw, L.py, nullptr);
        hipLaunchKernelGGL(amg_sub_norms_kernel, dim3(nsub), dim3(256), 0, st, L.sub_off, L.d, L.py, L.pw, nrm);
        if (it + 1 < kPowerIters)
            hipLaunchKernelGGL(amg_power_scale_kernel, dim3(ablk(n)), dim3(256), 0, st, n, L.row_sub, nrm, L.d, L.py,
                               L.pw);
    }
    hipLaunchKernelGGL(amg_weights_kernel, dim3(ablk(n)), dim3(256), 0, st, n, L.row_sub, nrm, L.d, L.w);
    ACHECK();
    return PYHIP_OK;
}

int amg_setup(AmgDev &A, hipStream_t st, const double *vals0, KernelTimer *tm, const int32_t *need_h) {
    if (!A.built) {
        set_error("multigrid plan not built");
        return PYHIP_ERR_INVALID;
    }
    PYHIP_CHECK_HIP(hipMemsetAsync(A.chol_fail, 0, std::max(1, A.nsub) * sizeof(int32_t), st));
    const double *vals = vals0;
    int rc;
    for (size_t l = 0; l < A.lv.size(); l++) {
        AmgDevLevel &L = A.lv[l];
        if (l > 0) vals = L.own_vals;
        L.vals = vals;
        if ((rc = level_weights(st, L, vals, A.nrm, A.nsub, tm))) return rc;
        if (L.nends)
            hipLaunchKernelGGL(A.nonsym ? amg_gj_inverse_kernel : amg_chol_factor_kernel, dim3(L.nends),
                               dim3(kCholThreads), 0, st, L.ends, L.sub_off, L.rs, L.col, vals, A.chol, A.chol_fail);
        if (L.nnzp > 0)
            hipLaunchKernelGGL(amg_pvals_kernel, dim3(ablk(L.nnzp)), dim3(256), 0, st, L.nnzp, L.prow,
                               L.pcptr, L.pcidx, L.pfix, vals, L.w, L.tinv, L.pv, L.ptv);
        // Q = A_l P (terms: A's row entries; source rows: P's rows)
        launch_rowprod<kQCap>(L.n ? (double)L.nnzp / L.n : 0.0, st, L.n, L.qrs, L.qcol, L.qv, L.rs, L.col, vals,
                              L.prs, L.pcol, L.pv);
        AmgDevLevel &C = (l + 1 < A.lv.size()) ? A.lv[l + 1] : A.bot;
        // A_{l+1} = P^T Q (terms: P^T's row entries; source rows: Q's rows)
        if (L.nc)
            launch_rowprod<kCCap>(L.n ? (double)L.nnzq / L.n : 0.0, st, L.nc, C.rs, C.col, C.own_vals, L.trs, L.trow,
                                  L.ptv, L.qrs, L.qcol, L.qv);
        ACHECK();
    }
    A.bot.vals = A.lv.empty() ? vals0 : A.bot.own_vals;
    if (A.bot.nends)
        hipLaunchKernelGGL(A.nonsym ? amg_gj_inverse_kernel : amg_chol_factor_kernel, dim3(A.bot.nends),
                           dim3(kCholThreads), 0, st, A.bot.ends, A.bot.sub_off, A.bot.rs, A.bot.col, A.bot.vals,
                           A.chol, A.chol_fail);
    ACHECK();
    // A coarsest matrix that is not positive definite (an indefinite Jacobian, e.g. past a
    // fold) makes the V-cycle an invalid preconditioner: report it instead of iterating on.
    std::vector<int32_t> fail_h(std::max(1, A.nsub), 0);
    PYHIP_CHECK_HIP(hipMemcpyAsync(fail_h.data(), A.chol_fail, fail_h.size() * sizeof(int32_t),
                                   hipMemcpyDeviceToHost, st));
    PYHIP_CHECK_HIP(hipStreamSynchronize(st));
    for (int s = 0; s < A.nsub; s++)
        if (fail_h[s] && (!need_h || need_h[s])) {
            set_error("multigrid: the coarsest matrix of sub-problem " + std::to_string(s) +
                      (A.nonsym ? std::string(" has a zero pivot") :
                                  std::string(" is not positive definite (indefinite Jacobian, e.g. past a fold); use "
                                              "the ILU(0) or Jacobi preconditioner there")));
            return PYHIP_ERR_BREAKDOWN;
        }
    if (!A.lv.empty() && !A.lv[0].lpr_a && (rc = tune_levels(A, st))) return rc;
    return PYHIP_OK;
}

int amg_set_fine(AmgDev &A, const double *vals) {
    if (!A.built) {
        set_error("multigrid plan not built");
        return PYHIP_ERR_INVALID;
    }
    if (A.lv.empty()) A.bot.vals = vals;
    else A.lv[0].vals = vals;
    return PYHIP_OK;
}

// Per-precision views of a level's values and work vectors (the FP64 hierarchy, or its
// FP32 copy made by amg_make_f32).
template <typename T> struct AmgView;
template <> struct AmgView<double> {
    static const double *vals(const AmgDevLevel &L) { return L.vals; }
    static const double *w(const AmgDevLevel &L) { return L.w; }
    static const double *pv(const AmgDevLevel &L) { return L.pv; }
    static const double *ptv(const AmgDevLevel &L) { return L.ptv; }
    static double *b(AmgDevLevel &L) { return L.b; }
    static double *&x(AmgDevLevel &L) { return L.x; }
    static double *&x2(AmgDevLevel &L) { return L.x2; }
    static double *r(AmgDevLevel &L) { return L.r; }
    static const double *chol(const AmgDev &A) { return A.chol; }
};
template <> struct AmgView<float> {
    static const float *vals(const AmgDevLevel &L) { return L.vals32; }
    static const float *w(const AmgDevLevel &L) { return L.w32; }
    static const float *pv(const AmgDevLevel &L) { return L.pv32; }
    static const float *ptv(const AmgDevLevel &L) { return L.ptv32; }
    static float *b(AmgDevLevel &L) { return L.b32; }
    static float *&x(AmgDevLevel &L) { return L.x32; }
    static float *&x2(AmgDevLevel &L) { return L.x2_32; }
    static float *r(AmgDevLevel &L) { return L.r32; }
    static const float *chol(const AmgDev &A) { return A.chol32; }
};

// z = M^-1 b (both [n0]) in precision T
template <typename T>
static int amg_apply_t(AmgDev &A, hipStream_t st, const T *b, T *z, KernelTimer *tm, const int32_t *act,
                       const int32_t *act_h, bool x0_ready) {
    using V = AmgView<T>;
    const int nl = (int)A.lv.size();
    AmgDevLevel &Bt = A.bot;
    if (nl == 0) {
        if (Bt.nends)
            hipLaunchKernelGGL((amg_chol_solve_kernel<T>), dim3(Bt.nends), dim3(kCholThreads), 0, st, Bt.ends,
                               Bt.sub_off, V::chol(A), b, z, act);
        ACHECK();
        return PYHIP_OK;
    }
    // visited rows, entries and P entries per level (all sub-problems without a filter)
    std::vector<int64_t> rows(nl + 1), nnz(nl), nnzp(nl);
    for (int l = 0; l < nl; l++) {
        rows[l] = act ? active_sum(A.lv[l].h_rows, act_h) : A.lv[l].n;
        nnz[l] = act ? active_sum(A.lv[l].h_nnz, act_h) : A.lv[l].nnz;
        nnzp[l] = act ? active_sum(A.lv[l].h_nnzp, act_h) : A.lv[l].nnzp;
    }
    rows[nl] = act ? active_sum(Bt.h_rows, act_h) : Bt.n;
    // down: level l works on right-hand side bl (level 0: b). The rows of sub-problems that
    // end on level l are computed too (their couplings stay in their sub-problem) and are
    // then overwritten by their coarsest solve.
    for (int l = 0; l < nl; l++) {
        AmgDevLevel &L = A.lv[l];
        const T *bl = l == 0 ? b : V::b(L);
        AmgDevLevel &C = (l + 1 < nl) ? A.lv[l + 1] : A.bot;
        // x = W b (on coarser levels the restriction above has written it)
        if (l == 0 && !x0_ready)
            hipLaunchKernelGGL((amg_wb_kernel<T>), dim3(ablk(L.n)), dim3(256), 0, st, L.n, V::w(L), bl, V::x(L),
                               L.row_sub, act);
        const int la = L.lpr_a ? L.lpr_a : lanes_for(L.nnz, L.n);
        level_spmv<T>(tm, kSpmvSub, la, st, L.n, nnz[l], rows[l], rows[l], L.rs, L.col, V::vals(L), V::x(L),
                      V::r(L), bl, L.row_sub, act);
        // b_c = P^T r, and x_c = W_c b_c where the coarse level is not the coarsest
        if (L.nc && l + 1 < nl)
            level_spmv<T>(tm, kSpmvScale, L.lpr_t ? L.lpr_t : 8, st, L.nc, nnzp[l], rows[l + 1], rows[l], L.trs,
                          L.trow, V::ptv(L), V::r(L), V::b(C), nullptr, C.row_sub, act, V::w(C), V::x(C));
        else if (L.nc)
            level_spmv<T>(tm, kSpmvPlain, L.lpr_t ? L.lpr_t : 8, st, L.nc, nnzp[l], rows[l + 1], rows[l], L.trs,
                          L.trow, V::ptv(L), V::r(L), V::b(C), nullptr, C.row_sub, act);
    }
    if (Bt.nends)
        hipLaunchKernelGGL((amg_chol_solve_kernel<T>), dim3(Bt.nends), dim3(kCholThreads), 0, st, Bt.ends,
                           Bt.sub_off, V::chol(A), V::b(Bt), V::x(Bt), act);
    // up: coarse correction and post-sweep; sub-problems ending on a level get their
    // coarsest solve there
    for (int l = nl - 1; l >= 0; l--) {
        AmgDevLevel &L = A.lv[l];
        const T *bl = l == 0 ? b : V::b(L);
        AmgDevLevel &C = (l + 1 < nl) ? A.lv[l + 1] : A.bot;
        if (L.nc)
            level_spmv<T>(tm, kSpmvAdd, L.lpr_p ? L.lpr_p : 4, st, L.n, nnzp[l], rows[l], rows[l + 1], L.prs, L.pcol,
                          V::pv(L), V::x(C), V::x(L), V::x(L), L.row_sub, act);
        // post-sweep out = x + W (b - A x) in the spmv's epilogue
        T *out = l == 0 ? z : V::x2(L);
        level_spmv<T>(tm, kSpmvJacobi, L.lpr_a ? L.lpr_a : lanes_for(L.nnz, L.n), st, L.n, nnz[l], rows[l], rows[l],
                      L.rs, L.col, V::vals(L), V::x(L), nullptr, bl, L.row_sub, act, V::w(L), out);
        if (l > 0) std::swap(V::x(L), V::x2(L));   // the level's result is in x for the prolongation above
        if (L.nends)
            hipLaunchKernelGGL((amg_chol_solve_kernel<T>), dim3(L.nends), dim3(kCholThreads), 0, st, L.ends,
                               L.sub_off, V::chol(A), bl, l == 0 ? z : V::x(L), act);
    }
    ACHECK();
    return PYHIP_OK;
}

// Row-dependent spmv-family bytes one V-cycle moves for each sub-problem (the terms
// level_spmv counts, less the 4 bytes per launch), element size esize.
std::vector<int64_t> amg_vcycle_bytes(const AmgDev &A, int esize) {
    std::vector<int64_t> out(std::max(0, A.nsub), 0);
    const int nl = (int)A.lv.size();
    const int64_t E = esize;
    for (int s = 0; s < A.nsub; s++) {
        int64_t b = 0;
        for (int l = 0; l < nl; l++) {
            const AmgDevLevel &L = A.lv[l];
            const AmgDevLevel &C = (l + 1 < nl) ? A.lv[l + 1] : A.bot;
            const int64_t r = L.h_rows[s], z = L.h_nnz[s], zp = L.h_nnzp[s], rc = C.h_rows[s];
            b += (E + 4) * z + 4 * r + E * r + 2 * E * r;             // r = b - A x
            if (L.nc) b += (E + 4) * zp + 4 * rc + E * r + (l + 1 < nl ? 3 : 1) * E * rc;   // restriction
            if (L.nc) b += (E + 4) * zp + 4 * r + E * rc + 2 * E * r; // prolongation x += P x_c
            b += (E + 4) * z + 4 * r + E * r + 4 * E * r;             // post-sweep
        }
        out[s] = b;
    }
    return out;
}

int amg_apply(AmgDev &A, hipStream_t st, const double *b, double *z, KernelTimer *tm, const int32_t *act,
              const int32_t *act_h, bool x0_ready) {
    return amg_apply_t<double>(A, st, b, z, tm, act, act_h, x0_ready);
}

int amg_make_f32(AmgDev &A, hipStream_t st, int64_t &B) {
    if (!A.built) {
        set_error("multigrid plan not built");
        return PYHIP_ERR_INVALID;
    }
    int rc;
    auto conv = [&](const double *src, float *&dst, int64_t n) -> int {
        if (!dst && (rc = dev_alloc(&dst, n, B))) return rc;
        if (n > 0) hipLaunchKernelGGL(amg_to_f32_kernel, dim3(ablk(n)), dim3(256), 0, st, n, src, dst);
        return PYHIP_OK;
    };
    auto vecs = [&](AmgDevLevel &L) -> int {
        if (!L.b32 && ((rc = dev_alloc(&L.b32, L.n, B)) || (rc = dev_alloc(&L.x32, L.n, B)) ||
                       (rc = dev_alloc(&L.x2_32, L.n, B)) || (rc = dev_alloc(&L.r32, L.n, B))))
            return rc;
        return PYHIP_OK;
    };
    for (size_t l = 0; l < A.lv.size(); l++) {
        AmgDevLevel &L = A.lv[l];
        if (l > 0) {
            if ((rc = conv(L.own_vals, L.own_vals32, L.nnz))) return rc;
            L.vals32 = L.own_vals32;
        }
        if ((rc = conv(L.w, L.w32, L.n)) || (rc = conv(L.pv, L.pv32, L.nnzp)) || (rc = conv(L.ptv, L.ptv32, L.nnzp)) ||
            (rc = vecs(L)))
            return rc;
    }
    if (!A.lv.empty()) {
        if ((rc = conv(A.bot.own_vals, A.bot.own_vals32, A.bot.nnz))) return rc;
        A.bot.vals32 = A.bot.own_vals32;
    }
    if ((rc = vecs(A.bot))) return rc;
    if ((rc = conv(A.chol, A.chol32, (int64_t)std::max(1, A.nsub) * kBottomMaxRows * kBottomMaxRows))) return rc;
    ACHECK();
    return PYHIP_OK;
}

int amg_set_fine32(AmgDev &A, const float *vals) {
    if (!A.built) {
        set_error("multigrid plan not built");
        return PYHIP_ERR_INVALID;
    }
    if (A.lv.empty()) A.bot.vals32 = vals;
    else A.lv[0].vals32 = vals;
    return PYHIP_OK;
}

int amg_apply32(AmgDev &A, hipStream_t st, const float *b, float *z, KernelTimer *tm, const int32_t *act,
                const int32_t *act_h) {
    return amg_apply_t<float>(A, st, b, z, tm, act, act_h, false);
}

}  // namespace pyhip
