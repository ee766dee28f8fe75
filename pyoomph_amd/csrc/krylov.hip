// Batched preconditioned GMRES on the device.
//
// Every vector spans all sub-problems of a problem (block-diagonal system). Every scalar
// (dot products, norms, Hessenberg entries, Givens rotations, convergence state) is per
// sub-problem. Reductions are two-level and deterministic: fixed 1024-row chunks that
// never straddle a sub-problem, a fixed-tree block reduction per chunk, then a
// sequential sum of a sub-problem's chunk partials in chunk order. A sub-problem's
// arithmetic therefore depends only on its own data: it gives the same bits alone and in
// any batch.
//
// GMRES(m) with right preconditioning and classical Gram-Schmidt with one
// re-orthogonalisation (CGS2); Hessenberg least squares by Givens rotations on the
// device (one thread per sub-problem). The host reads one counter per iteration (the
// number of sub-problems still iterating).
//
// Preconditioners: Jacobi, and ILU(0) on the structural pattern, level-scheduled
// (rows of one level are independent; a level is one launch; levels of all
// sub-problems are merged by level index). Taylor-Hood sub-problems are factorised with
// the velocity unknowns ordered before the pressure unknowns (natural order within
// each), which keeps ILU(0) pivots non-zero for the zero pressure-pressure block.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "ilu_plan.h"
#include "krylov.h"
#include "spmv.h"

namespace pyhip {

// ------------------------------------------------------------------------------------
// kernels

// partial[c][j] = sum over chunk c of V_j . w, for the vectors j of block column
// blockIdx.y (kDotVecs of them): wave q of the block takes j = q, q + 4, ...; its lanes
// stride the chunk's rows, and the wave sums them with a fixed xor-shuffle tree.
constexpr int kDotVecs = 16;
template <typename T>
__global__ __launch_bounds__(256) void multidot_kernel(int nvec, const T *__restrict__ V,
                                                      int64_t ldv, const T *__restrict__ w,
                                                      const int64_t *__restrict__ cb,
                                                      const int64_t *__restrict__ ce,
                                                      double *__restrict__ partial, int pstride,
                                                      const int32_t *__restrict__ csub,
                                                      const int32_t *__restrict__ act) {
    const int64_t c = blockIdx.x;
    // chunks of sub-problems that stopped (act[csub] == 0) are skipped: their partials are
    // left as they are and nothing reads their sums
    if (act && !act[csub[c]]) return;
    const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
    const int64_t b0 = cb[c], b1 = ce[c];
    for (int jj = q; jj < kDotVecs; jj += 4) {
        const int j = blockIdx.y * kDotVecs + jj;
        if (j >= nvec) break;
        const T *v = V + (int64_t)j * ldv;
        double acc = 0.0;
#pragma unroll 4
        for (int64_t i = b0 + lane; i < b1; i += 64) acc += double(v[i]) * double(w[i]);
        for (int off = 32; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
        if (lane == 0) partial[c * pstride + j] = acc;
    }
}

// max |v| per chunk (non-finite -> +inf), DoubleVector::max (double_vector.cc:611-630)
__global__ __launch_bounds__(256) void maxabs_kernel(const double *__restrict__ v,
                                                    const int64_t *__restrict__ cb,
                                                    const int64_t *__restrict__ ce,
                                                    double *__restrict__ partial, int pstride) {
    __shared__ double red[256];
    const int64_t c = blockIdx.x;
    double m = 0.0;
    for (int64_t i = cb[c] + threadIdx.x; i < ce[c]; i += 256) {
        const double a = v[i];
        if (!isfinite(a)) m = INFINITY;
        else m = fmax(m, fabs(a));
    }
    red[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[c * pstride] = red[0];
}

// out[j][s] = sum (or max) of the chunk partials of sub-problem s, vector j: one 64-lane
// block per (j, s); lane t takes chunks t, t+64, ... in order, then a fixed xor tree.
__global__ __launch_bounds__(64) void reduce_partials_kernel(int nvec, int nsub,
                                                            const int32_t *__restrict__ scp,
                                                            const double *__restrict__ partial,
                                                            int pstride, double *__restrict__ out,
                                                            int is_max) {
    const int j = blockIdx.x / nsub, s = blockIdx.x % nsub;
    double acc = 0.0;
    for (int c = scp[s] + threadIdx.x; c < scp[s + 1]; c += 64) {
        const double v = partial[(int64_t)c * pstride + j];
        acc = is_max ? fmax(acc, v) : acc + v;
    }
    for (int off = 32; off > 0; off >>= 1) {
        const double o = __shfl_xor(acc, off, 64);
        acc = is_max ? fmax(acc, o) : acc + o;
    }
    if (threadIdx.x == 0) out[j * nsub + s] = acc;
}

// w -= sum_j h[j][sub] V_j  (h = dots of this Gram-Schmidt pass)
template <typename T>
__global__ __launch_bounds__(256) void cgs_update_kernel(int64_t n, int nvec, int nsub,
                                                         const T *__restrict__ V, int64_t ldv,
                                                         const double *__restrict__ h,
                                                         const int32_t *__restrict__ dsub,
                                                         const int32_t *__restrict__ active,
                                                         T *__restrict__ w) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int s = dsub[i];
    if (!active[s]) return;
    double acc = double(w[i]);
    for (int j = 0; j < nvec; j++) acc -= h[j * nsub + s] * double(V[(int64_t)j * ldv + i]);
    w[i] = T(acc);
}

// out = in * (1 / nrm[sub]) (0 where nrm == 0); only active sub-problems
template <typename T>
__global__ __launch_bounds__(256) void scale_kernel(int64_t n, const T *in,
                                                    const double *__restrict__ nrm2,
                                                    const int32_t *__restrict__ dsub,
                                                    const int32_t *__restrict__ active,
                                                    T *out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int s = dsub[i];
    if (!active[s]) return;
    const double nr = sqrt(nrm2[s]);
    out[i] = nr > 0.0 ? T(double(in[i]) / nr) : T(0);
}

__global__ void sqrt_copy_kernel(int n, const double *__restrict__ a, double *__restrict__ b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = sqrt(a[i]);
}

// Start of a cycle: beta = |r|, g = (beta, 0, ...), active if the sub-problem still has
// work (need[s]) and beta > tol * |b|.
// Start of a restart cycle. With stag > 0 (inner solves of iterative refinement), a
// sub-problem whose last cycle reduced |r| by less than the factor stag stops: its
// precision floor is reached and the outer refinement takes over.
__global__ void gmres_start_kernel(int nsub, int m, const double *__restrict__ nrm2,
                                   const double *__restrict__ bnorm, double tol, double stag, int first,
                                   const int32_t *__restrict__ need, double *__restrict__ g,
                                   int32_t *__restrict__ active, int32_t *__restrict__ kcols,
                                   double *__restrict__ beta0, int32_t *__restrict__ nactive) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsub) return;
    const double beta = sqrt(nrm2[s]);
    const bool stalled = !first && stag > 0.0 && !(beta < stag * beta0[s]);
    kcols[s] = 0;
    beta0[s] = beta;
    for (int i = 0; i <= m; i++) g[s * (m + 1) + i] = 0.0;
    g[s * (m + 1)] = beta;
    const int a = need[s] && (beta > tol * bnorm[s]) && isfinite(beta) && !stalled;
    active[s] = a;
    if (a) atomicAdd(nactive, 1);
}

// Column k of the Hessenberg: h[0..k] = dots (CGS pass 1 + pass 2), h[k+1] = |w|.
// m: stride of the per-sub-problem Hessenberg / Givens arrays (the workspace's restart
// capacity); mr <= m: the restart length of this solve (the cycle ends after column mr-1).
__global__ void gmres_givens_kernel(int nsub, int m, int mr, int k, const double *__restrict__ d1,
                                    const double *__restrict__ d2, const double *__restrict__ wn2,
                                    double *__restrict__ H, double *__restrict__ cs,
                                    double *__restrict__ sn, double *__restrict__ g,
                                    const double *__restrict__ bnorm, double tol,
                                    int32_t *__restrict__ active, int32_t *__restrict__ kcols,
                                    int32_t *__restrict__ nactive) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsub || !active[s]) return;
    double *Hs = H + (int64_t)s * (m + 1) * m;
    double *cs_s = cs + (int64_t)s * m, *sn_s = sn + (int64_t)s * m, *gs = g + (int64_t)s * (m + 1);
    // column k is rotated in place: Hs[i * m + k], i = 0..k (row k + 1 is only carried in
    // a register, since the rotation zeroes it)
    double hi = d1[s] + d2[s];
    for (int i = 0; i < k; i++) {
        const double hn = d1[(i + 1) * nsub + s] + d2[(i + 1) * nsub + s];
        Hs[i * m + k] = cs_s[i] * hi + sn_s[i] * hn;
        hi = -sn_s[i] * hi + cs_s[i] * hn;
    }
    const double hk1 = sqrt(wn2[s]);
    const double den = hypot(hi, hk1);
    double c = 1.0, sv = 0.0;
    if (den > 0.0) {
        c = hi / den;
        sv = hk1 / den;
    }
    cs_s[k] = c;
    sn_s[k] = sv;
    Hs[k * m + k] = den;
    gs[k + 1] = -sv * gs[k];
    gs[k] = c * gs[k];
    kcols[s] = k + 1;
    const bool done = !(fabs(gs[k + 1]) > tol * bnorm[s]) || hk1 == 0.0 || den == 0.0 ||
                      !isfinite(gs[k + 1]) || (k + 1 == mr);
    if (done) active[s] = 0;
    else atomicAdd(nactive, 1);
}

// y = R^{-1} g (upper triangular back substitution of the rotated Hessenberg)
__global__ void gmres_backsolve_kernel(int nsub, int m, const double *__restrict__ H,
                                       const double *__restrict__ g,
                                       const int32_t *__restrict__ kcols,
                                       double *__restrict__ y) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsub) return;
    const double *Hs = H + (int64_t)s * (m + 1) * m;
    const double *gs = g + (int64_t)s * (m + 1);
    double *ys = y + (int64_t)s * m;
    const int k = kcols[s];
    for (int i = k - 1; i >= 0; i--) {
        double acc = gs[i];
        for (int j = i + 1; j < k; j++) acc -= Hs[i * m + j] * ys[j];
        const double d = Hs[i * m + i];
        ys[i] = d != 0.0 ? acc / d : 0.0;
    }
    for (int i = k; i < m; i++) ys[i] = 0.0;
}

// t = sum_{j < kcols[sub]} y_j V_j
template <typename T>
__global__ __launch_bounds__(256) void combine_kernel(int64_t n, int m, const T *__restrict__ V,
                                                     int64_t ldv, const double *__restrict__ y,
                                                     const int32_t *__restrict__ kcols,
                                                     const int32_t *__restrict__ dsub,
                                                     T *__restrict__ t) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int s = dsub[i];
    const int k = kcols[s];
    double acc = 0.0;
    for (int j = 0; j < k; j++) acc += y[(int64_t)s * m + j] * double(V[(int64_t)j * ldv + i]);
    t[i] = T(acc);
}

template <typename TA, typename TB>
__global__ __launch_bounds__(256) void axpy_kernel(int64_t n, const TA *__restrict__ a,
                                                   TB *__restrict__ b, double alpha,
                                                   const int32_t *__restrict__ dsub,
                                                   const int32_t *__restrict__ mask) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    if (mask && !mask[dsub[i]]) return;
    b[i] = TB(double(b[i]) + alpha * double(a[i]));
}

template <typename TA, typename TB>
__global__ __launch_bounds__(256) void convert_kernel(int64_t n, const TA *__restrict__ a,
                                                      TB *__restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = TB(a[i]);
}

// v[i] = mask[sub] ? v : 0
template <typename T>
__global__ __launch_bounds__(256) void mask_kernel(int64_t n, T *__restrict__ v,
                                                   const int32_t *__restrict__ dsub,
                                                   const int32_t *__restrict__ mask) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n && !mask[dsub[i]]) v[i] = T(0);
}

template <typename TV, typename T>
__global__ __launch_bounds__(256) void jacobi_setup_kernel(int64_t n, const TV *__restrict__ vals,
                                                           const int32_t *__restrict__ diag,
                                                           T *__restrict__ jd) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const double d = diag[i] >= 0 ? double(vals[diag[i]]) : 0.0;
    jd[i] = T(d != 0.0 ? 1.0 / d : 1.0);
}

template <typename T>
__global__ __launch_bounds__(256) void jacobi_apply_kernel(int64_t n, const T *__restrict__ jd,
                                                           const T *__restrict__ v,
                                                           T *__restrict__ z) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) z[i] = jd[i] * v[i];
}

// ---- ILU(0) -------------------------------------------------------------------------

template <typename TV, typename T>
__global__ __launch_bounds__(256) void ilu_copy_kernel(int64_t nnz, const int32_t *__restrict__ src,
                                                       const TV *__restrict__ vals,
                                                       T *__restrict__ lu) {
    const int64_t k = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (k < nnz) lu[k] = T(vals[src[k]]);
}

// One level of the row-wise (IKJ) ILU(0) factorisation; one wave per row.
template <typename T>
__global__ __launch_bounds__(256) void ilu_factor_level_kernel(int64_t nrows,
                                                               const int32_t *__restrict__ rows,
                                                               const int32_t *__restrict__ rs,
                                                               const int32_t *__restrict__ col,
                                                               const int32_t *__restrict__ diag,
                                                               T *__restrict__ lu,
                                                               int32_t *__restrict__ err,
                                                               const int32_t *__restrict__ perm,
                                                               const int32_t *__restrict__ row_sub,
                                                               const int32_t *__restrict__ need) {
    const int64_t w = (int64_t)blockIdx.x * 4 + threadIdx.x / 64;
    const int lane = threadIdx.x % 64;
    if (w >= nrows) return;
    const int32_t i = rows[w];
    // rows of sub-problems that are not being solved keep their old factor (sub-problems
    // never couple, so no needed row reads them)
    if (need && !need[row_sub[perm[i]]]) return;
    const int32_t r0 = rs[i], r1 = rs[i + 1], di = diag[i];
    for (int32_t kk = r0; kk < di; kk++) {
        const int32_t k = col[kk];
        const T piv = lu[diag[k]];
        const T l = lu[kk] / piv;  // every lane computes the same value
        if (!isfinite(double(l))) {
            if (lane == 0) err[0] = 1;
        }
        if (lane == 0) lu[kk] = l;
        const int32_t k0 = diag[k] + 1, k1 = rs[k + 1];
        for (int32_t jj = kk + 1 + lane; jj < r1; jj += 64) {
            const int32_t c = col[jj];
            int32_t lo = k0, hi = k1;
            while (lo < hi) {
                const int32_t mid = (lo + hi) >> 1;
                if (col[mid] < c) lo = mid + 1;
                else hi = mid;
            }
            if (lo < k1 && col[lo] == c) lu[jj] = lu[jj] - l * lu[lo];
        }
        // the next pivot column was updated by other lanes of this wave
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
    }
    if (lane == 0 && !(double(lu[di]) != 0.0 && isfinite(double(lu[di])))) err[0] = 1;
}

// One level of L y = b (unit lower) or U z = y; 16 lanes per row.
template <typename T, bool LOWER>
__global__ __launch_bounds__(256) void trsv_level_kernel(int64_t nrows, const int32_t *__restrict__ rows,
                                                         const int32_t *__restrict__ rs,
                                                         const int32_t *__restrict__ col,
                                                         const int32_t *__restrict__ diag,
                                                         const T *__restrict__ lu,
                                                         const T *__restrict__ b,
                                                         T *__restrict__ x) {
    const int64_t w = (int64_t)blockIdx.x * 16 + threadIdx.x / 16;
    const int lane = threadIdx.x % 16;
    if (w >= nrows) return;
    const int32_t i = rows[w];
    const int32_t k0 = LOWER ? rs[i] : diag[i] + 1;
    const int32_t k1 = LOWER ? diag[i] : rs[i + 1];
    T acc = T(0);
    for (int32_t k = k0 + lane; k < k1; k += 16) acc += lu[k] * x[col[k]];
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 16);
    if (lane == 0) x[i] = LOWER ? b[i] - acc : (b[i] - acc) / lu[diag[i]];
}

template <typename TA, typename TB>
__global__ __launch_bounds__(256) void permute_kernel(int64_t n, const int32_t *__restrict__ perm,
                                                      const TA *__restrict__ a, TB *__restrict__ b,
                                                      int forward) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    if (forward) b[i] = TB(a[perm[i]]);   // permuted[i] = original[perm[i]]
    else b[perm[i]] = TB(a[i]);
}

// ---- arclength continuation vector operations (per sub-problem scalars, masked) ------
// op 0: out = a + alpha[s] b        (predictor: Dof_current + Ds Dof_derivative)
// op 1: out = out - (a - alpha[s] b) (Newton update: x -= y - dparam z)
// op 2: out = -alpha[s] a            (Dof_derivative = -Parameter_derivative z)
// op 3: out = a - b                  (x - Dof_current)
// op 4: out = a                      (masked copy)
// op 5: out = 0
__global__ __launch_bounds__(256) void cont_vec_kernel(int op, int64_t n, const int32_t *__restrict__ dsub,
                                                       const int32_t *__restrict__ mask,
                                                       const double *__restrict__ alpha,
                                                       const double *__restrict__ a,
                                                       const double *__restrict__ b, double *out) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int s = dsub[i];
    if (!mask[s]) return;
    if (op == 0) out[i] = a[i] + alpha[s] * b[i];
    else if (op == 1) out[i] = out[i] - (a[i] - alpha[s] * b[i]);
    else if (op == 2) out[i] = -alpha[s] * a[i];
    else if (op == 3) out[i] = a[i] - b[i];
    else if (op == 4) out[i] = a[i];
    else out[i] = 0.0;
}

// ------------------------------------------------------------------------------------
// host side

static inline unsigned nblk(int64_t n, int per = 256) { return (unsigned)((n + per - 1) / per); }

static bool debug_enabled() {
    static const bool v = getenv("PYHIP_DEBUG") != nullptr;
    return v;
}

#define KCHECK() PYHIP_CHECK_HIP(hipGetLastError())

template <typename T>
static int kalloc(T **p, size_t n, KrylovWork *W) {
    return dev_alloc(p, n, W->bytes);
}

static int krylov_setup(pyhip_problem *p, int m) {
    PYHIP_REQUIRE(m >= 1 && m <= kMaxRestart, "restart must be in 1..1000");
    if (p->kw && p->kw->m >= m) return PYHIP_OK;
    KrylovWork *W = p->kw;
    if (!W) {
        W = new KrylovWork();
        p->kw = W;
        W->n = p->ndof;
        W->nsub = p->nsub;
        // chunks
        std::vector<int32_t> csub, scp(p->nsub + 1, 0);
        std::vector<int64_t> cb, ce;
        for (int s = 0; s < p->nsub; s++) {
            scp[s] = (int32_t)cb.size();
            for (int64_t b = p->dof_off_h[s]; b < p->dof_off_h[s + 1]; b += kChunk) {
                cb.push_back(b);
                ce.push_back(std::min<int64_t>(b + kChunk, p->dof_off_h[s + 1]));
                csub.push_back(s);
            }
        }
        scp[p->nsub] = (int32_t)cb.size();
        W->nchunk = (int64_t)cb.size();
        W->chunk_ptr_h = scp;
        int rc;
#define KUP(dst, v)                                                                        \
    if ((rc = kalloc(&(dst), (v).size(), W))) return rc;                                   \
    if (!(v).empty())                                                                      \
        PYHIP_CHECK_HIP(hipMemcpy(dst, (v).data(), (v).size() * sizeof((v)[0]), hipMemcpyHostToDevice));
        KUP(W->chunk_sub, csub);
        KUP(W->chunk_begin, cb);
        KUP(W->chunk_end, ce);
        KUP(W->sub_chunk_ptr, scp);
#undef KUP
        const int64_t n = W->n, ns = W->nsub;
        if ((rc = kalloc(&W->beta0, ns, W)) || (rc = kalloc(&W->bnorm, ns, W)) ||
            (rc = kalloc(&W->kcols, ns, W)) || (rc = kalloc(&W->active, ns, W)) ||
            (rc = kalloc(&W->nactive_d, 1, W)) || (rc = kalloc(&W->r64, n, W)) ||
            (rc = kalloc(&W->x64, n, W)) || (rc = kalloc(&W->d64, n, W)) ||
            (rc = kalloc(&W->b64, n, W)) || (rc = kalloc(&W->w64, n, W)) ||
            (rc = kalloc(&W->t64, n, W)) || (rc = kalloc(&W->r32, n, W)) ||
            (rc = kalloc(&W->d32, n, W)) || (rc = kalloc(&W->w32, n, W)) ||
            (rc = kalloc(&W->t32, n, W)) || (rc = kalloc(&W->res64, n, W)) ||
            (rc = kalloc(&W->maxres, ns, W)) || (rc = kalloc(&W->sub_scalar, ns, W)) ||
            (rc = kalloc(&W->jdiag64, n, W)) || (rc = kalloc(&W->jdiag32, n, W)) ||
            (rc = kalloc(&W->dof_mask, ns, W)) || (rc = kalloc(&W->need, ns, W)) ||
            (rc = kalloc(&W->cg_rz, ns, W)) || (rc = kalloc(&W->cg_pq, ns, W)) ||
            (rc = kalloc(&W->cg_beta, ns, W)) || (rc = kalloc(&W->nfail_d, 1, W)))
            return rc;
        // single-reduction CG work list
        W->sub_nnz_h.resize(ns);
        for (int s = 0; s < ns; s++)
            W->sub_nnz_h[s] = (int64_t)p->row_start_h[p->dof_off_h[s + 1]] - p->row_start_h[p->dof_off_h[s]];
        const size_t ngp = (size_t)3 * W->nchunk * kWlBlocksPerChunk;
        if ((rc = kalloc(&W->achunk, std::max<int64_t>(1, W->nchunk), W)) ||
            (rc = kalloc(&W->gpart, ngp, W)) || (rc = kalloc(&W->cg_iters, ns, W)) ||
            (rc = kalloc(&W->cg_iters2, ns, W)) ||
            (rc = kalloc(&W->cg_fail, ns, W)))
            return rc;
        if (ngp) PYHIP_CHECK_HIP(hipMemset(W->gpart, 0, ngp * sizeof(double)));
        PYHIP_CHECK_HIP(hipHostMalloc((void **)&W->achunk_h, std::max<int64_t>(1, W->nchunk) * sizeof(int32_t),
                                      hipHostMallocDefault));
        PYHIP_CHECK_HIP(hipHostMalloc((void **)&W->flags_h, std::max<int64_t>(1, 3 * ns) * sizeof(int32_t),
                                      hipHostMallocDefault));
        PYHIP_CHECK_HIP(hipHostMalloc((void **)&W->nactive_h, sizeof(int32_t), hipHostMallocDefault));
        PYHIP_CHECK_HIP(hipHostMalloc((void **)&W->act_h, std::max<int64_t>(1, ns) * sizeof(int32_t),
                                      hipHostMallocDefault));
        PYHIP_CHECK_HIP(hipHostMalloc((void **)&W->maxres_h, ns * sizeof(double), hipHostMallocDefault));
    }
    // restart-length dependent buffers
    int rc;
    dfree(W->partial);
    dfree(W->dots);
    dfree(W->H);
    dfree(W->cs);
    dfree(W->sn);
    dfree(W->g);
    dfree(W->y);
    dfree(W->V);
    W->V = nullptr;
    W->vprec = -1;
    W->vcount = 0;
    const int64_t ns = W->nsub;
    if ((rc = kalloc(&W->partial, (size_t)W->nchunk * (m + 2) + 1, W)) ||
        (rc = kalloc(&W->dots, (size_t)3 * (m + 2) * ns, W)) ||
        (rc = kalloc(&W->H, (size_t)ns * (m + 1) * m, W)) || (rc = kalloc(&W->cs, (size_t)ns * m, W)) ||
        (rc = kalloc(&W->sn, (size_t)ns * m, W)) || (rc = kalloc(&W->g, (size_t)ns * (m + 1), W)) ||
        (rc = kalloc(&W->y, (size_t)ns * m, W)))
        return rc;
    W->m = m;
    return PYHIP_OK;
}

// Krylov basis / work vectors V: at least nvec vectors of n entries in precision T.
template <typename T>
static int ensure_basis(KrylovWork *W, int64_t n, int nvec) {
    const int prec = std::is_same<T, float>::value ? PYHIP_F32 : PYHIP_F64;
    if (W->V && W->vprec == prec && W->vcount >= nvec) return PYHIP_OK;
    dfree(W->V);
    W->V = nullptr;
    W->vprec = -1;
    W->vcount = 0;
    int rc;
    if ((rc = dev_alloc((T **)&W->V, (size_t)nvec * n, W->bytes))) return rc;
    W->vprec = prec;
    W->vcount = nvec;
    return PYHIP_OK;
}

// Per-sub-problem dot products of nvec vectors with w: out[j][sub].
template <typename T>
// With act (device [nsub]) the chunks of sub-problems with act[s] == 0 are skipped and
// their sums are not defined.
static int batched_dots(pyhip_problem *p, KrylovWork *W, int nvec, const T *V, int64_t ldv,
                        const T *w, double *out, const int32_t *act = nullptr) {
    TimeScope ts(&p->timer, "vector", p->st);
    const int pstride = W->m + 2;
    if (W->nchunk == 0) return PYHIP_OK;
    hipLaunchKernelGGL((multidot_kernel<T>), dim3((unsigned)W->nchunk, (unsigned)((nvec + kDotVecs - 1) / kDotVecs)),
                       dim3(256), 0, p->st, nvec,
                       V, ldv, w, W->chunk_begin, W->chunk_end, W->partial, pstride, W->chunk_sub, act);
    KCHECK();
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((unsigned)(nvec * W->nsub)), dim3(64), 0, p->st, nvec,
                       W->nsub, W->sub_chunk_ptr, W->partial, pstride, out, 0);
    KCHECK();
    return PYHIP_OK;
}

int batched_maxabs(pyhip_problem *p, const double *v, double *out) {
    KrylovWork *W = p->kw;
    const int pstride = W->m + 2;
    if (W->nchunk == 0) return PYHIP_OK;
    hipLaunchKernelGGL(maxabs_kernel, dim3((unsigned)W->nchunk), dim3(256), 0, p->st, v,
                       W->chunk_begin, W->chunk_end, W->partial, pstride);
    KCHECK();
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((unsigned)W->nsub), dim3(64), 0, p->st, 1, W->nsub,
                       W->sub_chunk_ptr, W->partial, pstride, out, 1);
    KCHECK();
    return PYHIP_OK;
}

// Variant of this problem's spmv: the fastest of the canonical spmv's variants (1, 2, 4 or
// 8 lanes per row, LDS staging) over two timed launches each, all giving the same bits, or
// PYHIP_SPMV_LPR. The work-list CG kernel takes the fastest lane count. y is scratch here.
template <typename TV, typename TX, typename TY>
static int spmv_tune(pyhip_problem *p, KrylovWork *W, const TV *vals, const TX *x, TY *y) {
    if (W->smode) return PYHIP_OK;
    if (const char *e = getenv("PYHIP_SPMV_LPR")) {
        const int v = atoi(e);
        if (spmv_variant_ok(v)) {
            W->smode = v;
            W->lpr = v <= 8 ? v : 8;
            return PYHIP_OK;
        }
    }
    const int64_t n = p->ndof;
    W->smode = W->lpr = 8;
    if (n == 0) return PYHIP_OK;
    hipEvent_t e0, e1;
    PYHIP_CHECK_HIP(hipEventCreate(&e0));
    PYHIP_CHECK_HIP(hipEventCreate(&e1));
    float best = 1e30f, best_lanes = 1e30f;
    for (int v : kSpmvVariants) {
        launch_spmv<TV, TX, TY>(v, p->st, n, p->row_start, p->col, vals, x, y, nullptr);
        (void)hipEventRecord(e0, p->st);
        for (int k = 0; k < 2; k++)
            launch_spmv<TV, TX, TY>(v, p->st, n, p->row_start, p->col, vals, x, y, nullptr);
        (void)hipEventRecord(e1, p->st);
        float ms = 0.f;
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess) continue;
        if (ms < best) {
            best = ms;
            W->smode = v;
        }
        if (v <= 8 && ms < best_lanes) {
            best_lanes = ms;
            W->lpr = v;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    KCHECK();
    if (debug_enabled()) fprintf(stderr, "[pyhip] spmv variant: %d, work-list lanes per row: %d\n", W->smode, W->lpr);
    return PYHIP_OK;
}

// y = A x, or y = b - A x when b is given. With act (device [nsub], host copy act_h)
// only the rows of sub-problems with act[s] != 0 are computed.
template <typename TV, typename TX, typename TY>
static int spmv(pyhip_problem *p, const TV *vals, const TX *x, TY *y, const TY *b, const int32_t *act = nullptr,
                const int32_t *act_h = nullptr) {
    const int64_t n = p->ndof;
    if (n == 0) return PYHIP_OK;
    int rc;
    if ((rc = spmv_tune<TV, TX, TY>(p, p->kw, vals, x, y))) return rc;
    TimeScope ts(&p->timer, "spmv", p->st);
    launch_spmv<TV, TX, TY>(p->kw->smode, p->st, n, p->row_start, p->col, vals, x, y, b, act ? p->dof_sub : nullptr,
                            act);
    KCHECK();
    if (p->timer.on) {
        int64_t rows = n, nnz = p->nnz;
        if (act) {
            rows = nnz = 0;
            for (int s = 0; s < p->nsub; s++)
                if (act_h[s]) {
                    rows += p->dof_off_h[s + 1] - p->dof_off_h[s];
                    nnz += p->kw->sub_nnz_h[s];
                }
        }
        p->timer.add_work("spmv", (int64_t)(sizeof(TV) + 4) * nnz +
                                      (int64_t)(4 + sizeof(TX) + sizeof(TY) + (b ? sizeof(TY) : 0)) * rows + 4, 4);
    }
    return PYHIP_OK;
}

// ---- ILU(0) plan: permutation, permuted pattern, level sets (host, once per pattern)
//
// Order within a sub-problem: Taylor-Hood velocity unknowns first, then its pressure
// unknowns (other forms: one group). Within a group either the natural order, or a
// multicolour order: a greedy colouring of the group's rows (in natural order, each row
// takes the smallest colour none of its already coloured neighbours in the group has),
// rows sorted by (colour, index). Rows of one colour of one group are independent, so
// the triangular solves have at most as many levels as there are colours, against
// O(rows per side) levels in the natural order, at the price of a weaker factor.
// PYHIP_ILU_ORDER_AUTO takes the multicolour order for each sub-problem whose natural
// order has more than kIluAutoLevels lower-solve levels (decided per sub-problem, so a
// sub-problem gets the same factor alone and in any batch).
constexpr int kIluAutoLevels = 64;
constexpr int kGmresCheck = 8;   // GMRES iterations between host reads of the active count

static void ilu_free(KrylovWork *W) {
    IluPlan &L = W->ilu;
    void *ptrs[] = {L.perm, L.rs, L.col, L.src, L.diag, L.lo_rows, L.up_rows, L.lu,
                    L.tmp64, L.tmp64b, L.tmp32, L.tmp32b};
    for (void *q : ptrs)
        if (q) (void)hipFree(q);
    L = IluPlan();
}

static int ilu_plan_build(pyhip_problem *p, KrylovWork *W) {
    IluPlan &L = W->ilu;
    if (L.built && L.requested == W->ilu_order) return PYHIP_OK;
    if (L.built) {
        PYHIP_CHECK_HIP(hipStreamSynchronize(p->st));
        ilu_free(W);
    }
    const int64_t n = p->ndof;
    const auto &rs = p->row_start_h;
    const auto &col = p->col_h;
    // group of each row: 0 velocity / other, 1 Taylor-Hood pressure or EnforcedBC multiplier
    std::vector<char> is_p(n, 0);
    {
        if ((int64_t)p->ilu_last_h.size() == n) is_p = p->ilu_last_h;
        // pressure unknowns: eqn of the pressure slots (2, 7, 16, 21) of Taylor-Hood elements
        for (auto &G : p->groups) {
            if (!th_layout(G.form)) continue;
            std::vector<int32_t> seqn((size_t)G.ns * G.nel);
            PYHIP_CHECK_HIP(hipMemcpy(seqn.data(), G.seqn, seqn.size() * sizeof(int32_t), hipMemcpyDeviceToHost));
            const int ps[4] = {2, 7, 16, 21};
            for (int k = 0; k < 4; k++)
                for (int64_t e = 0; e < G.nel; e++) {
                    const int32_t q = seqn[(size_t)ps[k] * G.nel + e];
                    if (q >= 0) is_p[q] = 1;
                }
        }
    }
    // EnforcedBC multipliers (value 1 of a scalar form's nodes) go last as well: their
    // pivots come from the Schur complement of the rows before them
    // (problems made from a bare CSR matrix, pyhip_solve_serial, carry no sub-problem data)
    if ((int64_t)p->dof_val_h.size() == n)
        for (size_t s = 0; s < p->subs.size(); s++)
            if (p->subs[s].nenf > 0)
                for (int64_t r = p->dof_off_h[s]; r < p->dof_off_h[s + 1]; r++)
                    if (p->dof_val_h[r] == 1) is_p[r] = 1;
    IluHostPlan H;
    std::string err;
    if (ilu_plan_host(n, rs.data(), col.data(), p->dof_off_h, is_p, W->ilu_order, kIluAutoLevels, H, err)) {
        set_error(err);
        return PYHIP_ERR_BREAKDOWN;
    }
    const auto &perm = H.perm, &prs = H.prs, &pcol = H.pcol, &psrc = H.psrc, &pdiag = H.pdiag;
    const auto &lo_rows = H.lo_rows, &up_rows = H.up_rows;
    L.lo_ptr = H.lo_ptr;
    L.up_ptr = H.up_ptr;
    L.requested = W->ilu_order;
    L.multicolor = H.multicolor;
    int rc;
#define LUP(dst, v)                                                                        \
    if ((rc = kalloc(&(dst), (v).size(), W))) return rc;                                   \
    PYHIP_CHECK_HIP(hipMemcpy(dst, (v).data(), (v).size() * sizeof((v)[0]), hipMemcpyHostToDevice));
    LUP(L.perm, perm);
    LUP(L.rs, prs);
    LUP(L.col, pcol);
    LUP(L.src, psrc);
    LUP(L.diag, pdiag);
    LUP(L.lo_rows, lo_rows);
    LUP(L.up_rows, up_rows);
#undef LUP
    if ((rc = kalloc(&L.tmp64, n, W)) || (rc = kalloc(&L.tmp64b, n, W)) ||
        (rc = kalloc(&L.tmp32, n, W)) || (rc = kalloc(&L.tmp32b, n, W)))
        return rc;
    L.n = n;
    L.nnz = p->nnz;
    L.built = true;
    return PYHIP_OK;
}

template <typename TV, typename T>
static int ilu_factor(pyhip_problem *p, KrylovWork *W, const TV *vals, const int32_t *need) {
    IluPlan &L = W->ilu;
    int rc;
    const int prec = std::is_same<T, float>::value ? PYHIP_F32 : PYHIP_F64;
    if (L.lu_prec != prec) {
        dfree(L.lu);
        L.lu = nullptr;
        if ((rc = dev_alloc((T **)&L.lu, L.nnz, W->bytes))) return rc;
        L.lu_prec = prec;
    }
    T *lu = (T *)L.lu;
    hipLaunchKernelGGL((ilu_copy_kernel<TV, T>), dim3(nblk(L.nnz)), dim3(256), 0, p->st, L.nnz, L.src,
                       vals, lu);
    KCHECK();
    PYHIP_CHECK_HIP(hipMemsetAsync(W->nactive_d, 0, sizeof(int32_t), p->st));
    for (size_t l = 0; l + 1 < L.lo_ptr.size(); l++) {
        const int64_t nr = L.lo_ptr[l + 1] - L.lo_ptr[l];
        hipLaunchKernelGGL((ilu_factor_level_kernel<T>), dim3(nblk(nr, 4)), dim3(256), 0, p->st, nr,
                           L.lo_rows + L.lo_ptr[l], L.rs, L.col, L.diag, lu, W->nactive_d, L.perm, p->dof_sub,
                           need);
    }
    KCHECK();
    PYHIP_CHECK_HIP(hipMemcpyAsync(W->nactive_h, W->nactive_d, sizeof(int32_t), hipMemcpyDeviceToHost, p->st));
    PYHIP_CHECK_HIP(hipStreamSynchronize(p->st));
    if (*W->nactive_h) {
        set_error("ILU(0) breakdown: zero or non-finite pivot");
        return PYHIP_ERR_BREAKDOWN;
    }
    return PYHIP_OK;
}

template <typename T>
static int ilu_apply(pyhip_problem *p, KrylovWork *W, const T *v, T *z) {
    IluPlan &L = W->ilu;
    const T *lu = (const T *)L.lu;
    T *a = std::is_same<T, float>::value ? (T *)L.tmp32 : (T *)L.tmp64;
    T *b = std::is_same<T, float>::value ? (T *)L.tmp32b : (T *)L.tmp64b;
    const int64_t n = L.n;
    hipLaunchKernelGGL((permute_kernel<T, T>), dim3(nblk(n)), dim3(256), 0, p->st, n, L.perm, v, a, 1);
    for (size_t l = 0; l + 1 < L.lo_ptr.size(); l++) {
        const int64_t nr = L.lo_ptr[l + 1] - L.lo_ptr[l];
        hipLaunchKernelGGL((trsv_level_kernel<T, true>), dim3(nblk(nr, 16)), dim3(256), 0, p->st, nr,
                           L.lo_rows + L.lo_ptr[l], L.rs, L.col, L.diag, lu, a, b);
    }
    for (size_t l = 0; l + 1 < L.up_ptr.size(); l++) {
        const int64_t nr = L.up_ptr[l + 1] - L.up_ptr[l];
        hipLaunchKernelGGL((trsv_level_kernel<T, false>), dim3(nblk(nr, 16)), dim3(256), 0, p->st, nr,
                           L.up_rows + L.up_ptr[l], L.rs, L.col, L.diag, lu, b, a);
    }
    hipLaunchKernelGGL((permute_kernel<T, T>), dim3(nblk(n)), dim3(256), 0, p->st, n, L.perm, a, z, 0);
    KCHECK();
    return PYHIP_OK;
}

static int need_on_host(pyhip_problem *p, KrylovWork *W, const int32_t *need);

// ---- block-triangular preconditioner of Taylor-Hood Jacobians (BlockPlan in krylov.h) ----
//
// J = [A G; D 0] (velocity rows / pressure rows). The upper block-triangular
// [A G; 0 -S]^-1 with the Schur complement S = -D A^-1 G approximated by the lumped Q1
// pressure mass over the viscosity (a mesh-independent equivalent of the Stokes Schur
// complement, here with the sign the Navier-Stokes Jacobian's Schur diagonal has) and A^-1
// by one multigrid V-cycle on A_d, the velocity block without the coupling of the two
// velocity components. Sub-problems of scalar forms have no pressure rows: all of their
// matrix is in A_d, so for them this is the multigrid preconditioner.

__global__ void blk_gather_kernel(int64_t n, const int32_t *__restrict__ src, const double *__restrict__ a,
                                  double *__restrict__ b) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) b[i] = a[src[i]];
}

// sinv = mu / m_p per pressure unknown (mu: the first parameter of its sub-problem)
__global__ void blk_sinv_kernel(int64_t np, const int32_t *__restrict__ psub, const double *__restrict__ pmass,
                                const double *__restrict__ params, int pstride, double *__restrict__ sinv) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i < np) sinv[i] = params[(int64_t)psub[i] * pstride] / pmass[i];
}

// z_p = sinv * r_p, into z and into pz (whose velocity rows stay 0)
__global__ void blk_pressure_kernel(int64_t np, const int32_t *__restrict__ prow, const int32_t *__restrict__ psub,
                                    const double *__restrict__ sinv, const double *__restrict__ r,
                                    double *__restrict__ pz, double *__restrict__ z,
                                    const int32_t *__restrict__ act) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= np || (act && !act[psub[i]])) return;
    const int32_t row = prow[i];
    const double v = r[row] * sinv[i];
    pz[row] = v;
    z[row] = v;
}

// tv = r_u - (G z_p)_u (t = J pz; without pressure rows t is not read)
__global__ void blk_vgather_kernel(int64_t nv, const int32_t *__restrict__ vrows, const int32_t *__restrict__ vsub,
                                   const double *__restrict__ r, const double *__restrict__ t,
                                   double *__restrict__ tv, const int32_t *__restrict__ act) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nv || (act && !act[vsub[i]])) return;
    const int32_t row = vrows[i];
    tv[i] = t ? r[row] - t[row] : r[row];
}

__global__ void blk_vscatter_kernel(int64_t nv, const int32_t *__restrict__ vrows, const int32_t *__restrict__ vsub,
                                    const double *__restrict__ zv, double *__restrict__ z,
                                    const int32_t *__restrict__ act) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= nv || (act && !act[vsub[i]])) return;
    z[vrows[i]] = zv[i];
}

// host plan, once per pattern: the velocity / pressure split, A_d's pattern with the source
// of each entry, and A_d's multigrid plan
static int blk_plan_build(pyhip_problem *p, KrylovWork *W, bool element_level) {
    BlockPlan &B = W->blk;
    if (B.built) return PYHIP_OK;
    const int64_t n = p->ndof;
    const int ns = p->nsub;
    PYHIP_REQUIRE((int64_t)p->dof_val_h.size() == n && (int64_t)p->pmass_h.size() == n,
                  "block preconditioner: unknown kinds missing");
    std::vector<int32_t> vmap(n, -1), vrows, vsub, prow, psub, vsrc;
    std::vector<double> pmass;
    B.off_h.assign(ns + 1, 0);
    for (int s = 0; s < ns; s++) {
        const bool th = th_layout(p->subs[s].form);
        for (int64_t r = p->dof_off_h[s]; r < p->dof_off_h[s + 1]; r++) {
            if (th && p->dof_val_h[r] == 2) {
                PYHIP_REQUIRE(p->pmass_h[r] > 0.0, "block preconditioner: pressure unknown with no element");
                prow.push_back((int32_t)r);
                psub.push_back(s);
                pmass.push_back(p->pmass_h[r]);
            } else {
                vmap[r] = (int32_t)vrows.size();
                vrows.push_back((int32_t)r);
                vsub.push_back(s);
            }
        }
        B.off_h[s + 1] = (int64_t)vrows.size();
    }
    B.rs_h.assign(1, 0);
    B.col_h.clear();
    for (const int32_t r : vrows) {
        for (int32_t e = p->row_start_h[r]; e < p->row_start_h[r + 1]; e++) {
            const int32_t c = p->col_h[e];
            if (vmap[c] >= 0 && p->dof_val_h[c] == p->dof_val_h[r]) {
                B.col_h.push_back(vmap[c]);
                vsrc.push_back(e);
            }
        }
        B.rs_h.push_back((int32_t)B.col_h.size());
    }
    B.nv = (int64_t)vrows.size();
    B.np = (int64_t)prow.size();
    B.nnzv = (int64_t)B.col_h.size();
    int rc;
    auto up = [&](int32_t **dst, const std::vector<int32_t> &v) -> int {
        if ((rc = dev_alloc(dst, std::max<size_t>(1, v.size()), W->bytes))) return rc;
        if (!v.empty()) PYHIP_CHECK_HIP(hipMemcpy(*dst, v.data(), v.size() * sizeof(int32_t), hipMemcpyHostToDevice));
        return PYHIP_OK;
    };
    if ((rc = up(&B.vrows, vrows)) || (rc = up(&B.vsub, vsub)) || (rc = up(&B.vsrc, vsrc)) ||
        (rc = up(&B.prow, prow)) || (rc = up(&B.psub, psub)))
        return rc;
    if ((rc = dev_alloc(&B.pmass, std::max<int64_t>(1, B.np), W->bytes)) ||
        (rc = dev_alloc(&B.sinv, std::max<int64_t>(1, B.np), W->bytes)) ||
        (rc = dev_alloc(&B.vals, std::max<int64_t>(1, B.nnzv), W->bytes)) ||
        (rc = dev_alloc(&B.tv, std::max<int64_t>(1, B.nv), W->bytes)) ||
        (rc = dev_alloc(&B.zv, std::max<int64_t>(1, B.nv), W->bytes)) ||
        (rc = dev_alloc(&B.pz, std::max<int64_t>(1, n), W->bytes)) ||
        (rc = dev_alloc(&B.t, std::max<int64_t>(1, n), W->bytes)))
        return rc;
    if (B.np) PYHIP_CHECK_HIP(hipMemcpy(B.pmass, pmass.data(), B.np * sizeof(double), hipMemcpyHostToDevice));
    PYHIP_CHECK_HIP(hipMemset(B.pz, 0, std::max<int64_t>(1, n) * sizeof(double)));
    B.amg.nonsym = true;
    // the element level of A_d: the problem's rows restricted to the velocity rows
    AmgFixedInterp F;
    const AmgFixedInterp &G = p->amg_fix_vel.rs.empty() ? p->amg_fix : p->amg_fix_vel;   // no Taylor-Hood: scalar rows
    const bool fix = element_level && G.rs.size() == (size_t)n + 1;
    if (fix) {
        F.on = G.on;
        F.nc = G.nc;
        F.rs.assign(1, 0);
        for (const int32_t r : vrows) {
            for (int32_t t = G.rs[r]; t < G.rs[r + 1]; t++) {
                F.col.push_back(G.col[t]);
                F.w.push_back(G.w[t]);
            }
            F.rs.push_back((int32_t)F.col.size());
        }
    }
    if ((rc = amg_build(B.amg, B.nv, B.rs_h.data(), B.col_h.data(), B.off_h, W->bytes, fix ? &F : nullptr)))
        return rc;
    B.element_level = element_level;
    B.built = true;
    return PYHIP_OK;
}

static void blk_free(KrylovWork *W) {
    BlockPlan &B = W->blk;
    void *ptrs[] = {B.vrows, B.vsrc, B.prow, B.psub, B.vsub, B.pmass, B.sinv, B.vals, B.tv, B.zv, B.pz, B.t};
    for (void *q : ptrs) dfree(q);
    amg_free(B.amg);
    W->blk = BlockPlan();
}

// numeric set-up from the FP64 Jacobian values: A_d's values and hierarchy, mu / m_p
static int blk_setup(pyhip_problem *p, KrylovWork *W, const double *vals, const int32_t *need, bool element_level) {
    if (W->blk.built && W->blk.element_level != element_level) blk_free(W);
    int rc = blk_plan_build(p, W, element_level);
    if (rc) return rc;
    BlockPlan &B = W->blk;
    B.jvals = vals;
    if (B.nnzv)
        hipLaunchKernelGGL(blk_gather_kernel, dim3(nblk(B.nnzv)), dim3(256), 0, p->st, B.nnzv, B.vsrc, vals, B.vals);
    if (B.np)
        hipLaunchKernelGGL(blk_sinv_kernel, dim3(nblk(B.np)), dim3(256), 0, p->st, B.np, B.psub, B.pmass, p->params,
                           kParamStride, B.sinv);
    KCHECK();
    if ((rc = need_on_host(p, W, need))) return rc;
    return amg_setup(B.amg, p->st, B.vals, &p->timer, need ? W->need_h.data() : nullptr);
}

static int blk_apply(pyhip_problem *p, KrylovWork *W, const double *r, double *z, const int32_t *act,
                     const int32_t *act_h) {
    BlockPlan &B = W->blk;
    int rc;
    if (B.np) {
        hipLaunchKernelGGL(blk_pressure_kernel, dim3(nblk(B.np)), dim3(256), 0, p->st, B.np, B.prow, B.psub, B.sinv,
                           r, B.pz, z, act);
        KCHECK();
        if ((rc = spmv<double, double, double>(p, B.jvals, B.pz, B.t, (const double *)nullptr, act, act_h))) return rc;
    }
    if (B.nv) {
        hipLaunchKernelGGL(blk_vgather_kernel, dim3(nblk(B.nv)), dim3(256), 0, p->st, B.nv, B.vrows, B.vsub, r,
                           B.np ? (const double *)B.t : nullptr, B.tv, act);
        KCHECK();
        if ((rc = amg_apply(B.amg, p->st, B.tv, B.zv, &p->timer, act, act_h))) return rc;
        hipLaunchKernelGGL(blk_vscatter_kernel, dim3(nblk(B.nv)), dim3(256), 0, p->st, B.nv, B.vrows, B.vsub, B.zv, z,
                           act);
        KCHECK();
    }
    return PYHIP_OK;
}

template <typename T>
// z = M^-1 v. The multigrid V-cycle computes only the rows of sub-problems with
// act[s] != 0 when act (device, host copy act_h) is given; the other preconditioners
// compute every row.
static int precond_apply(pyhip_problem *p, KrylovWork *W, int pc, const T *v, T *z, const int32_t *act = nullptr,
                         const int32_t *act_h = nullptr, bool x0_ready = false) {
    TimeScope ts(&p->timer, "precond", p->st);
    if (pc == PYHIP_PRECOND_ILU0) return ilu_apply<T>(p, W, v, z);
    if (pc == PYHIP_PRECOND_BLOCK) {
        if constexpr (std::is_same<T, double>::value) return blk_apply(p, W, v, z, act, act_h);
        set_error("the block preconditioner runs FP64 solves");
        return PYHIP_ERR_INVALID;
    }
    if (pc == PYHIP_PRECOND_AMG) {
        if constexpr (std::is_same<T, double>::value)
            return amg_apply(W->amg, p->st, v, z, &p->timer, act, act_h, x0_ready);
        else
            return amg_apply32(W->amg, p->st, v, z, &p->timer, act, act_h);
    }
    const T *jd = std::is_same<T, float>::value ? (const T *)W->jdiag32 : (const T *)W->jdiag64;
    hipLaunchKernelGGL((jacobi_apply_kernel<T>), dim3(nblk(p->ndof)), dim3(256), 0, p->st, p->ndof, jd, v, z);
    KCHECK();
    return PYHIP_OK;
}

// W->need_h = the device flags need[nsub] (waits for the stream)
static int need_on_host(pyhip_problem *p, KrylovWork *W, const int32_t *need) {
    if (!need) return PYHIP_OK;
    W->need_h.resize(W->nsub);
    PYHIP_CHECK_HIP(hipMemcpyAsync(W->need_h.data(), need, W->nsub * sizeof(int32_t), hipMemcpyDeviceToHost, p->st));
    PYHIP_CHECK_HIP(hipStreamSynchronize(p->st));
    return PYHIP_OK;
}

template <typename TV, typename T>
static int precond_setup(pyhip_problem *p, KrylovWork *W, int pc, const TV *vals, const int32_t *need = nullptr,
                         bool element_level = true) {
    TimeScope ts(&p->timer, "precond", p->st);
    if (pc == PYHIP_PRECOND_JACOBI || pc == PYHIP_PRECOND_AMG)
        for (const auto &S : p->subs)
            PYHIP_REQUIRE(S.nenf == 0, "EnforcedBC multipliers give rows with a zero diagonal: "
                                       "use the ILU(0) preconditioner");
    if (pc == PYHIP_PRECOND_BLOCK) {
        if constexpr (std::is_same<T, double>::value && std::is_same<TV, double>::value)
            return blk_setup(p, W, vals, need, element_level);
        set_error("the block preconditioner runs FP64 solves");
        return PYHIP_ERR_INVALID;
    }
    if (pc == PYHIP_PRECOND_AMG) {
        // the element level where the problem has one (problems built from elements)
        const AmgFixedInterp *fix =
            element_level && p->amg_fix.rs.size() == (size_t)p->ndof + 1 ? &p->amg_fix : nullptr;
        if (W->amg.built && W->amg.element_level != (fix != nullptr)) {
            amg_free(W->amg);
            W->amg_stale = true;
        }
        if constexpr (std::is_same<T, double>::value && std::is_same<TV, double>::value) {
            int rc = amg_build(W->amg, p->ndof, p->row_start_h.data(), p->col_h.data(), p->dof_off_h, W->bytes, fix);
            if (rc) return rc;
            // the hierarchy (weights, prolongators, coarse operators) is computed from the
            // first matrix of a solve and kept for its later Newton steps; the finest level
            // always uses the current matrix
            if (!W->amg_stale) return amg_set_fine(W->amg, vals);
            if ((rc = need_on_host(p, W, need))) return rc;
            rc = amg_setup(W->amg, p->st, vals, &p->timer, need ? W->need_h.data() : nullptr);
            W->amg_stale = rc != PYHIP_OK;   // a failed set-up is redone by the next solve
            return rc;
        } else if constexpr (std::is_same<T, float>::value && std::is_same<TV, float>::value) {
            // FP32 V-cycle: the hierarchy is set up in FP64 (from the FP64 Jacobian when the
            // caller has one, else from the FP32 values widened), then rounded to FP32
            int rc = amg_build(W->amg, p->ndof, p->row_start_h.data(), p->col_h.data(), p->dof_off_h, W->bytes, fix);
            if (rc) return rc;
            if (W->amg_stale) {
                const double *src = W->amg_src64;
                if (!src) {
                    if (!W->amg_conv64 && (rc = dev_alloc(&W->amg_conv64, p->nnz, W->bytes))) return rc;
                    hipLaunchKernelGGL((convert_kernel<float, double>), dim3(nblk(p->nnz)), dim3(256), 0, p->st,
                                       p->nnz, vals, W->amg_conv64);
                    KCHECK();
                    src = W->amg_conv64;
                }
                if ((rc = need_on_host(p, W, need))) return rc;
                rc = amg_setup(W->amg, p->st, src, &p->timer, need ? W->need_h.data() : nullptr);
                W->amg_stale = rc != PYHIP_OK;
                if (rc || (rc = amg_make_f32(W->amg, p->st, W->bytes))) return rc;
            }
            return amg_set_fine32(W->amg, vals);
        }
        set_error("the multigrid preconditioner runs with matching solve and matrix precision");
        return PYHIP_ERR_INVALID;
    }
    if (pc == PYHIP_PRECOND_ILU0) {
        int rc = ilu_plan_build(p, W);
        if (rc) return rc;
        return ilu_factor<TV, T>(p, W, vals, need);
    }
    T *jd = std::is_same<T, float>::value ? (T *)W->jdiag32 : (T *)W->jdiag64;
    hipLaunchKernelGGL((jacobi_setup_kernel<TV, T>), dim3(nblk(p->ndof)), dim3(256), 0, p->st, p->ndof,
                       vals, p->diag, jd);
    KCHECK();
    return PYHIP_OK;
}

// GMRES in precision T on matrix values TV: x = A^{-1} b per sub-problem with need[s].
// b, x: device vectors of type T. Returns the iterations of the longest-running one.
template <typename T, typename TV>
static int gmres(pyhip_problem *p, KrylovWork *W, const TV *vals, const T *b, T *x,
                 const pyhip_linear_opts &o, double rtol, double stag, const int32_t *need, int *iters,
                 int *failed) {
    const int64_t n = p->ndof;
    // m: layout stride of the workspace (sized for the longest restart asked so far);
    // mr: this solve's restart length, so the result does not depend on earlier solves
    const int m = W->m, ns = W->nsub;
    const int mr = o.restart;
    PYHIP_REQUIRE(mr >= 1 && mr <= m, "GMRES restart longer than the workspace");
    const int prec = std::is_same<T, float>::value ? PYHIP_F32 : PYHIP_F64;
    int rc;
    if ((rc = ensure_basis<T>(W, n, mr + 1))) return rc;
    T *V = (T *)W->V;
    T *w = std::is_same<T, float>::value ? (T *)W->w32 : (T *)W->w64;
    T *t = std::is_same<T, float>::value ? (T *)W->t32 : (T *)W->t64;
    double *d1 = W->dots, *d2 = W->dots + (m + 2) * ns, *d3 = W->dots + 2 * (m + 2) * ns;
    // |b| per sub-problem
    if ((rc = batched_dots<T>(p, W, 1, b, n, b, d3))) return rc;
    hipLaunchKernelGGL(sqrt_copy_kernel, dim3(nblk(ns, 64)), dim3(64), 0, p->st, ns, d3, W->bnorm);
    KCHECK();
    // x0 = 0
    PYHIP_CHECK_HIP(hipMemsetAsync(x, 0, n * sizeof(T), p->st));
    int total = 0;
    const int max_cycles = std::max(1, (o.max_iters + mr - 1) / mr);
    for (int cyc = 0; cyc < max_cycles; cyc++) {
        // r = b - A x  -> V_0 (unnormalised)
        if ((rc = spmv<TV, T, T>(p, vals, x, V, b))) return rc;
        if ((rc = batched_dots<T>(p, W, 1, V, n, V, d3))) return rc;
        PYHIP_CHECK_HIP(hipMemsetAsync(W->nactive_d, 0, sizeof(int32_t), p->st));
        hipLaunchKernelGGL(gmres_start_kernel, dim3(nblk(ns, 64)), dim3(64), 0, p->st, ns, m, d3,
                           W->bnorm, rtol, stag, cyc == 0 ? 1 : 0, need, W->g, W->active, W->kcols, W->beta0,
                           W->nactive_d);
        KCHECK();
        // the running flags come to the host with the count: between reads they only lose
        // sub-problems, so the host copy stays a superset (it sizes the byte counts of the
        // filtered launches; the device flags decide which rows are computed)
        int32_t *act_h = W->act_h;
        PYHIP_CHECK_HIP(hipMemcpyAsync(W->nactive_h, W->nactive_d, sizeof(int32_t), hipMemcpyDeviceToHost, p->st));
        PYHIP_CHECK_HIP(hipMemcpyAsync(act_h, W->active, ns * sizeof(int32_t), hipMemcpyDeviceToHost, p->st));
        PYHIP_CHECK_HIP(hipStreamSynchronize(p->st));
        if (debug_enabled()) {
            double b0 = 0, bn = 0;
            (void)hipMemcpy(&b0, W->beta0, sizeof(double), hipMemcpyDeviceToHost);
            (void)hipMemcpy(&bn, W->bnorm, sizeof(double), hipMemcpyDeviceToHost);
            fprintf(stderr, "[pyhip] gmres cycle %d: active %d |r| %.3e |b| %.3e iters %d\n", cyc,
                    *W->nactive_h, b0, bn, total);
        }
        if (*W->nactive_h == 0) break;
        hipLaunchKernelGGL((scale_kernel<T>), dim3(nblk(n)), dim3(256), 0, p->st, n, V, d3, p->dof_sub,
                           W->active, V);
        KCHECK();
        int k = 0;
        for (; k < mr; k++) {
            T *vk = V + (int64_t)k * n;
            // the preconditioner (multigrid) and the spmv skip the rows of stopped
            // sub-problems: nothing reads their t and w
            if ((rc = precond_apply<T>(p, W, o.precond, vk, t, W->active, act_h))) return rc;
            if ((rc = spmv<TV, T, T>(p, vals, t, w, (const T *)nullptr, W->active, act_h))) return rc;
            // CGS2
            // (the Gram-Schmidt dots skip the sub-problems that have stopped: every kernel
            // that reads them skips those too)
            if ((rc = batched_dots<T>(p, W, k + 1, V, n, w, d1, W->active))) return rc;
            {
                TimeScope tsu(&p->timer, "vector", p->st);
                hipLaunchKernelGGL((cgs_update_kernel<T>), dim3(nblk(n)), dim3(256), 0, p->st, n, k + 1, ns, V,
                                   n, d1, p->dof_sub, W->active, w);
            }
            if ((rc = batched_dots<T>(p, W, k + 1, V, n, w, d2, W->active))) return rc;
            {
                TimeScope tsu(&p->timer, "vector", p->st);
                hipLaunchKernelGGL((cgs_update_kernel<T>), dim3(nblk(n)), dim3(256), 0, p->st, n, k + 1, ns, V,
                                   n, d2, p->dof_sub, W->active, w);
            }
            if ((rc = batched_dots<T>(p, W, 1, w, n, w, d3, W->active))) return rc;
            {
                TimeScope tsu(&p->timer, "vector", p->st);
                hipLaunchKernelGGL((scale_kernel<T>), dim3(nblk(n)), dim3(256), 0, p->st, n, w, d3, p->dof_sub,
                                   W->active, V + (int64_t)(k + 1) * n);
            }
            PYHIP_CHECK_HIP(hipMemsetAsync(W->nactive_d, 0, sizeof(int32_t), p->st));
            hipLaunchKernelGGL(gmres_givens_kernel, dim3(nblk(ns, 64)), dim3(64), 0, p->st, ns, m, mr, k, d1, d2,
                               d3, W->H, W->cs, W->sn, W->g, W->bnorm, rtol, W->active, W->kcols,
                               W->nactive_d);
            KCHECK();
            total++;
            // the host reads the number of sub-problems still iterating every kGmresCheck
            // iterations (and at the end of the cycle); in between, a sub-problem that
            // has converged keeps its state (every kernel skips it) while the others go on
            if ((k + 1) % kGmresCheck == 0 || k + 1 == mr) {
                PYHIP_CHECK_HIP(hipMemcpyAsync(W->nactive_h, W->nactive_d, sizeof(int32_t), hipMemcpyDeviceToHost,
                                               p->st));
                PYHIP_CHECK_HIP(hipMemcpyAsync(act_h, W->active, ns * sizeof(int32_t), hipMemcpyDeviceToHost, p->st));
                PYHIP_CHECK_HIP(hipStreamSynchronize(p->st));
                if (*W->nactive_h == 0) break;
            }
        }
        // x += M^{-1} V y
        hipLaunchKernelGGL(gmres_backsolve_kernel, dim3(nblk(ns, 64)), dim3(64), 0, p->st, ns, m, W->H, W->g,
                           W->kcols, W->y);
        hipLaunchKernelGGL((combine_kernel<T>), dim3(nblk(n)), dim3(256), 0, p->st, n, m, V, n, W->y,
                           W->kcols, p->dof_sub, w);
        KCHECK();
        if ((rc = precond_apply<T>(p, W, o.precond, w, t))) return rc;
        hipLaunchKernelGGL((axpy_kernel<T, T>), dim3(nblk(n)), dim3(256), 0, p->st, n, t, x, 1.0,
                           p->dof_sub, need);
        KCHECK();
        if (total >= o.max_iters) break;
    }
    // sub-problems still above tolerance (true residual of the returned x)
    if ((rc = spmv<TV, T, T>(p, vals, x, V, b))) return rc;
    if ((rc = batched_dots<T>(p, W, 1, V, n, V, d3))) return rc;
    PYHIP_CHECK_HIP(hipMemsetAsync(W->nactive_d, 0, sizeof(int32_t), p->st));
    hipLaunchKernelGGL(gmres_start_kernel, dim3(nblk(ns, 64)), dim3(64), 0, p->st, ns, m, d3, W->bnorm,
                       rtol * 1.01, 0.0, 1, need, W->g, W->active, W->kcols, W->beta0, W->nactive_d);
    KCHECK();
    PYHIP_CHECK_HIP(hipMemcpyAsync(W->nactive_h, W->nactive_d, sizeof(int32_t), hipMemcpyDeviceToHost, p->st));
    PYHIP_CHECK_HIP(hipMemcpyAsync(W->act_h, W->active, ns * sizeof(int32_t), hipMemcpyDeviceToHost, p->st));
    PYHIP_CHECK_HIP(hipStreamSynchronize(p->st));
    // a sub-problem still marked active is above its tolerance
    W->lfail_h.assign(W->act_h, W->act_h + ns);
    *failed = *W->nactive_h;
    *iters = total;
    return PYHIP_OK;
}



// ---- preconditioned conjugate gradients (symmetric positive definite systems) ----------

template <typename T>
__global__ __launch_bounds__(256) void dot2_kernel(const T *__restrict__ a, const T *__restrict__ b,
                                                   const T *__restrict__ c, const T *__restrict__ d,
                                                   const int64_t *__restrict__ cb,
                                                   const int64_t *__restrict__ ce,
                                                   double *__restrict__ partial, int pstride,
                                                   const int32_t *__restrict__ csub,
                                                   const int32_t *__restrict__ act) {
    __shared__ double r0[256], r1[256];
    const int64_t ch = blockIdx.x;
    if (act && !act[csub[ch]]) return;   // a stopped sub-problem's chunk (whole block)
    double s0 = 0.0, s1 = 0.0;
    for (int64_t i = cb[ch] + threadIdx.x; i < ce[ch]; i += 256) {
        s0 += double(a[i]) * double(b[i]);
        s1 += double(c[i]) * double(d[i]);
    }
    r0[threadIdx.x] = s0;
    r1[threadIdx.x] = s1;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            r0[threadIdx.x] += r0[threadIdx.x + s];
            r1[threadIdx.x] += r1[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partial[ch * pstride] = r0[0];
        partial[ch * pstride + 1] = r1[0];
    }
}

// nactive = number of sub-problems with active[s] != 0 (one block, fixed order)
__global__ __launch_bounds__(256) void count_active_kernel(int nsub, const int32_t *__restrict__ active,
                                                           int32_t *__restrict__ nactive) {
    __shared__ int32_t c[256];
    int32_t k = 0;
    for (int s = threadIdx.x; s < nsub; s += 256) k += active[s] != 0;
    c[threadIdx.x] = k;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) c[threadIdx.x] += c[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) *nactive = c[0];
}

constexpr int kPcgRead = 4;   // preconditioned-CG iterations between host reads of the flags

__global__ void zero_i32_kernel(int n, int32_t *__restrict__ v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = 0;
}

// start: bnorm = |r0|, rz = (r0, z0); active where needed and |b| > 0
__global__ void cg_start_kernel(int nsub, const double *__restrict__ d2,
                                const int32_t *__restrict__ need, double *__restrict__ bnorm,
                                double *__restrict__ rz, int32_t *__restrict__ active,
                                int32_t *__restrict__ nactive) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsub) return;
    bnorm[s] = sqrt(d2[s]);
    rz[s] = d2[nsub + s];
    const int a = need[s] && d2[s] > 0.0 && isfinite(d2[s]);
    active[s] = a;
    if (a) atomicAdd(nactive, 1);
}

// alpha = rz / (p, Ap); a non-positive curvature stops the sub-problem (breakdown)
// (also counts, per sub-problem, the iterations whose spmv visited it)
__global__ void cg_alpha_kernel(int nsub, const double *__restrict__ pq_in, double *__restrict__ pq,
                                int32_t *__restrict__ active, int32_t *__restrict__ nfail,
                                int32_t *__restrict__ fail, int32_t *__restrict__ nspmv) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsub || !active[s]) return;
    nspmv[s] += 1;
    const double v = pq_in[s];
    pq[s] = v;
    if (!(v > 0.0) || !isfinite(v)) {
        active[s] = 0;
        fail[s] = 1;
        atomicAdd(nfail, 1);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void cg_xr_kernel(int64_t n, const T *__restrict__ p,
                                                    const T *__restrict__ q, T *__restrict__ x,
                                                    T *__restrict__ r, const double *__restrict__ rz,
                                                    const double *__restrict__ pq,
                                                    const int32_t *__restrict__ dsub,
                                                    const int32_t *__restrict__ active,
                                                    const double *__restrict__ w0, double *__restrict__ x0) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int s = dsub[i];
    if (!active[s]) return;
    const double a = rz[s] / pq[s];
    x[i] = T(double(x[i]) + a * double(p[i]));
    const T rn = T(double(r[i]) - a * double(q[i]));
    r[i] = rn;
    // the multigrid V-cycle's first smoothing x0 = W r of the new residual
    if (x0) x0[i] = w0[i] * double(rn);
}

// convergence test on |r|, then beta = (r,z)_new / (r,z)_old (also counts, per
// sub-problem, the preconditioner applies that visited it)
__global__ void cg_check_kernel(int nsub, const double *__restrict__ d2,
                                const double *__restrict__ bnorm, double rtol,
                                double *__restrict__ rz, double *__restrict__ beta,
                                int32_t *__restrict__ active, int32_t *__restrict__ nprec) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nsub || !active[s]) return;
    nprec[s] += 1;
    const double rr = d2[s], rzn = d2[nsub + s];
    if (!(sqrt(rr) > rtol * bnorm[s]) || !isfinite(rr)) {
        active[s] = 0;
        return;
    }
    beta[s] = rzn / rz[s];
    rz[s] = rzn;
}

template <typename T>
__global__ __launch_bounds__(256) void cg_p_kernel(int64_t n, const T *__restrict__ z,
                                                   T *__restrict__ p, const double *__restrict__ beta,
                                                   const int32_t *__restrict__ dsub,
                                                   const int32_t *__restrict__ active) {
    const int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= n) return;
    const int s = dsub[i];
    if (!active[s]) return;
    p[i] = T(double(z[i]) + beta[s] * double(p[i]));
}

template <typename T>
static int batched_dot2(pyhip_problem *p, KrylovWork *W, const T *a, const T *b, const T *c,
                        const T *d, double *out, const int32_t *act = nullptr) {
    TimeScope ts(&p->timer, "vector", p->st);
    const int pstride = W->m + 2;
    if (W->nchunk == 0) return PYHIP_OK;
    hipLaunchKernelGGL((dot2_kernel<T>), dim3((unsigned)W->nchunk), dim3(256), 0, p->st, a, b, c, d,
                       W->chunk_begin, W->chunk_end, W->partial, pstride, W->chunk_sub, act);
    KCHECK();
    hipLaunchKernelGGL(reduce_partials_kernel, dim3((unsigned)(2 * W->nsub)), dim3(64), 0, p->st, 2,
                       W->nsub, W->sub_chunk_ptr, W->partial, pstride, out, 0);
    KCHECK();
    return PYHIP_OK;
}

static int read_counter(pyhip_problem *p, KrylovWork *W, int32_t *dev, int32_t *out,
                        const int32_t *flags = nullptr, int32_t *flags_h = nullptr) {
    PYHIP_CHECK_HIP(hipMemcpyAsync(W->nactive_h, dev, sizeof(int32_t), hipMemcpyDeviceToHost, p->st));
    if (flags)
        PYHIP_CHECK_HIP(hipMemcpyAsync(flags_h, flags, W->nsub * sizeof(int32_t), hipMemcpyDeviceToHost, p->st));
    PYHIP_CHECK_HIP(hipStreamSynchronize(p->st));
    *out = *W->nactive_h;
    return PYHIP_OK;
}

// x = A^{-1} b per sub-problem with need[s]; CG in precision T on values TV, with any
// preconditioner (used for ILU(0); Jacobi goes through cg_worklist below).
template <typename T, typename TV>
static int cg_generic(pyhip_problem *p, KrylovWork *W, const TV *vals, const T *b, T *x,
              const pyhip_linear_opts &o, double rtol, const int32_t *need, int *iters, int *failed) {
    const int64_t n = p->ndof;
    const int ns = W->nsub;
    const int prec = std::is_same<T, float>::value ? PYHIP_F32 : PYHIP_F64;
    int rc;
    if ((rc = ensure_basis<T>(W, n, W->m + 1))) return rc;
    T *pv = (T *)W->V;
    T *qv = pv + n;
    T *r = std::is_same<T, float>::value ? (T *)W->w32 : (T *)W->w64;
    T *z = std::is_same<T, float>::value ? (T *)W->t32 : (T *)W->t64;
    double *d2 = W->dots;
    PYHIP_CHECK_HIP(hipMemsetAsync(x, 0, n * sizeof(T), p->st));
    hipLaunchKernelGGL((convert_kernel<T, T>), dim3(nblk(n)), dim3(256), 0, p->st, n, b, r);
    // Matrix and preconditioner work visits only the sub-problems being solved (need, then
    // those still iterating): the spmv and the multigrid V-cycle skip the rows of the
    // others, whose scalars the per-sub-problem kernels leave alone.
    int32_t *act_h = W->act_h;
    PYHIP_CHECK_HIP(hipMemcpyAsync(act_h, need, ns * sizeof(int32_t), hipMemcpyDeviceToHost, p->st));
    PYHIP_CHECK_HIP(hipStreamSynchronize(p->st));
    if ((rc = precond_apply<T>(p, W, o.precond, r, z, need, act_h))) return rc;
    if ((rc = batched_dot2<T>(p, W, r, r, r, z, d2))) return rc;
    PYHIP_CHECK_HIP(hipMemsetAsync(W->nactive_d, 0, sizeof(int32_t), p->st));
    PYHIP_CHECK_HIP(hipMemsetAsync(W->nfail_d, 0, sizeof(int32_t), p->st));
    PYHIP_CHECK_HIP(hipMemsetAsync(W->cg_fail, 0, ns * sizeof(int32_t), p->st));
    hipLaunchKernelGGL(cg_start_kernel, dim3(nblk(ns, 64)), dim3(64), 0, p->st, ns, d2, need, W->bnorm,
                       W->cg_rz, W->active, W->nactive_d);
    hipLaunchKernelGGL((convert_kernel<T, T>), dim3(nblk(n)), dim3(256), 0, p->st, n, z, pv);
    KCHECK();
    int32_t nact = 0;
    if ((rc = read_counter(p, W, W->nactive_d, &nact, W->active, act_h))) return rc;
    // with the multigrid preconditioner the x/r update also writes the V-cycle's first
    // smoothing W r of the finest level
    const double *w0 = nullptr;
    double *x0 = nullptr;
    if (std::is_same<T, double>::value && o.precond == PYHIP_PRECOND_AMG) {
        w0 = amg_fine_w(W->amg);
        x0 = amg_fine_x(W->amg);
    }
    // The host reads the running count every kPcgRead iterations: in between, a converged
    // sub-problem's flag is cleared on the device and every kernel skips its rows, so the
    // host keeps enqueueing instead of waiting. The spmv-family bytes of those launches
    // depend on which sub-problems were live: the launches add their per-launch part now
    // and the rows visited are counted per sub-problem on the device (spmv: nspmv, V-cycle:
    // nprec) and added at the end.
    int32_t *nspmv = W->cg_iters, *nprec = W->cg_iters2;
    hipLaunchKernelGGL(zero_i32_kernel, dim3(nblk(ns, 64)), dim3(64), 0, p->st, ns, nspmv);
    hipLaunchKernelGGL(zero_i32_kernel, dim3(nblk(ns, 64)), dim3(64), 0, p->st, ns, nprec);
    const bool defer_saved = p->timer.defer;
    p->timer.defer = true;
    int it = 0;
    while (nact > 0 && it < o.max_iters) {
        const int batch = std::min(kPcgRead, o.max_iters - it);
        for (int k = 0; k < batch; k++, it++) {
            if ((rc = spmv<TV, T, T>(p, vals, pv, qv, (const T *)nullptr, W->active, act_h))) return rc;
            if ((rc = batched_dots<T>(p, W, 1, pv, n, qv, d2 + 2 * ns, W->active))) return rc;
            hipLaunchKernelGGL(cg_alpha_kernel, dim3(nblk(ns, 64)), dim3(64), 0, p->st, ns, d2 + 2 * ns,
                               W->cg_pq, W->active, W->nfail_d, W->cg_fail, nspmv);
            {
                TimeScope ts(&p->timer, "vector", p->st);
                hipLaunchKernelGGL((cg_xr_kernel<T>), dim3(nblk(n)), dim3(256), 0, p->st, n, pv, qv, x, r,
                                   W->cg_rz, W->cg_pq, p->dof_sub, W->active, w0, x0);
            }
            KCHECK();
            if ((rc = precond_apply<T>(p, W, o.precond, r, z, W->active, act_h, x0 != nullptr))) return rc;
            if ((rc = batched_dot2<T>(p, W, r, r, r, z, d2, W->active))) return rc;
            hipLaunchKernelGGL(cg_check_kernel, dim3(nblk(ns, 64)), dim3(64), 0, p->st, ns, d2, W->bnorm, rtol,
                               W->cg_rz, W->cg_beta, W->active, nprec);
            {
                TimeScope ts(&p->timer, "vector", p->st);
                hipLaunchKernelGGL((cg_p_kernel<T>), dim3(nblk(n)), dim3(256), 0, p->st, n, z, pv, W->cg_beta,
                                   p->dof_sub, W->active);
            }
            KCHECK();
        }
        // the running count of this read: recount from the flags
        hipLaunchKernelGGL(count_active_kernel, dim3(1), dim3(256), 0, p->st, ns, W->active, W->nactive_d);
        if ((rc = read_counter(p, W, W->nactive_d, &nact, W->active, act_h))) return rc;
    }
    p->timer.defer = defer_saved;
    int32_t nf = 0;
    PYHIP_CHECK_HIP(hipMemcpyAsync(W->flags_h, W->cg_fail, ns * sizeof(int32_t), hipMemcpyDeviceToHost, p->st));
    PYHIP_CHECK_HIP(hipMemcpyAsync(W->flags_h + ns, nspmv, ns * sizeof(int32_t), hipMemcpyDeviceToHost, p->st));
    PYHIP_CHECK_HIP(hipMemcpyAsync(W->flags_h + 2 * ns, nprec, ns * sizeof(int32_t), hipMemcpyDeviceToHost, p->st));
    if ((rc = read_counter(p, W, W->nfail_d, &nf, W->active, act_h))) return rc;
    if (p->timer.on) {
        // rows visited: each spmv of the loop (y = A p: matrix, row starts, p, q) and each
        // V-cycle, per sub-problem
        std::vector<int64_t> vb;
        if (o.precond == PYHIP_PRECOND_AMG) vb = amg_vcycle_bytes(W->amg, (int)sizeof(T));
        int64_t bytes = 0;
        for (int s = 0; s < ns; s++) {
            const int64_t rows = p->dof_off_h[s + 1] - p->dof_off_h[s];
            bytes += (int64_t)W->flags_h[ns + s] *
                     ((int64_t)(sizeof(TV) + 4) * W->sub_nnz_h[s] + (int64_t)(4 + 2 * sizeof(T)) * rows);
            if (!vb.empty()) bytes += (int64_t)W->flags_h[2 * ns + s] * vb[s];
        }
        p->timer.add_work("spmv", bytes);
    }
    // per sub-problem: broke down, or still above tolerance; iterations of the longest-
    // running sub-problem (the loop may have run a few no-op iterations past the last one)
    W->lfail_h.assign(ns, 0);
    int maxit = 0;
    for (int s = 0; s < ns; s++) {
        W->lfail_h[s] = (W->flags_h[s] != 0 || act_h[s] != 0) ? 1 : 0;
        maxit = std::max(maxit, (int)W->flags_h[ns + s]);
    }
    *iters = maxit;
    *failed = nf + nact;   // breakdowns + sub-problems still above tolerance
    return PYHIP_OK;
}

// ---- single-reduction CG on a work list of active chunks (Jacobi preconditioner) -------
//
// Chronopoulos-Gear CG: one fused reduction per iteration instead of two. With u = M r
// and w = A u (M = diag(A)^-1):
//   gamma = (r, u), delta = (w, u), rho = (r, r)
//   beta = gamma / gamma_old (0 first), alpha = gamma / (delta - beta gamma / alpha_old)
//   p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s, u = M r
// which equals preconditioned CG in exact arithmetic. Per iteration: work-list spmv
// (w, delta partials), one reduction block per sub-problem (convergence, alpha, beta),
// and one fused vector update (gamma, rho partials).
//
// Work list: the chunks of the sub-problems still iterating, built by the host each time
// it reads the flags (every kCgCheck iterations). Between reads the list can only hold
// extra chunks, whose blocks see their sub-problem's flag cleared and return. Dot
// products are summed per 256-row block of a chunk (fixed tree), then per sub-problem
// over its blocks (strided lanes, fixed tree): a sub-problem gets the same bits alone and
// in any batch.

// Sum of one value per thread over a 256-thread block in a fixed order: xor tree within
// each 32-lane group, then the 8 group sums in increasing order. Result in thread 0.
__device__ __forceinline__ double block_sum(double v, double *sg) {
    v += __shfl_xor(v, 16, 32);
    v += __shfl_xor(v, 8, 32);
    v += __shfl_xor(v, 4, 32);
    v += __shfl_xor(v, 2, 32);
    v += __shfl_xor(v, 1, 32);
    if ((threadIdx.x & 31) == 0) sg[threadIdx.x >> 5] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int g = 0; g < kWlRows / 32; g++) s += sg[g];
    return s;
}

// Work-list block -> (chunk, part). Blocks are dispatched round-robin over the XCDs (block
// b runs on XCD b % kXcd), so each XCD is given a contiguous slice of the work list and
// the entries of the gathered vector its rows touch stay in that XCD's L2. The grid is
// wl_grid(nachunk). Rows [r0, r1); false when the block has none or its sub-problem has
// stopped iterating.
constexpr int kXcd = 8;

__device__ __forceinline__ bool wl_block(const int32_t *__restrict__ achunk, int nachunk,
                                         const int32_t *__restrict__ chunk_sub,
                                         const int32_t *__restrict__ active,
                                         const int64_t *__restrict__ cb, const int64_t *__restrict__ ce,
                                         int32_t &c, int &sb, int &part, int64_t &r0, int64_t &r1) {
    const int64_t A = (int64_t)nachunk * kWlBlocksPerChunk;
    const int64_t per = (A + kXcd - 1) / kXcd;
    const int64_t k = blockIdx.x / kXcd;
    if (k >= per) return false;
    const int64_t lb = (int64_t)(blockIdx.x % kXcd) * per + k;
    if (lb >= A) return false;
    c = achunk[lb / kWlBlocksPerChunk];
    sb = chunk_sub[c];
    if (!active[sb]) return false;
    part = (int)(lb % kWlBlocksPerChunk);
    r0 = cb[c] + (int64_t)part * kWlRows;
    r1 = ce[c];
    return r0 < r1;
}

static inline unsigned wl_grid(int64_t chunks) {
    const int64_t A = std::max<int64_t>(1, chunks * kWlBlocksPerChunk);
    return (unsigned)(((A + kXcd - 1) / kXcd) * kXcd);
}

// r = b, u = M b, x = p = s = 0; gamma and rho partials
template <typename T>
__global__ __launch_bounds__(256) void cg_init_kernel(const int32_t *__restrict__ achunk, int nachunk,
                                                      const int32_t *__restrict__ chunk_sub,
                                                      const int32_t *__restrict__ need,
                                                      const int64_t *__restrict__ cb,
                                                      const int64_t *__restrict__ ce,
                                                      const T *__restrict__ b, const T *__restrict__ jd,
                                                      T *__restrict__ x, T *__restrict__ r,
                                                      T *__restrict__ u, T *__restrict__ pv,
                                                      T *__restrict__ sv, double *__restrict__ part,
                                                      int64_t pstride) {
    int32_t c;
    int sb, part_i;
    int64_t r0, r1;
    if (!wl_block(achunk, nachunk, chunk_sub, need, cb, ce, c, sb, part_i, r0, r1)) return;
    const int64_t row = r0 + threadIdx.x;
    double g = 0.0, rho = 0.0;
    if (row < r1) {
        const T ri = b[row];
        const T ui = T(double(jd[row]) * double(ri));
        r[row] = ri;
        u[row] = ui;
        x[row] = T(0);
        pv[row] = T(0);
        sv[row] = T(0);
        g = double(ri) * double(ui);
        rho = double(ri) * double(ri);
    }
    __shared__ double sg[2][kWlRows / 32];
    g = block_sum(g, sg[0]);
    rho = block_sum(rho, sg[1]);
    if (threadIdx.x == 0) {
        const int64_t slot = (int64_t)c * kWlBlocksPerChunk + part_i;
        part[slot] = g;
        part[pstride + slot] = rho;
    }
}

// w = A u on the work list; delta = (w, u) partials
template <typename TV, typename T, int LPR>
__global__ __launch_bounds__(256) void cg_spmv_kernel(const int32_t *__restrict__ achunk, int nachunk,
                                                      const int32_t *__restrict__ chunk_sub,
                                                      const int32_t *__restrict__ active,
                                                      const int64_t *__restrict__ cb,
                                                      const int64_t *__restrict__ ce,
                                                      const int32_t *__restrict__ rs,
                                                      const int32_t *__restrict__ col,
                                                      const TV *__restrict__ val, const T *__restrict__ u,
                                                      T *__restrict__ w, double *__restrict__ part_delta) {
    __shared__ double sp[kWlRows];
    __shared__ double sg[kWlRows / 32];
    int32_t c;
    int sb, part_i;
    int64_t r0, r1;
    if (!wl_block(achunk, nachunk, chunk_sub, active, cb, ce, c, sb, part_i, r0, r1)) return;
    constexpr int RPP = kWlRows / LPR;
    T v[LPR];
    bool live[LPR];
    rows_dot<TV, T, T, LPR>(r0, r1, rs, col, val, u, nullptr, nullptr, v, live);
    if (threadIdx.x % LPR == 0) {
#pragma unroll
        for (int p = 0; p < LPR; p++) {
            const int loc = p * RPP + threadIdx.x / LPR;
            const int64_t row = r0 + loc;
            double d = 0.0;
            if (row < r1) {
                w[row] = v[p];
                d = double(u[row]) * double(v[p]);
            }
            sp[loc] = d;
        }
    }
    __syncthreads();
    const double d = block_sum(sp[threadIdx.x], sg);
    if (threadIdx.x == 0) part_delta[(int64_t)c * kWlBlocksPerChunk + part_i] = d;
}

// One 64-lane block per sub-problem: sums its gamma / rho / delta block partials and
// updates its scalars. mode 0: start (bnorm, gamma; active = need and r != 0); mode 1:
// first iteration (beta = 0); mode 2: later iterations. Counts the iterations taken.
__global__ __launch_bounds__(64) void cg_reduce_kernel(int mode, double rtol, const int32_t *__restrict__ scp,
                                                       const double *__restrict__ part, int64_t pstride,
                                                       const int32_t *__restrict__ need,
                                                       double *__restrict__ bnorm, double *__restrict__ gam,
                                                       double *__restrict__ alpha, double *__restrict__ beta,
                                                       int32_t *__restrict__ active, int32_t *__restrict__ iters,
                                                       int32_t *__restrict__ fail) {
    const int sb = blockIdx.x, t = threadIdx.x;
    const bool run = mode == 0 ? need[sb] != 0 : active[sb] != 0;
    if (!run) {
        if (mode == 0 && t == 0) {
            active[sb] = 0;
            iters[sb] = 0;
            fail[sb] = 0;
        }
        return;
    }
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    const int64_t g1 = (int64_t)scp[sb + 1] * kWlBlocksPerChunk;
    for (int64_t g = (int64_t)scp[sb] * kWlBlocksPerChunk + t; g < g1; g += 64) {
        a0 += part[g];
        a1 += part[pstride + g];
        a2 += part[2 * pstride + g];
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a0 += __shfl_xor(a0, off, 64);
        a1 += __shfl_xor(a1, off, 64);
        a2 += __shfl_xor(a2, off, 64);
    }
    if (t != 0) return;
    const double g = a0, rho = a1, dl = a2;
    int a = 1;
    if (mode == 0) {
        bnorm[sb] = sqrt(rho);
        gam[sb] = g;
        a = rho > 0.0 && isfinite(rho);
        iters[sb] = 0;
        fail[sb] = 0;
    } else if (!isfinite(rho) || !isfinite(g)) {
        a = 0;
        fail[sb] = 1;
    } else if (!(sqrt(rho) > rtol * bnorm[sb])) {
        a = 0;   // converged
    } else {
        const double bt = mode == 1 ? 0.0 : g / gam[sb];
        const double den = mode == 1 ? dl : dl - bt * g / alpha[sb];
        if (!(g > 0.0) || !(den > 0.0) || !isfinite(den)) {
            a = 0;   // indefinite system or preconditioner
            fail[sb] = 1;
        } else {
            alpha[sb] = g / den;
            beta[sb] = bt;
            gam[sb] = g;
            iters[sb] = iters[sb] + 1;
        }
    }
    active[sb] = a;
}

// p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s, u = M r; gamma, rho partials
template <typename T>
__global__ __launch_bounds__(256) void cg_update_kernel(const int32_t *__restrict__ achunk, int nachunk,
                                                        const int32_t *__restrict__ chunk_sub,
                                                        const int32_t *__restrict__ active,
                                                        const int64_t *__restrict__ cb,
                                                        const int64_t *__restrict__ ce,
                                                        const double *__restrict__ alpha,
                                                        const double *__restrict__ beta,
                                                        const T *__restrict__ jd, const T *__restrict__ w,
                                                        T *__restrict__ u, T *__restrict__ pv,
                                                        T *__restrict__ sv, T *__restrict__ x,
                                                        T *__restrict__ r, double *__restrict__ part,
                                                        int64_t pstride) {
    int32_t c;
    int sb, part_i;
    int64_t r0, r1;
    if (!wl_block(achunk, nachunk, chunk_sub, active, cb, ce, c, sb, part_i, r0, r1)) return;
    const double a = alpha[sb], bt = beta[sb];
    const int64_t row = r0 + threadIdx.x;
    double g = 0.0, rho = 0.0;
    if (row < r1) {
        const T pi = T(double(u[row]) + bt * double(pv[row]));
        const T si = T(double(w[row]) + bt * double(sv[row]));
        x[row] = T(double(x[row]) + a * double(pi));
        const T ri = T(double(r[row]) - a * double(si));
        const T ui = T(double(jd[row]) * double(ri));
        pv[row] = pi;
        sv[row] = si;
        r[row] = ri;
        u[row] = ui;
        g = double(ri) * double(ui);
        rho = double(ri) * double(ri);
    }
    __shared__ double sg[2][kWlRows / 32];
    g = block_sum(g, sg[0]);
    rho = block_sum(rho, sg[1]);
    if (threadIdx.x == 0) {
        const int64_t slot = (int64_t)c * kWlBlocksPerChunk + part_i;
        part[slot] = g;
        part[pstride + slot] = rho;
    }
}

template <typename TV, typename T>
static void launch_cg_spmv(int lpr, hipStream_t st, unsigned grid, int nachunk, KrylovWork *W,
                           pyhip_problem *p, const TV *vals, const T *u, T *w, double *pd) {
    const dim3 g(grid), t(256);
#define CG_SPMV(L)                                                                                  \
    hipLaunchKernelGGL((cg_spmv_kernel<TV, T, L>), g, t, 0, st, W->achunk, nachunk, W->chunk_sub,    \
                       W->active, W->chunk_begin, W->chunk_end, p->row_start, p->col, vals, u, w, pd)
    switch (lpr) {
    case 1: CG_SPMV(1); break;
    case 2: CG_SPMV(2); break;
    case 4: CG_SPMV(4); break;
    default: CG_SPMV(8); break;
    }
#undef CG_SPMV
}

// Host side of the work list: reads flags[] (ns int32 from dev) and uploads the chunks of
// the sub-problems with a non-zero flag. Returns the number of chunks.
static int cg_refresh(pyhip_problem *p, KrylovWork *W, const int32_t *flags_dev, bool with_stats,
                      int *nachunk, int *nactive) {
    const int ns = W->nsub;
    int32_t *f = W->flags_h;
    PYHIP_CHECK_HIP(hipMemcpyAsync(f, flags_dev, ns * sizeof(int32_t), hipMemcpyDeviceToHost, p->st));
    if (with_stats) {
        PYHIP_CHECK_HIP(hipMemcpyAsync(f + ns, W->cg_fail, ns * sizeof(int32_t), hipMemcpyDeviceToHost, p->st));
        PYHIP_CHECK_HIP(hipMemcpyAsync(f + 2 * ns, W->cg_iters, ns * sizeof(int32_t), hipMemcpyDeviceToHost, p->st));
    }
    PYHIP_CHECK_HIP(hipStreamSynchronize(p->st));
    // the stream is idle here, so the pinned list can be rewritten
    int n = 0, na = 0;
    for (int s = 0; s < ns; s++) {
        if (!f[s]) continue;
        na++;
        const int64_t c0 = W->chunk_ptr_h[s], c1 = W->chunk_ptr_h[s + 1];
        for (int64_t c = c0; c < c1; c++) W->achunk_h[n++] = (int32_t)c;
    }
    if (n) PYHIP_CHECK_HIP(hipMemcpyAsync(W->achunk, W->achunk_h, n * sizeof(int32_t), hipMemcpyHostToDevice, p->st));
    *nachunk = n;
    *nactive = na;
    return PYHIP_OK;
}

// x = A^{-1} b for the sub-problems with need[s], Jacobi-preconditioned single-reduction
// CG in precision T on values TV. The host refreshes the work list every kCgCheck
// iterations.
constexpr int kCgCheck = 8;

template <typename T, typename TV>
static int cg_worklist(pyhip_problem *p, KrylovWork *W, const TV *vals, const T *b, T *x,
                       const pyhip_linear_opts &o, double rtol, const int32_t *need, int *iters,
                       int *failed) {
    const int64_t n = p->ndof;
    const int ns = W->nsub;
    int rc;
    *iters = 0;
    *failed = 0;
    W->lfail_h.assign(ns, 0);
    if (n == 0 || W->nchunk == 0) return PYHIP_OK;
    PYHIP_REQUIRE(W->achunk && W->achunk_h && W->gpart && W->cg_iters && W->cg_fail && W->flags_h &&
                      W->active && need && W->sub_chunk_ptr && W->chunk_begin && W->chunk_end &&
                      W->chunk_sub && (int)W->chunk_ptr_h.size() == ns + 1,
                  "CG work-list buffers missing");
    if ((rc = ensure_basis<T>(W, n, std::max(W->m + 1, 3)))) return rc;
    T *pv = (T *)W->V, *sv = pv + n, *wv = pv + 2 * n;
    T *r = std::is_same<T, float>::value ? (T *)W->w32 : (T *)W->w64;
    T *u = std::is_same<T, float>::value ? (T *)W->t32 : (T *)W->t64;
    const T *jd = std::is_same<T, float>::value ? (const T *)W->jdiag32 : (const T *)W->jdiag64;
    if ((rc = spmv_tune<TV, T, T>(p, W, vals, b, wv))) return rc;
    const int64_t pstride = W->nchunk * kWlBlocksPerChunk;
    double *pg = W->gpart, *pd = W->gpart + 2 * pstride;
    int nachunk = 0, nactive = 0;
    if ((rc = cg_refresh(p, W, need, false, &nachunk, &nactive))) return rc;
    if (nachunk) {
        TimeScope ts(&p->timer, "vector", p->st);
        hipLaunchKernelGGL((cg_init_kernel<T>), dim3(wl_grid(nachunk)), dim3(256), 0, p->st, W->achunk, nachunk,
                           W->chunk_sub, need, W->chunk_begin, W->chunk_end, b, jd, x, r, u, pv, sv, pg, pstride);
    }
    hipLaunchKernelGGL(cg_reduce_kernel, dim3(ns), dim3(64), 0, p->st, 0, rtol, W->sub_chunk_ptr, W->gpart,
                       pstride, need, W->bnorm, W->cg_rz, W->cg_pq, W->cg_beta, W->active, W->cg_iters,
                       W->cg_fail);
    KCHECK();
    if ((rc = cg_refresh(p, W, W->active, false, &nachunk, &nactive))) return rc;
    const std::vector<int32_t> started(W->flags_h, W->flags_h + ns);
    int it = 0;
    int64_t rows_done = 0, nnz_done = 0;
    while (nactive > 0 && it < o.max_iters) {
        const unsigned grid = wl_grid(nachunk);
        const int batch = std::min(kCgCheck, o.max_iters - it);
        for (int k = 0; k < batch; k++, it++) {
            {
                TimeScope ts(&p->timer, "spmv", p->st);
                launch_cg_spmv<TV, T>(W->lpr, p->st, grid, nachunk, W, p, vals, u, wv, pd);
            }
            TimeScope ts(&p->timer, "vector", p->st);
            hipLaunchKernelGGL(cg_reduce_kernel, dim3(ns), dim3(64), 0, p->st, it == 0 ? 1 : 2, rtol,
                               W->sub_chunk_ptr, W->gpart, pstride, need, W->bnorm, W->cg_rz, W->cg_pq,
                               W->cg_beta, W->active, W->cg_iters, W->cg_fail);
            hipLaunchKernelGGL((cg_update_kernel<T>), dim3(grid), dim3(256), 0, p->st, W->achunk, nachunk,
                               W->chunk_sub, W->active, W->chunk_begin, W->chunk_end, W->cg_pq, W->cg_beta,
                               jd, wv, u, pv, sv, x, r, pg, pstride);
        }
        KCHECK();
        if ((rc = cg_refresh(p, W, W->active, false, &nachunk, &nactive))) return rc;
    }
    // statistics: per-sub iterations (each one spmv over the sub-problem's rows) and failures
    int nfail = 0, maxit = 0;
    if ((rc = cg_refresh(p, W, W->active, true, &nachunk, &nactive))) return rc;
    for (int s = 0; s < ns; s++) {
        const int its = W->flags_h[2 * ns + s];
        nfail += W->flags_h[ns + s] != 0;
        maxit = std::max(maxit, its);
        // every iteration is an spmv then a reduction; a sub-problem stopped by its
        // reduction (converged or failed) ran one spmv more than its iteration count
        const int64_t sp = its + (started[s] && !W->flags_h[s] ? 1 : 0);
        rows_done += sp * (p->dof_off_h[s + 1] - p->dof_off_h[s]);
        nnz_done += sp * W->sub_nnz_h[s];
    }
    p->timer.add_work("spmv", (int64_t)(sizeof(TV) + 4) * nnz_done + (int64_t)(4 + 2 * sizeof(T)) * rows_done);
    W->lfail_h.assign(ns, 0);
    for (int s = 0; s < ns; s++) W->lfail_h[s] = (W->flags_h[ns + s] != 0 || W->flags_h[s] != 0) ? 1 : 0;
    if (debug_enabled())
        fprintf(stderr, "[pyhip] cg: %d iterations (longest sub-problem %d), %d still active, %d failed\n", it,
                maxit, nactive, nfail);
    *iters = it;
    *failed = nfail + nactive;
    return PYHIP_OK;
}

}  // namespace pyhip

// ------------------------------------------------------------------------------------
// Newton

namespace pyhip {

__global__ void newton_state_kernel(int nsub, const int32_t *__restrict__ running,
                                    int32_t *__restrict__ mask) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < nsub) mask[s] = running[s];
}

// Solve J d = r for the masked sub-problems. r: FP64 right-hand side; d: FP64 result.
template <typename T, typename TV>
static int krylov(pyhip_problem *p, KrylovWork *W, const TV *vals, const T *b, T *x,
                  const pyhip_linear_opts &o, double rtol, double stag, const int32_t *need, int *iters,
                  int *failed) {
    if (o.krylov == PYHIP_KRYLOV_CG && o.precond == PYHIP_PRECOND_JACOBI)
        return cg_worklist<T, TV>(p, W, vals, b, x, o, rtol, need, iters, failed);
    if (o.krylov == PYHIP_KRYLOV_CG) return cg_generic<T, TV>(p, W, vals, b, x, o, rtol, need, iters, failed);
    return gmres<T, TV>(p, W, vals, b, x, o, rtol, stag, need, iters, failed);
}

// An FP32 inner GMRES of the refinement stops a sub-problem whose restart cycle reduced the
// residual by less than this factor (FP32 floor reached; the FP64 refinement continues).
constexpr double kInnerStagnation = 0.9;

static int linear_solve(pyhip_problem *p, KrylovWork *W, const pyhip_linear_opts &o, int jac_prec,
                        const double *r, double *d, const int32_t *mask, int *iters,
                        int *refinements, int *failed) {
    int rc;
    const int64_t n = p->ndof;
    *iters = 0;
    *refinements = 0;
    *failed = 0;
    if (o.precision == PYHIP_F64) {
        PYHIP_REQUIRE(jac_prec == PYHIP_F64, "an FP64 Krylov solve needs the FP64 Jacobian");
        W->ilu_order = o.ilu_order;
        if ((rc = precond_setup<double, double>(p, W, o.precond, W->vals64, mask, o.amg_element_level != 0)))
            return rc;
        return krylov<double, double>(p, W, W->vals64, r, d, o, o.rtol, 0.0, mask, iters, failed);
    }
    // FP32 inner solves inside FP64 iterative refinement:
    //   d = 0; repeat: s = A32^{-1} rr (FP32, to rtol_f32); d += s; rr = r - J d (FP64)
    if (jac_prec == PYHIP_F64) {
        hipLaunchKernelGGL((convert_kernel<double, float>), dim3(nblk(p->nnz)), dim3(256), 0, p->st,
                           p->nnz, W->vals64, W->vals32);
        KCHECK();
    }
    W->ilu_order = o.ilu_order;
    W->amg_src64 = jac_prec == PYHIP_F64 ? W->vals64 : nullptr;
    if ((rc = precond_setup<float, float>(p, W, o.precond, W->vals32, mask, o.amg_element_level != 0))) return rc;
    PYHIP_CHECK_HIP(hipMemsetAsync(d, 0, n * sizeof(double), p->st));
    hipLaunchKernelGGL((convert_kernel<double, double>), dim3(nblk(n)), dim3(256), 0, p->st, n, r, W->r64);
    KCHECK();
    const int ns = W->nsub;
    double *d3 = W->dots + 2 * (W->m + 2) * ns;
    if ((rc = batched_dots<double>(p, W, 1, r, n, r, W->dots))) return rc;
    std::vector<double> rn0(ns), rn(ns);
    PYHIP_CHECK_HIP(hipMemcpyAsync(rn0.data(), W->dots, ns * sizeof(double), hipMemcpyDeviceToHost, p->st));
    std::vector<int32_t> need_h(ns);
    PYHIP_CHECK_HIP(hipMemcpyAsync(need_h.data(), mask, ns * sizeof(int32_t), hipMemcpyDeviceToHost, p->st));
    PYHIP_CHECK_HIP(hipStreamSynchronize(p->st));
    PYHIP_CHECK_HIP(hipMemcpyAsync(W->need, need_h.data(), ns * sizeof(int32_t), hipMemcpyHostToDevice, p->st));
    int left = 0;
    for (int it = 0; it < std::max(1, o.max_refinements); it++) {
        hipLaunchKernelGGL((convert_kernel<double, float>), dim3(nblk(n)), dim3(256), 0, p->st, n, W->r64, W->r32);
        KCHECK();
        int its = 0, f = 0;
        if ((rc = krylov<float, float>(p, W, W->vals32, W->r32, W->d32, o, o.rtol_f32, kInnerStagnation,
                                       W->need, &its, &f)))
            return rc;
        *iters += its;
        (*refinements)++;
        hipLaunchKernelGGL((axpy_kernel<float, double>), dim3(nblk(n)), dim3(256), 0, p->st, n, W->d32, d, 1.0,
                           p->dof_sub, W->need);
        // rr = r - J d  (FP64 arithmetic on the stored Jacobian)
        if (jac_prec == PYHIP_F64) rc = spmv<double, double, double>(p, W->vals64, d, W->r64, r);
        else rc = spmv<float, double, double>(p, W->vals32, d, W->r64, r);
        if (rc) return rc;
        if ((rc = batched_dots<double>(p, W, 1, W->r64, n, W->r64, d3))) return rc;
        PYHIP_CHECK_HIP(hipMemcpyAsync(rn.data(), d3, ns * sizeof(double), hipMemcpyDeviceToHost, p->st));
        PYHIP_CHECK_HIP(hipStreamSynchronize(p->st));
        left = 0;
        for (int s = 0; s < ns; s++) {
            if (!need_h[s]) continue;
            if (!(std::sqrt(rn[s]) > o.refine_rtol * std::sqrt(rn0[s]))) need_h[s] = 0;
            left += need_h[s];
        }
        if (!left) break;
        PYHIP_CHECK_HIP(hipMemcpyAsync(W->need, need_h.data(), ns * sizeof(int32_t), hipMemcpyHostToDevice, p->st));
    }
    W->lfail_h = need_h;   // still above the refinement tolerance
    *failed = left;
    return PYHIP_OK;
}

int plugin_setup(pyhip_problem *p, const pyhip_linear_opts &o) {
    int rc = krylov_setup(p, o.restart);
    if (rc) return rc;
    KrylovWork *W = p->kw;
    if (!W->vals64 && (rc = dev_alloc(&W->vals64, p->nnz, W->bytes))) return rc;
    return PYHIP_OK;
}
double *plugin_values(pyhip_problem *p) { return p->kw->vals64; }
int plugin_factor(pyhip_problem *p, const pyhip_linear_opts &o) {
    p->kw->ilu_order = o.ilu_order;
    return precond_setup<double, double>(p, p->kw, o.precond, p->kw->vals64, nullptr, o.amg_element_level != 0);
}
int plugin_solve(pyhip_problem *p, const pyhip_linear_opts &o, const double *b, double *x, int *iters) {
    KrylovWork *W = p->kw;
    int32_t one = 1;
    int failed = 0;
    PYHIP_CHECK_HIP(hipMemcpy(W->dof_mask, &one, sizeof(int32_t), hipMemcpyHostToDevice));
    int rc = krylov<double, double>(p, W, W->vals64, b, x, o, o.rtol, 0.0, W->dof_mask, iters, &failed);
    if (rc) return rc;
    if (failed) {
        set_error("Krylov solver did not reach its tolerance");
        return PYHIP_ERR_NOT_CONVERGED;
    }
    return PYHIP_OK;
}

}  // namespace pyhip

using namespace pyhip;

extern "C" {

void pyhip_linear_opts_default(pyhip_linear_opts *o) {
    o->krylov = PYHIP_KRYLOV_GMRES;
    o->precond = PYHIP_PRECOND_ILU0;
    o->precision = PYHIP_F64;
    o->restart = 50;
    o->max_iters = 5000;
    o->rtol = 1e-10;
    o->rtol_f32 = 1e-5;
    o->max_refinements = 8;
    o->refine_rtol = 1e-10;
    o->ilu_order = PYHIP_ILU_ORDER_AUTO;
    o->amg_element_level = 1;
}

void pyhip_newton_opts_default(pyhip_newton_opts *o) {
    o->tol = 1e-8;
    o->max_iter = 10;
    o->max_residuals = 1e10;
    o->always_take_one_newton_step = 1;
    o->relaxation = 1.0;
    o->jacobian_precision = PYHIP_F64;
    pyhip_linear_opts_default(&o->linear);
    o->correction_tol = 0.0;
}

int64_t pyhip_krylov_bytes(const pyhip_problem *p) { return p->kw ? p->kw->bytes : 0; }

void pyhip_krylov_free(pyhip_problem *p) {
    KrylovWork *W = p->kw;
    if (!W) return;
    void *ptrs[] = {W->chunk_sub, W->chunk_begin, W->chunk_end, W->sub_chunk_ptr, W->partial, W->dots,
                    W->H, W->cs, W->sn, W->g, W->beta0, W->bnorm, W->kcols, W->active, W->nactive_d,
                    W->y, W->V, W->r64, W->x64, W->d64, W->b64, W->r32, W->d32, W->w32, W->t32,
                    W->w64, W->t64, W->vals32, W->vals64, W->res64, W->maxres, W->sub_scalar,
                    W->jdiag64, W->jdiag32, W->dof_mask, W->need, W->cg_rz, W->cg_pq,
                    W->cg_beta, W->nfail_d, W->ilu.perm, W->ilu.rs, W->ilu.col,
                    W->ilu.src, W->ilu.diag, W->ilu.lo_rows, W->ilu.up_rows, W->ilu.lu,
                    W->ilu.tmp64, W->ilu.tmp64b, W->ilu.tmp32, W->ilu.tmp32b, W->achunk,
                    W->gpart, W->cg_iters, W->cg_iters2, W->cg_fail, W->ac_cur, W->ac_der, W->ac_y, W->ac_z,
                    W->ac_t, W->ac_zs, W->ac_alpha, W->ac_mask, W->amg_conv64};
    for (void *q : ptrs)
        dfree(q);
    if (W->nactive_h) (void)hipHostFree(W->nactive_h);
    if (W->act_h) (void)hipHostFree(W->act_h);
    if (W->maxres_h) (void)hipHostFree(W->maxres_h);
    amg_free(W->amg);
    blk_free(W);
    if (W->achunk_h) (void)hipHostFree(W->achunk_h);
    if (W->flags_h) (void)hipHostFree(W->flags_h);
    delete W;
    p->kw = nullptr;
}

int pyhip_newton_solve(pyhip_problem *p, double *x, const pyhip_newton_opts *o,
                       pyhip_newton_stats *stats) {
    PYHIP_REQUIRE(p && x && o && stats, "null argument");
    PYHIP_REQUIRE(o->linear.restart >= 1 && o->linear.restart <= kMaxRestart,
                  "restart must be in 1..1000");
    PYHIP_REQUIRE(o->jacobian_precision == PYHIP_F64 || o->jacobian_precision == PYHIP_F32,
                  "bad Jacobian precision");
    int rc;
    if ((rc = krylov_setup(p, o->linear.restart))) return rc;
    KrylovWork *W = p->kw;
    W->amg_stale = true;
    W->step_res_ok.assign(W->step_res_ok.size(), 0);
    const int ns = p->nsub;
    const int64_t n = p->ndof;
    if (o->jacobian_precision == PYHIP_F64 && !W->vals64)
        if ((rc = dev_alloc(&W->vals64, p->nnz, W->bytes))) return rc;
    if (o->linear.precision == PYHIP_F32 && !W->vals32)
        if ((rc = dev_alloc(&W->vals32, p->nnz, W->bytes))) return rc;
    memset(stats, 0, sizeof(pyhip_newton_stats) * ns);
    const auto t_start = std::chrono::steady_clock::now();
    auto elapsed_ms = [&]() {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    };
    std::vector<int32_t> running(ns, 1);
    std::vector<int> count(ns, 0);
    std::vector<double> dmax(ns, 0.0), xmax(ns, 0.0);
    auto record = [&](int s, double v) {
        auto &S = stats[s];
        if (S.nhistory < 16) S.history[S.nhistory++] = v;
        S.final_maxres = v;
    };
    auto residual_norms = [&]() -> int {
        int r2;
        if ((r2 = pyhip_assemble(p, x, 0, -1, PYHIP_F64, W->res64, nullptr))) return r2;
        if ((r2 = batched_maxabs(p, W->res64, W->maxres))) return r2;
        PYHIP_CHECK_HIP(hipMemcpyAsync(W->maxres_h, W->maxres, ns * sizeof(double), hipMemcpyDeviceToHost, p->st));
        PYHIP_CHECK_HIP(hipStreamSynchronize(p->st));
        return PYHIP_OK;
    };
    // first residual (count == 1)
    if ((rc = residual_norms())) return rc;
    for (int s = 0; s < ns; s++) {
        record(s, W->maxres_h[s]);
        if (p->dof_off_h[s + 1] == p->dof_off_h[s]) running[s] = 0, stats[s].converged = 1;
        else if (W->maxres_h[s] < o->tol && !o->always_take_one_newton_step)
            running[s] = 0, stats[s].converged = 1;
        if (!running[s]) stats[s].done_ms = elapsed_ms();
    }
    int any_failed = 0;
    while (true) {
        int nrun = 0;
        for (int s = 0; s < ns; s++) nrun += running[s];
        if (!nrun) break;
        for (int s = 0; s < ns; s++) count[s] += running[s];
        PYHIP_CHECK_HIP(hipMemcpyAsync(W->dof_mask, running.data(), ns * sizeof(int32_t), hipMemcpyHostToDevice, p->st));
        // Jacobian at x (its residual is the one computed last, bit-identical)
        if (o->jacobian_precision == PYHIP_F64) {
            if ((rc = pyhip_assemble(p, x, 1, -1, PYHIP_F64, W->w64, W->vals64))) return rc;
        } else {
            if ((rc = pyhip_assemble(p, x, 1, -1, PYHIP_F32, W->w64, W->vals32))) return rc;
        }
        int its = 0, refs = 0, lfail = 0;
        if ((rc = linear_solve(p, W, o->linear, o->jacobian_precision, W->res64, W->d64, W->dof_mask,
                               &its, &refs, &lfail)))
            return rc;
        for (int s = 0; s < ns; s++)
            if (running[s]) {
                stats[s].linear_iters += its;
                stats[s].linear_failures += (s < (int)W->lfail_h.size() && W->lfail_h[s]) ? 1 : 0;
            }
        hipLaunchKernelGGL((axpy_kernel<double, double>), dim3(nblk(n)), dim3(256), 0, p->st, n, W->d64, x,
                           -o->relaxation, p->dof_sub, W->dof_mask);
        KCHECK();
        if (o->correction_tol > 0.0) {
            // max |dx| and max |x| per sub-problem (FP64)
            if ((rc = batched_maxabs(p, W->d64, W->sub_scalar))) return rc;
            PYHIP_CHECK_HIP(hipMemcpyAsync(dmax.data(), W->sub_scalar, ns * sizeof(double), hipMemcpyDeviceToHost,
                                           p->st));
            if ((rc = batched_maxabs(p, x, W->sub_scalar))) return rc;
            PYHIP_CHECK_HIP(hipMemcpyAsync(xmax.data(), W->sub_scalar, ns * sizeof(double), hipMemcpyDeviceToHost,
                                           p->st));
        }
        if ((rc = residual_norms())) return rc;
        for (int s = 0; s < ns; s++) {
            if (!running[s]) continue;
            const double mr = W->maxres_h[s];
            record(s, mr);
            stats[s].newton_iters = count[s];
            const bool small_update = o->correction_tol <= 0.0 ||
                                      dmax[s] * std::fabs(o->relaxation) <= o->correction_tol * xmax[s];
            if (mr < o->tol && small_update) {
                running[s] = 0;
                stats[s].converged = 1;
            } else if (mr > o->max_residuals || count[s] >= o->max_iter) {
                running[s] = 0;
                stats[s].converged = 0;
                any_failed = 1;
            }
            if (!running[s]) stats[s].done_ms = elapsed_ms();
        }
    }
    if (any_failed) {
        set_error("NewtonSolverError: a sub-problem did not converge (see stats)");
        return PYHIP_ERR_NEWTON;
    }
    return PYHIP_OK;
}

// One Newton iteration (declared in pyoomph_hip.h) for the sub-problems with mask[s] that
// have neither converged nor failed: the body of pyhip_newton_solve's loop, with each
// sub-problem's progress kept by the caller between calls, so that sub-problems can start
// and finish at any call (slots of a serving loop).
int pyhip_newton_step(pyhip_problem *p, double *x, const pyhip_newton_opts *o, const int32_t *mask,
                      pyhip_newton_step_state *st) {
    PYHIP_REQUIRE(p && x && o && mask && st, "null argument");
    PYHIP_REQUIRE(o->linear.restart >= 1 && o->linear.restart <= kMaxRestart, "restart must be in 1..1000");
    PYHIP_REQUIRE(o->jacobian_precision == PYHIP_F64 || o->jacobian_precision == PYHIP_F32,
                  "bad Jacobian precision");
    int rc;
    if ((rc = krylov_setup(p, o->linear.restart))) return rc;
    KrylovWork *W = p->kw;
    const int ns = p->nsub;
    const int64_t n = p->ndof;
    if (o->jacobian_precision == PYHIP_F64 && !W->vals64)
        if ((rc = dev_alloc(&W->vals64, p->nnz, W->bytes))) return rc;
    if (o->linear.precision == PYHIP_F32 && !W->vals32)
        if ((rc = dev_alloc(&W->vals32, p->nnz, W->bytes))) return rc;
    std::vector<int32_t> running(ns, 0);
    int nrun = 0;
    bool need_res = false;
    W->step_res_ok.resize(ns, 0);
    for (int s = 0; s < ns; s++) {
        const bool empty = p->dof_off_h[s + 1] == p->dof_off_h[s];
        if (!mask[s] || st[s].converged || st[s].failed) continue;
        if (empty) {
            st[s].converged = 1;
            continue;
        }
        running[s] = 1;
        nrun++;
        need_res = need_res || st[s].newton_iters == 0 || !W->step_res_ok[s];
    }
    if (!nrun) return PYHIP_OK;
    // Residuals, Jacobians and solves cover the running sub-problems only (W->dof_mask):
    // the element kernels skip the others' blocks, the Krylov kernels their rows.
    auto upload_running = [&]() -> int {
        PYHIP_CHECK_HIP(hipMemcpyAsync(W->dof_mask, running.data(), ns * sizeof(int32_t), hipMemcpyHostToDevice,
                                       p->st));
        return PYHIP_OK;
    };
    auto residual_norms = [&]() -> int {
        int r2;
        if ((r2 = assemble_subset(p, x, 0, -1, PYHIP_F64, W->res64, nullptr, running.data()))) return r2;
        if ((r2 = batched_maxabs(p, W->res64, W->maxres))) return r2;
        PYHIP_CHECK_HIP(hipMemcpyAsync(W->maxres_h, W->maxres, ns * sizeof(double), hipMemcpyDeviceToHost, p->st));
        PYHIP_CHECK_HIP(hipStreamSynchronize(p->st));
        for (int s = 0; s < ns; s++) W->step_res_ok[s] = (uint8_t)running[s];
        return PYHIP_OK;
    };
    if ((rc = upload_running())) return rc;
    // a sub-problem starting its solve: the first residual (count 1: without
    // always_take_one_newton_step a small enough start is already converged); otherwise
    // the residual the call that last stepped it ended with is still the one at its dofs
    if (need_res && (rc = residual_norms())) return rc;
    bool dropped = false;
    for (int s = 0; s < ns; s++) {
        if (!running[s] || st[s].newton_iters) continue;
        st[s].maxres = W->maxres_h[s];
        if (W->maxres_h[s] < o->tol && !o->always_take_one_newton_step) {
            st[s].converged = 1;
        } else if (!std::isfinite(W->maxres_h[s])) {
            // a non-finite start residual (non-finite data) cannot converge; failing it here
            // keeps its Jacobian out of the shared preconditioner set-up, whose breakdown
            // would stop every sub-problem of the call
            st[s].failed = 1;
        } else {
            continue;
        }
        running[s] = 0;
        nrun--;
        dropped = true;
    }
    if (!nrun) return PYHIP_OK;
    // the multigrid hierarchy comes from this step's Jacobian, so that no sub-problem's
    // preconditioner depends on when the others started their solves
    W->amg_stale = true;
    if (dropped && (rc = upload_running())) return rc;
    if (o->jacobian_precision == PYHIP_F64) {
        if ((rc = assemble_subset(p, x, 1, -1, PYHIP_F64, W->w64, W->vals64, running.data()))) return rc;
    } else {
        if ((rc = assemble_subset(p, x, 1, -1, PYHIP_F32, W->w64, W->vals32, running.data()))) return rc;
    }
    int its = 0, refs = 0, lfail = 0;
    if ((rc = linear_solve(p, W, o->linear, o->jacobian_precision, W->res64, W->d64, W->dof_mask, &its, &refs,
                           &lfail)))
        return rc;
    hipLaunchKernelGGL((axpy_kernel<double, double>), dim3(nblk(n)), dim3(256), 0, p->st, n, W->d64, x,
                       -o->relaxation, p->dof_sub, W->dof_mask);
    KCHECK();
    std::vector<double> dmax(ns, 0.0), xmax(ns, 0.0);
    if ((rc = batched_maxabs(p, W->d64, W->sub_scalar))) return rc;
    PYHIP_CHECK_HIP(hipMemcpyAsync(dmax.data(), W->sub_scalar, ns * sizeof(double), hipMemcpyDeviceToHost, p->st));
    if ((rc = batched_maxabs(p, x, W->sub_scalar))) return rc;
    PYHIP_CHECK_HIP(hipMemcpyAsync(xmax.data(), W->sub_scalar, ns * sizeof(double), hipMemcpyDeviceToHost, p->st));
    if ((rc = residual_norms())) return rc;
    for (int s = 0; s < ns; s++) {
        if (!running[s]) continue;
        auto &S = st[s];
        S.newton_iters++;
        S.linear_iters += its;
        S.linear_failures += (s < (int)W->lfail_h.size() && W->lfail_h[s]) ? 1 : 0;
        S.maxres = W->maxres_h[s];
        S.max_update = dmax[s] * std::fabs(o->relaxation);
        const bool small_update = o->correction_tol <= 0.0 || S.max_update <= o->correction_tol * xmax[s];
        if (S.maxres < o->tol && small_update) S.converged = 1;
        else if (S.maxres > o->max_residuals || S.newton_iters >= o->max_iter) S.failed = 1;
    }
    return PYHIP_OK;
}

// Arclength continuation (declared in pyoomph_hip.h): the reference's scalar recursion per
// sub-problem on the host, the vectors and solves batched on the device.
int pyhip_arclength_step(pyhip_problem *p, double *x, int param, const pyhip_newton_opts *o,
                         const int32_t *active, const double *ds, pyhip_arclength_state *st,
                         double *ds_next) {
    PYHIP_REQUIRE(p && x && o && active && ds && st && ds_next, "null argument");
    PYHIP_REQUIRE(param == 0, "unknown global parameter");
    PYHIP_REQUIRE(o->jacobian_precision == PYHIP_F64 && o->linear.precision == PYHIP_F64,
                  "arclength continuation runs FP64 solves");
    for (auto &G : p->groups)
        PYHIP_REQUIRE(G.form == PYHIP_FORM_BRATU || face_group_form(G.form) || gen_form_has_pderiv(G.form),
                      "arclength continuation needs sub-problems with a global parameter (Bratu lambda)");
    int rc;
    if ((rc = krylov_setup(p, o->linear.restart))) return rc;
    KrylovWork *W = p->kw;
    W->amg_stale = true;
    W->step_res_ok.assign(W->step_res_ok.size(), 0);
    const int ns = p->nsub;
    const int64_t n = p->ndof;
    if (!W->vals64 && (rc = dev_alloc(&W->vals64, p->nnz, W->bytes))) return rc;
    if (!W->ac_cur) {
        if ((rc = kalloc(&W->ac_cur, std::max<int64_t>(1, n), W)) || (rc = kalloc(&W->ac_der, std::max<int64_t>(1, n), W)) ||
            (rc = kalloc(&W->ac_y, std::max<int64_t>(1, n), W)) || (rc = kalloc(&W->ac_z, std::max<int64_t>(1, n), W)) ||
            (rc = kalloc(&W->ac_t, std::max<int64_t>(1, n), W)) || (rc = kalloc(&W->ac_zs, std::max<int64_t>(1, n), W)) ||
            (rc = kalloc(&W->ac_alpha, std::max(1, ns), W)) ||
            (rc = kalloc(&W->ac_mask, std::max(1, ns), W)))
            return rc;
        PYHIP_CHECK_HIP(hipMemsetAsync(W->ac_der, 0, std::max<int64_t>(1, n) * sizeof(double), p->st));
    }
    const double kMinDs = 1.0e-10, kDesiredProp = 0.5;
    const int kDesiredIters = 5;
    std::vector<int32_t> mask(ns, 0);
    std::vector<double> alpha(ns, 0.0), lam(ns), acr(ns), maxres(ns), dots(2 * ns);
    auto upload = [&](const std::vector<int32_t> &m, const std::vector<double> &a) -> int {
        PYHIP_CHECK_HIP(hipMemcpyAsync(W->ac_mask, m.data(), ns * sizeof(int32_t), hipMemcpyHostToDevice, p->st));
        PYHIP_CHECK_HIP(hipMemcpyAsync(W->ac_alpha, a.data(), ns * sizeof(double), hipMemcpyHostToDevice, p->st));
        return PYHIP_OK;
    };
    auto vec = [&](int op, const double *a, const double *b, double *out) -> int {
        hipLaunchKernelGGL(cont_vec_kernel, dim3(nblk(n)), dim3(256), 0, p->st, op, n, p->dof_sub, W->ac_mask,
                           W->ac_alpha, a, b, out);
        KCHECK();
        return PYHIP_OK;
    };
    auto set_lambda = [&]() -> int {
        for (int s = 0; s < ns; s++) p->params_h[(size_t)s * kParamStride + param] = lam[s];
        PYHIP_CHECK_HIP(hipMemcpyAsync(p->params, p->params_h.data(), p->params_h.size() * sizeof(double),
                                       hipMemcpyHostToDevice, p->st));
        return PYHIP_OK;
    };
    // per-sub-problem dot of two vectors, on the host
    auto sub_dot = [&](const double *a, const double *b, double *out) -> int {
        int r2;
        if ((r2 = batched_dots<double>(p, W, 1, a, n, b, W->dots))) return r2;
        PYHIP_CHECK_HIP(hipMemcpyAsync(out, W->dots, ns * sizeof(double), hipMemcpyDeviceToHost, p->st));
        PYHIP_CHECK_HIP(hipStreamSynchronize(p->st));
        return PYHIP_OK;
    };
    // max |R| per sub-problem and the arclength constraint residual (newton_solve_continuation)
    auto residuals = [&](const std::vector<int32_t> &m) -> int {
        int r2;
        if ((r2 = pyhip_assemble(p, x, 0, -1, PYHIP_F64, W->res64, nullptr))) return r2;
        if ((r2 = batched_maxabs(p, W->res64, W->maxres))) return r2;
        PYHIP_CHECK_HIP(hipMemcpyAsync(maxres.data(), W->maxres, ns * sizeof(double), hipMemcpyDeviceToHost, p->st));
        if ((r2 = upload(m, alpha)) || (r2 = vec(3, x, W->ac_cur, W->ac_t))) return r2;
        if ((r2 = sub_dot(W->ac_der, W->ac_t, dots.data()))) return r2;
        for (int s = 0; s < ns; s++) {
            if (!m[s]) continue;
            double c = dots[s] * st[s].theta_squared;
            c += st[s].parameter_derivative * (lam[s] - st[s].parameter_current) - st[s].ds_current;
            acr[s] = c;
            if (std::fabs(c) > maxres[s]) maxres[s] = std::fabs(c);
        }
        return PYHIP_OK;
    };

    // entry: Parameter_current, Dof_current, Ds_current (problem.cc:10835-10853)
    std::vector<int32_t> act(ns, 0);
    for (int s = 0; s < ns; s++) {
        act[s] = active[s] && p->dof_off_h[s + 1] > p->dof_off_h[s];
        lam[s] = p->params_h[(size_t)s * kParamStride + param];
        st[s].newton_iters = 0;
        st[s].rejected = 0;
        ds_next[s] = ds[s];
        if (!act[s]) continue;
        st[s].parameter_current = lam[s];
        st[s].ds_current = ds[s];
        alpha[s] = 0.0;
    }
    // Dof_current = x; a sub-problem without a step taken starts from Dof_derivative = 0
    {
        std::vector<int32_t> fresh(ns, 0);
        for (int s = 0; s < ns; s++) fresh[s] = act[s] && !st[s].step_taken;
        if ((rc = upload(act, alpha)) || (rc = vec(4, x, nullptr, W->ac_cur))) return rc;
        if ((rc = upload(fresh, alpha)) || (rc = vec(5, nullptr, nullptr, W->ac_der))) return rc;
    }
    std::vector<int32_t> pending = act, failed_sub(ns, 0), lin_failed(ns, 0);
    std::vector<int> count(ns, 0);
    // a linear solve that ends above its tolerance is an error for its sub-problem, not a
    // rejected step (problem.cc:10905-10915 throws on a linear-solver error)
    auto note_linear_failures = [&](std::vector<int32_t> &run) {
        for (int s = 0; s < ns; s++)
            if (run[s] && s < (int)W->lfail_h.size() && W->lfail_h[s]) lin_failed[s] = 1;
    };
    while (true) {
        int npend = 0;
        for (int s = 0; s < ns; s++) npend += pending[s];
        if (!npend) break;
        // predictor (problem.cc:10870-10899)
        for (int s = 0; s < ns; s++) {
            if (!pending[s]) continue;
            if (std::fabs(st[s].ds_current) < kMinDs) {
                set_error("DESIRED ARC-LENGTH STEP HAS FALLEN BELOW MINIMUM TOLERANCE");
                return PYHIP_ERR_NEWTON;
            }
            lam[s] = st[s].parameter_current + st[s].parameter_derivative * st[s].ds_current;
            st[s].ds_current = (lam[s] - st[s].parameter_current) / st[s].parameter_derivative;
            alpha[s] = st[s].ds_current;
        }
        if ((rc = set_lambda()) || (rc = upload(pending, alpha)) || (rc = vec(0, W->ac_cur, W->ac_der, x))) return rc;
        // newton_solve_continuation, batched over the pending sub-problems
        std::vector<int32_t> run = pending;
        for (int s = 0; s < ns; s++) count[s] = 0, failed_sub[s] = 0;
        if ((rc = residuals(run))) return rc;   // count == 1 check (always one step)
        while (true) {
            int nrun = 0;
            for (int s = 0; s < ns; s++) nrun += run[s];
            if (!nrun) break;
            for (int s = 0; s < ns; s++) count[s] += run[s];
            PYHIP_CHECK_HIP(hipMemcpyAsync(W->dof_mask, run.data(), ns * sizeof(int32_t), hipMemcpyHostToDevice, p->st));
            // J y = R and J z = dR/dparam with the same Jacobian
            if ((rc = pyhip_assemble(p, x, 1, -1, PYHIP_F64, W->w64, W->vals64))) return rc;
            int its = 0, refs = 0, lfail = 0;
            if ((rc = linear_solve(p, W, o->linear, PYHIP_F64, W->res64, W->ac_y, W->dof_mask, &its, &refs, &lfail)))
                return rc;
            note_linear_failures(run);
            if ((rc = pyhip_assemble(p, x, 0, param, PYHIP_F64, W->ac_t, nullptr))) return rc;
            if ((rc = linear_solve(p, W, o->linear, PYHIP_F64, W->ac_t, W->ac_z, W->dof_mask, &its, &refs, &lfail)))
                return rc;
            note_linear_failures(run);
            if ((rc = sub_dot(W->ac_der, W->ac_y, dots.data())) || (rc = sub_dot(W->ac_der, W->ac_z, dots.data() + ns)))
                return rc;
            for (int s = 0; s < ns; s++) {
                if (!run[s]) continue;
                const double uy = dots[s] * st[s].theta_squared, uz = dots[ns + s] * st[s].theta_squared;
                const double dparam = (acr[s] - uy) / (st[s].parameter_derivative - uz);
                lam[s] -= dparam;
                alpha[s] = dparam;
            }
            // x -= y - dparam z; z of this iteration kept per sub-problem (each solve
            // restarts every row of its output from 0)
            if ((rc = set_lambda()) || (rc = upload(run, alpha)) || (rc = vec(1, W->ac_y, W->ac_z, x)) ||
                (rc = vec(4, W->ac_z, nullptr, W->ac_zs)))
                return rc;
            if ((rc = residuals(run))) return rc;
            for (int s = 0; s < ns; s++) {
                if (!run[s]) continue;
                if (lin_failed[s]) run[s] = 0;
                else if (maxres[s] < o->tol) run[s] = 0;
                else if (maxres[s] > o->max_residuals || count[s] >= o->max_iter) run[s] = 0, failed_sub[s] = 1;
            }
        }
        // failed Newton solves retry with a shorter step (problem.cc:10918-10928)
        for (int s = 0; s < ns; s++) {
            if (!pending[s]) continue;
            if (lin_failed[s]) {
                pending[s] = 0;
            } else if (failed_sub[s]) {
                st[s].ds_current *= (2.0 / 3.0);
                st[s].rejected++;
            } else {
                pending[s] = 0;
                st[s].newton_iters = count[s];
            }
        }
    }
    // calculate_continuation_derivatives(z) with the last z (problem.cc:9995-10020, :10115-10200)
    std::vector<double> zz(ns), dz(ns);
    if ((rc = sub_dot(W->ac_zs, W->ac_zs, zz.data())) || (rc = sub_dot(W->ac_der, W->ac_zs, dz.data()))) return rc;
    // Continuation_direction = Parameter_derivative - Dof_derivative . z, then
    // Parameter_derivative = +-1/sqrt(1 + Theta^2 z.z) with its sign, Dof_derivative = -Pd z
    auto helper = [&](int s, double der_dot_z) {
        const double cdir = st[s].parameter_derivative - der_dot_z;
        double pd = 1.0 / std::sqrt(1.0 + st[s].theta_squared * zz[s]);
        if (pd * cdir < 0.0) pd = -pd;
        st[s].parameter_derivative = pd;
    };
    // a sub-problem whose linear solve failed keeps its state: dofs back to Dof_current,
    // parameter back to Parameter_current
    int nlin = 0;
    std::string lin_msg;
    for (int s = 0; s < ns; s++) {
        st[s].linear_failed = lin_failed[s];
        if (!lin_failed[s]) continue;
        act[s] = 0;
        lam[s] = st[s].parameter_current;
        ds_next[s] = ds[s];
        if (nlin++ < 8) lin_msg += (nlin > 1 ? ", " : " ") + std::to_string(s);
    }
    if (nlin) {
        std::vector<int32_t> back(ns, 0);
        for (int s = 0; s < ns; s++) back[s] = lin_failed[s];
        if ((rc = set_lambda()) || (rc = upload(back, alpha)) || (rc = vec(4, W->ac_cur, nullptr, x))) return rc;
    }
    for (int s = 0; s < ns; s++) st[s].turning_point = 0;
    for (int s = 0; s < ns; s++) {
        if (!act[s]) continue;
        const double pd_before = st[s].parameter_derivative;
        helper(s, dz[s]);
        const double pd = st[s].parameter_derivative;
        if (pd != 1.0) {
            st[s].theta_squared *= (pd * pd / kDesiredProp) * ((1.0 - kDesiredProp) / (1.0 - pd * pd));
            helper(s, -pd * zz[s]);   // Dof_derivative is -pd z after the first call
        }
        alpha[s] = st[s].parameter_derivative;
        if (!st[s].step_taken)
            st[s].ds_current = (lam[s] - st[s].parameter_current) / st[s].parameter_derivative;
        // a fold between parameter_current and the new parameter: the sign of
        // Parameter_derivative follows that of det J along the branch
        st[s].turning_point = st[s].step_taken && pd_before * st[s].parameter_derivative < 0.0;
        st[s].step_taken = 1;
        const int c = st[s].newton_iters;
        ds_next[s] = c < kDesiredIters ? st[s].ds_current * 1.5
                                       : (c > kDesiredIters ? st[s].ds_current * (2.0 / 3.0) : st[s].ds_current);
    }
    if ((rc = upload(act, alpha)) || (rc = vec(2, W->ac_zs, nullptr, W->ac_der))) return rc;
    for (int s = 0; s < ns; s++) st[s].parameter = lam[s];
    PYHIP_CHECK_HIP(hipStreamSynchronize(p->st));
    if (nlin) {
        set_error("ERROR IN THE LINEAR SOLVER during an arclength step of sub-problem(s)" + lin_msg +
                  (nlin > 8 ? ", ..." : "") + " (state kept; see linear_failed)");
        return PYHIP_ERR_NOT_CONVERGED;
    }
    return PYHIP_OK;
}

}  // extern "C"
