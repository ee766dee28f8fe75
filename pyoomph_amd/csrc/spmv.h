// CSR sparse matrix-vector products in one canonical summation order (shared by the
// Krylov solvers and the multigrid preconditioner).
#pragma once
#include <algorithm>

#include "common.h"

namespace pyhip {

__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }

// Row sums of val[k] * x[col[k]] in one canonical order: stripe t (t = 0..7) of a row
// accumulates its entries k0+t, k0+t+8, ... with fused multiply-adds, and the stripes
// combine as ((S0+S4)+(S2+S6)) + ((S1+S5)+(S3+S7)). LPR lanes per row (1, 2, 4 or 8;
// lane l owns stripes l, l+LPR, ...) all give the same bits, so the lane count can be
// tuned per problem without changing any result.
// A thread serves LPR rows of a 256-row block (rows r0 + p*(256/LPR) + threadIdx.x/LPR,
// p = 0..LPR-1) and issues the loads of all of them together: 8 (row, stripe) streams in
// flight per thread, whatever LPR. All lanes of a row call this; out[p] is the sum of
// row p (rows at or past r1: 0). With a row filter (rsub, act non-null) a row whose
// sub-problem rsub[row] has act[...] == 0 is skipped: live[p] tells which rows to store.
template <typename TV, typename TX, typename TY, int LPR>
__device__ __forceinline__ void rows_dot(int64_t r0, int64_t r1, const int32_t *__restrict__ rs,
                                         const int32_t *__restrict__ col, const TV *__restrict__ val,
                                         const TX *__restrict__ x, const int32_t *__restrict__ rsub,
                                         const int32_t *__restrict__ act, TY (&out)[LPR],
                                         bool (&live)[LPR]) {
    constexpr int NS = 8 / LPR, RPP = 256 / LPR;
    const int lane = threadIdx.x % LPR;
    int32_t k0[LPR], k1[LPR];
    int32_t len = 0;
#pragma unroll
    for (int p = 0; p < LPR; p++) {
        const int64_t row = r0 + p * RPP + threadIdx.x / LPR;
        const bool ok = row < r1 && (act == nullptr || act[rsub[row]] != 0);
        live[p] = ok;
        k0[p] = ok ? rs[row] : 0;
        k1[p] = ok ? rs[row + 1] : 0;
        len = max(len, k1[p] - k0[p]);
    }
    TY acc[LPR][NS];
#pragma unroll
    for (int p = 0; p < LPR; p++)
#pragma unroll
        for (int j = 0; j < NS; j++) acc[p][j] = TY(0);
    for (int32_t base = lane; base < len; base += 8) {
        int32_t cc[LPR][NS];
        TY vv[LPR][NS], xx[LPR][NS];
#pragma unroll
        for (int p = 0; p < LPR; p++)
#pragma unroll
            for (int j = 0; j < NS; j++) {
                const int32_t k = k0[p] + base + j * LPR;
                const bool in = k < k1[p];
                cc[p][j] = in ? col[k] : -1;
                vv[p][j] = in ? TY(val[k]) : TY(0);
            }
#pragma unroll
        for (int p = 0; p < LPR; p++)
#pragma unroll
            for (int j = 0; j < NS; j++) xx[p][j] = cc[p][j] >= 0 ? TY(x[cc[p][j]]) : TY(0);
#pragma unroll
        for (int p = 0; p < LPR; p++)
#pragma unroll
            for (int j = 0; j < NS; j++)
                if (cc[p][j] >= 0) acc[p][j] = fmadd(vv[p][j], xx[p][j], acc[p][j]);
    }
#pragma unroll
    for (int p = 0; p < LPR; p++) {
        TY v;
        if constexpr (LPR == 8) v = acc[p][0];
        else if constexpr (LPR == 4) v = acc[p][0] + acc[p][1];
        else if constexpr (LPR == 2) v = (acc[p][0] + acc[p][2]) + (acc[p][1] + acc[p][3]);
        else v = ((acc[p][0] + acc[p][4]) + (acc[p][2] + acc[p][6])) + ((acc[p][1] + acc[p][5]) + (acc[p][3] + acc[p][7]));
        if constexpr (LPR >= 8) v = v + __shfl_xor(v, 4, LPR);
        if constexpr (LPR >= 4) v = v + __shfl_xor(v, 2, LPR);
        if constexpr (LPR >= 2) v = v + __shfl_xor(v, 1, LPR);
        out[p] = v;
    }
}

// What a launch writes for row i with row sum v = (A x)_i (every variant):
//   kEpiPlain   y_i = v, or b_i - v when b is given
//   kEpiAdd     y_i = b_i + v            (y may alias b)
//   kEpiJacobi  out_i = x_i + w_i (b_i - v)   (a weighted-Jacobi step; y is not written)
//   kEpiScale   y_i = v and out_i = w_i v     (a restriction and the next level's x = W b)
enum { kEpiPlain = 0, kEpiAdd = 1, kEpiJacobi = 2, kEpiScale = 3 };

template <int EPI, typename TX, typename TY>
__device__ __forceinline__ void spmv_store(int64_t row, TY v, const TX *x, TY *y, const TY *b, const TY *w,
                                           TY *out) {
    if constexpr (EPI == kEpiPlain) y[row] = b ? b[row] - v : v;
    else if constexpr (EPI == kEpiAdd) y[row] = b[row] + v;
    else if constexpr (EPI == kEpiJacobi) out[row] = TY(x[row]) + w[row] * (b[row] - v);
    else {
        y[row] = v;
        out[row] = w[row] * v;
    }
}

// 256 rows per block (rows_dot); the grid is a multiple of 8 (spmv_grid). Rows the filter
// skips keep their outputs.
template <typename TV, typename TX, typename TY, int LPR, int EPI>
__global__ __launch_bounds__(256) void spmv_kernel(int64_t n, const int32_t *__restrict__ rs,
                                                   const int32_t *__restrict__ col,
                                                   const TV *__restrict__ val,
                                                   const TX *__restrict__ x, TY *y, const TY *b,
                                                   const TY *__restrict__ w, TY *__restrict__ out,
                                                   const int32_t *__restrict__ rsub,
                                                   const int32_t *__restrict__ act) {
    constexpr int RPP = 256 / LPR;
    // blocks round-robin over the 8 XCDs: XCD x gets the contiguous block range x*per..
    const int64_t nb = (n + 255) / 256, per = (nb + 7) / 8, k = blockIdx.x / 8;
    const int64_t lb = (int64_t)(blockIdx.x % 8) * per + k;
    if (k >= per || lb >= nb) return;
    const int64_t r0 = lb * 256;
    TY v[LPR];
    bool live[LPR];
    rows_dot<TV, TX, TY, LPR>(r0, n, rs, col, val, x, rsub, act, v, live);
    if (threadIdx.x % LPR == 0) {
#pragma unroll
        for (int p = 0; p < LPR; p++) {
            const int64_t row = r0 + p * RPP + threadIdx.x / LPR;
            if (live[p]) spmv_store<EPI>(row, v[p], x, y, b, w, out);
        }
    }
}

// The same row sums through LDS, for short rows: a block of 256 rows streams its entries
// in chunks of CAP with coalesced loads (each lane loads entries t, t+256, ... of the chunk:
// the value and x[col], both kept in LDS), then one thread per row adds the chunk's entries
// of its row to its eight stripe sums (entry k0+8g+j to stripe j, in increasing k), so the
// result has rows_dot's bits. Rows longer than a chunk span several chunks; their stripe
// sums stay in registers. Outputs as spmv_kernel (EPI). Rows the filter skips keep their
// outputs; a block none of whose rows is live returns at once.
template <typename TV, typename TX, typename TY, int CAP, int EPI>
__global__ __launch_bounds__(256) void spmv_lds_kernel(int64_t n, const int32_t *__restrict__ rs,
                                                       const int32_t *__restrict__ col,
                                                       const TV *__restrict__ val, const TX *__restrict__ x,
                                                       TY *y, const TY *b, const TY *__restrict__ w,
                                                       TY *__restrict__ out, const int32_t *__restrict__ rsub,
                                                       const int32_t *__restrict__ act) {
    constexpr int U = (CAP + 255) / 256;
    __shared__ TY sv[CAP], sx[CAP];
    const int64_t nb = (n + 255) / 256, per = (nb + 7) / 8, kb = blockIdx.x / 8;
    const int64_t lb = (int64_t)(blockIdx.x % 8) * per + kb;
    if (kb >= per || lb >= nb) return;
    const int64_t r0 = lb * 256, row = r0 + threadIdx.x, rend = r0 + 256 < n ? r0 + 256 : n;
    const bool live = row < n && (act == nullptr || act[rsub[row]] != 0);
    if (!__syncthreads_or(live)) return;
    const int32_t k0 = live ? rs[row] : 0, k1 = live ? rs[row + 1] : 0;
    const int32_t kb0 = rs[r0], kb1 = rs[rend];
    TY acc[8];
#pragma unroll
    for (int j = 0; j < 8; j++) acc[j] = TY(0);
    // values and columns of the next chunk are loaded while the current one is summed
    int32_t cc[U];
    TY vv[U];
    auto load_chunk = [&](int32_t cs) {
        const int32_t ce = min(cs + CAP, kb1);
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int32_t k = cs + u * 256 + (int32_t)threadIdx.x;
            const bool in = k < ce;
            cc[u] = in ? col[k] : -1;
            vv[u] = in ? TY(val[k]) : TY(0);
        }
    };
    if (kb0 < kb1) load_chunk(kb0);
    for (int32_t cs = kb0; cs < kb1; cs += CAP) {
        const int32_t ce = min(cs + CAP, kb1);
#pragma unroll
        for (int u = 0; u < U; u++) {
            if (cc[u] >= 0) {
                const int32_t t = u * 256 + (int32_t)threadIdx.x;
                sv[t] = vv[u];
                sx[t] = TY(x[cc[u]]);
            }
        }
        __syncthreads();
        if (cs + CAP < kb1) load_chunk(cs + CAP);
        const int32_t lo = max(k0, cs), hi = min(k1, ce);
        if (lo < hi) {
            for (int32_t g = k0 + ((lo - k0) & ~7); g < hi; g += 8) {
#pragma unroll
                for (int j = 0; j < 8; j++) {
                    const int32_t k = g + j;
                    if (k >= lo && k < hi) acc[j] = fmadd(sv[k - cs], sx[k - cs], acc[j]);
                }
            }
        }
        __syncthreads();
    }
    if (live) {
        const TY v = ((acc[0] + acc[4]) + (acc[2] + acc[6])) + ((acc[1] + acc[5]) + (acc[3] + acc[7]));
        spmv_store<EPI>(row, v, x, y, b, w, out);
    }
}

// Variants of the canonical spmv (all give the same bits): 1, 2, 4, 8 lanes per row
// (spmv_kernel), or LDS staging with chunks of 512, 1024, 2048 or 4096 entries
// (spmv_lds_kernel).
enum { kSpmvLds512 = 16, kSpmvLds1K = 32, kSpmvLds2K = 64, kSpmvLds4K = 128 };
static constexpr int kSpmvVariants[] = {1, 2, 4, 8, kSpmvLds512, kSpmvLds1K, kSpmvLds2K, kSpmvLds4K};
static inline bool spmv_variant_ok(int v) {
    for (int w : kSpmvVariants)
        if (v == w) return true;
    return false;
}

static inline unsigned spmv_grid(int64_t n) {
    const int64_t nb = std::max<int64_t>(1, (n + 255) / 256);
    return (unsigned)(((nb + 7) / 8) * 8);
}

template <typename TV, typename TX, typename TY, int EPI>
static void launch_spmv_epi(int variant, hipStream_t st, int64_t n, const int32_t *rs, const int32_t *col,
                            const TV *vals, const TX *x, TY *y, const TY *b, const TY *w, TY *out,
                            const int32_t *rsub, const int32_t *act) {
    const dim3 g(spmv_grid(n)), t(256);
#define PYHIP_SPMV_LANES(L)                                                                                   \
    hipLaunchKernelGGL((spmv_kernel<TV, TX, TY, L, EPI>), g, t, 0, st, n, rs, col, vals, x, y, b, w, out, rsub, act)
#define PYHIP_SPMV_LDS(C)                                                                                     \
    hipLaunchKernelGGL((spmv_lds_kernel<TV, TX, TY, C, EPI>), g, t, 0, st, n, rs, col, vals, x, y, b, w, out, rsub, \
                       act)
    switch (variant) {
    case 1: PYHIP_SPMV_LANES(1); break;
    case 2: PYHIP_SPMV_LANES(2); break;
    case 4: PYHIP_SPMV_LANES(4); break;
    case kSpmvLds512: PYHIP_SPMV_LDS(512); break;
    case kSpmvLds1K: PYHIP_SPMV_LDS(1024); break;
    case kSpmvLds2K: PYHIP_SPMV_LDS(2048); break;
    case kSpmvLds4K: PYHIP_SPMV_LDS(4096); break;
    default: PYHIP_SPMV_LANES(8); break;
    }
#undef PYHIP_SPMV_LANES
#undef PYHIP_SPMV_LDS
}

// y = A x, or y = b - A x when b is given
template <typename TV, typename TX, typename TY>
static void launch_spmv(int variant, hipStream_t st, int64_t n, const int32_t *rs, const int32_t *col,
                        const TV *vals, const TX *x, TY *y, const TY *b, const int32_t *rsub = nullptr,
                        const int32_t *act = nullptr) {
    launch_spmv_epi<TV, TX, TY, kEpiPlain>(variant, st, n, rs, col, vals, x, y, b, nullptr, nullptr, rsub, act);
}

// y = b + A x (y may alias b)
template <typename TV, typename TX, typename TY>
static void launch_spmv_add(int variant, hipStream_t st, int64_t n, const int32_t *rs, const int32_t *col,
                            const TV *vals, const TX *x, TY *y, const TY *b, const int32_t *rsub = nullptr,
                            const int32_t *act = nullptr) {
    launch_spmv_epi<TV, TX, TY, kEpiAdd>(variant, st, n, rs, col, vals, x, y, b, nullptr, nullptr, rsub, act);
}

}  // namespace pyhip
