// Deterministic, atomics-free assembly of element vectors / matrices into the residual
// and the CSR values, and export of the reference-form CSR.
//
// Reference semantics (oomph Problem::sparse_assemble_row_or_column_compressed_with_
// vectors_of_pairs, problem.cc:5459-5659): elements are visited in order; for each local
// row i, residuals[eqn_i] += el_res[i]; each (i, j) with fabs(el_jac(i,j)) > 0.0 is
// added to the pair (eqn_i, eqn_j) of the row, created on first touch. So every CSR
// value is the element-order sum of its non-zero element contributions, and a row's
// columns appear in the order of their first non-zero contribution.
//
// Here every entry of the column-sorted pattern carries the list of element-matrix entries
// that add into it, in element order (pattern.h). One thread owns one entry and sums its
// list in that order, skipping exact zeros as the reference does -- the same summation
// order, with no atomics, so the result is bit-for-bit reproducible and independent of
// how entries are spread over waves.
#include "common.h"
#include "pattern.h"

namespace pyhip {

constexpr int kRowLanes = 32;
constexpr int kRowsPerBlock = 256 / kRowLanes;
constexpr int kGatherBlock = 256;

// CSR values from element matrices: the value blocks own one entry per thread; the
// blocks after them (when res is given) own one residual row per thread.
template <typename TE, typename TV>
__global__ __launch_bounds__(kGatherBlock) void gather_values_kernel(
    int64_t nnz, int64_t nrow, int64_t nb_vals, const int32_t *__restrict__ cptr,
    const uint32_t *__restrict__ cidx, const int64_t *__restrict__ occ_ptr,
    const int64_t *__restrict__ occ_roff, const TE *__restrict__ Ej, const TE *__restrict__ Er,
    TV *__restrict__ vals, double *__restrict__ res) {
    // value blocks: the first nb_pad = 8 * per blocks, dealt round-robin over the 8 XCDs;
    // XCD x takes the contiguous range x*per.. so neighbouring entries (which share
    // elements) read the element matrices through the same L2
    const int64_t per = (nb_vals + 7) / 8, nb_pad = 8 * per;
    if ((int64_t)blockIdx.x < nb_pad) {
        const int64_t lb = (int64_t)(blockIdx.x % 8) * per + blockIdx.x / 8;
        const int64_t i = lb * kGatherBlock + threadIdx.x;
        if (lb >= nb_vals || i >= nnz) return;
        TV acc = TV(0);
        for (int32_t c = cptr[i]; c < cptr[i + 1]; c++) {
            const TV v = TV(Ej[cidx[c]]);
            // oomph skips contributions with !(fabs(v) > 0): zeros and NaN
            if (v > TV(0) || v < TV(0)) acc = acc + v;
        }
        vals[i] = acc;
    } else {
        const int64_t r = ((int64_t)blockIdx.x - nb_pad) * kGatherBlock + threadIdx.x;
        if (r >= nrow) return;
        double acc = 0.0;
        for (int64_t o = occ_ptr[r]; o < occ_ptr[r + 1]; o++) acc = acc + double(Er[occ_roff[o]]);
        res[r] = acc;
    }
}

// Residual: residual[r] = sum over occurrences (element order) of the element vector.
template <typename TE>
__global__ __launch_bounds__(256) void gather_residual_kernel(
    int64_t nrow, const int64_t *__restrict__ occ_ptr, const int64_t *__restrict__ occ_roff,
    const TE *__restrict__ Er, double *__restrict__ res) {
    const int64_t r = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (r >= nrow) return;
    double acc = 0.0;
    for (int64_t o = occ_ptr[r]; o < occ_ptr[r + 1]; o++) acc = acc + double(Er[occ_roff[o]]);
    res[r] = acc;
}

// Export, pass 1: key of every structural entry = element-matrix index of its first
// non-zero contribution (grows with the reference's visiting order within the row);
// UINT32_MAX if it has none (dropped). kept[r] = entries of row r that stay.
__global__ __launch_bounds__(256) void export_keys_kernel(
    int64_t nrow, const int32_t *__restrict__ row_start, const int32_t *__restrict__ cptr,
    const uint32_t *__restrict__ cidx, const double *__restrict__ Ej,
    uint32_t *__restrict__ keys, int32_t *__restrict__ kept) {
    const int g = threadIdx.x / kRowLanes, lane = threadIdx.x % kRowLanes;
    const int64_t r = (int64_t)blockIdx.x * kRowsPerBlock + g;
    if (r >= nrow) return;
    const int32_t s = row_start[r], len = row_start[r + 1] - s;
    int n = 0;
    for (int p = lane; p < len; p += kRowLanes) {
        uint32_t key = UINT32_MAX;
        for (int32_t c = cptr[s + p]; c < cptr[s + p + 1]; c++) {
            const double v = Ej[cidx[c]];
            if (v > 0.0 || v < 0.0) {
                key = cidx[c];
                break;
            }
        }
        keys[s + p] = key;
        n += (key != UINT32_MAX);
    }
    // sum over the 32 lanes of the group
    for (int off = 16; off > 0; off >>= 1) n += __shfl_xor(n, off, kRowLanes);
    if (lane == 0) kept[r] = n;
}

// Export, pass 2: place every kept entry at its rank within the row: the first-touch rank
// (keys in element order), or with sorted the rank among the kept entries in column order
// (the std::map rows of Problem::sparse_assemble_row_or_column_compressed_base_problem,
// src/problem.cpp:2232-2262).
__global__ __launch_bounds__(256) void export_place_kernel(
    int64_t nrow, const int32_t *__restrict__ row_start, const int32_t *__restrict__ col,
    const uint32_t *__restrict__ keys, const double *__restrict__ vals,
    const int32_t *__restrict__ new_start, int32_t *__restrict__ out_col,
    double *__restrict__ out_val, int sorted) {
    const int g = threadIdx.x / kRowLanes, lane = threadIdx.x % kRowLanes;
    const int64_t r = (int64_t)blockIdx.x * kRowsPerBlock + g;
    if (r >= nrow) return;
    const int32_t s = row_start[r], len = row_start[r + 1] - s;
    for (int p = lane; p < len; p += kRowLanes) {
        const uint32_t k = keys[s + p];
        if (k == UINT32_MAX) continue;
        int rank = 0;
        if (sorted) {
            for (int q = 0; q < p; q++) rank += (keys[s + q] != UINT32_MAX);
        } else {
            for (int q = 0; q < len; q++) rank += (keys[s + q] < k);
        }
        out_col[new_start[r] + rank] = col[s + p];
        out_val[new_start[r] + rank] = vals[s + p];
    }
}

// Exclusive prefix sum of n int32 counts (row lengths of the export): per block of
// kScanItems counts the block total, one block scans the totals, then every block scans
// its counts from its offset. Three launches, no atomics, no library.
constexpr int kScanThreads = 256, kScanPer = 8, kScanItems = kScanThreads * kScanPer;

// exclusive scan of one value per thread over the block (Hillis-Steele in LDS)
__device__ inline int32_t block_exclusive_scan(int32_t v, int32_t *sh, int32_t *total) {
    const int t = threadIdx.x;
    sh[t] = v;
    __syncthreads();
    for (int off = 1; off < kScanThreads; off <<= 1) {
        const int32_t add = t >= off ? sh[t - off] : 0;
        __syncthreads();
        sh[t] += add;
        __syncthreads();
    }
    const int32_t incl = sh[t];
    *total = sh[kScanThreads - 1];
    __syncthreads();
    return incl - v;
}

__global__ __launch_bounds__(kScanThreads) void scan_block_totals_kernel(int64_t n, const int32_t *__restrict__ in,
                                                                         int32_t *__restrict__ totals) {
    __shared__ int32_t sh[kScanThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanItems + (int64_t)threadIdx.x * kScanPer;
    int32_t v = 0;
    for (int k = 0; k < kScanPer; k++)
        if (base + k < n) v += in[base + k];
    int32_t total;
    block_exclusive_scan(v, sh, &total);
    if (threadIdx.x == 0) totals[blockIdx.x] = total;
}

// one block: totals[0..nb) -> their exclusive prefix sums, in place
__global__ __launch_bounds__(kScanThreads) void scan_totals_kernel(int64_t nb, int32_t *__restrict__ totals) {
    __shared__ int32_t sh[kScanThreads];
    int32_t carry = 0;
    for (int64_t c0 = 0; c0 < nb; c0 += kScanThreads) {
        const int64_t i = c0 + threadIdx.x;
        const int32_t v = i < nb ? totals[i] : 0;
        int32_t total;
        const int32_t ex = block_exclusive_scan(v, sh, &total);
        if (i < nb) totals[i] = carry + ex;
        carry += total;
    }
}

__global__ __launch_bounds__(kScanThreads) void scan_apply_kernel(int64_t n, const int32_t *__restrict__ in,
                                                                  const int32_t *__restrict__ offsets,
                                                                  int32_t *__restrict__ out) {
    __shared__ int32_t sh[kScanThreads];
    const int64_t base = (int64_t)blockIdx.x * kScanItems + (int64_t)threadIdx.x * kScanPer;
    int32_t v[kScanPer];
    int32_t mine = 0;
    for (int k = 0; k < kScanPer; k++) {
        v[k] = base + k < n ? in[base + k] : 0;
        mine += v[k];
    }
    int32_t total;
    int32_t run = offsets[blockIdx.x] + block_exclusive_scan(mine, sh, &total);
    for (int k = 0; k < kScanPer; k++) {
        if (base + k < n) out[base + k] = run;
        run += v[k];
    }
}

int64_t export_kept_ints(int64_t nrow) {
    return nrow + 1 + (nrow + 1 + kScanItems - 1) / kScanItems;
}

static int exclusive_scan_i32(int64_t n, const int32_t *in, int32_t *totals, int32_t *out, hipStream_t st) {
    const int64_t nb = (n + kScanItems - 1) / kScanItems;
    if (nb == 0) return PYHIP_OK;
    hipLaunchKernelGGL(scan_block_totals_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, st, n, in, totals);
    PYHIP_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(scan_totals_kernel, dim3(1), dim3(kScanThreads), 0, st, nb, totals);
    PYHIP_CHECK_HIP(hipGetLastError());
    hipLaunchKernelGGL(scan_apply_kernel, dim3((unsigned)nb), dim3(kScanThreads), 0, st, n, in, totals, out);
    PYHIP_CHECK_HIP(hipGetLastError());
    return PYHIP_OK;
}

template <typename T>
static int launch_gather_values(const GatherMap &m, const T *Ej, const T *Er, T *vals, double *res,
                                hipStream_t st) {
    const int64_t nb_vals = (m.nnz + kGatherBlock - 1) / kGatherBlock;
    const int64_t nb_pad = 8 * ((nb_vals + 7) / 8);
    const int64_t nb_res = res ? (m.nrow + kGatherBlock - 1) / kGatherBlock : 0;
    if (nb_pad + nb_res == 0) return PYHIP_OK;
    hipLaunchKernelGGL((gather_values_kernel<T, T>), dim3((unsigned)(nb_pad + nb_res)),
                       dim3(kGatherBlock), 0, st, m.nnz, m.nrow, nb_vals, m.cptr, m.cidx,
                       m.occ_ptr, m.occ_roff, Ej, Er, vals, res);
    PYHIP_CHECK_HIP(hipGetLastError());
    return PYHIP_OK;
}

int launch_gather_values_f64(const GatherMap &m, const double *Ej, const double *Er, double *vals,
                             double *res, hipStream_t st) {
    return launch_gather_values<double>(m, Ej, Er, vals, res, st);
}

int launch_gather_values_f32(const GatherMap &m, const float *Ej, const float *Er, float *vals,
                             double *res, hipStream_t st) {
    return launch_gather_values<float>(m, Ej, Er, vals, res, st);
}

int launch_gather_residual_f64(int64_t nrow, const int64_t *occ_ptr, const int64_t *occ_roff,
                               const double *Er, double *res, hipStream_t st) {
    if (nrow == 0) return PYHIP_OK;
    const int blocks = (int)((nrow + 255) / 256);
    hipLaunchKernelGGL((gather_residual_kernel<double>), dim3(blocks), dim3(256), 0, st, nrow,
                       occ_ptr, occ_roff, Er, res);
    PYHIP_CHECK_HIP(hipGetLastError());
    return PYHIP_OK;
}

int launch_gather_residual_f32(int64_t nrow, const int64_t *occ_ptr, const int64_t *occ_roff,
                               const float *Er, double *res, hipStream_t st) {
    if (nrow == 0) return PYHIP_OK;
    const int blocks = (int)((nrow + 255) / 256);
    hipLaunchKernelGGL((gather_residual_kernel<float>), dim3(blocks), dim3(256), 0, st, nrow,
                       occ_ptr, occ_roff, Er, res);
    PYHIP_CHECK_HIP(hipGetLastError());
    return PYHIP_OK;
}

// Reference-form CSR export. Scratch: keys [nnz] uint32, kept [export_kept_ints(nrow)] int32.
int export_reference_csr(const GatherMap &m, const int32_t *col, const double *Ej,
                         const double *vals, uint32_t *keys, int32_t *kept, int32_t *out_rs,
                         int32_t *out_col, double *out_val, int64_t *nnz_out, hipStream_t st,
                         bool sorted) {
    const int64_t nrow = m.nrow;
    const int32_t *rs = m.rs;
    if (nrow == 0) {
        *nnz_out = 0;
        PYHIP_CHECK_HIP(hipMemsetAsync(out_rs, 0, sizeof(int32_t), st));
        return PYHIP_OK;
    }
    const int blocks = (int)((nrow + kRowsPerBlock - 1) / kRowsPerBlock);
    hipLaunchKernelGGL(export_keys_kernel, dim3(blocks), dim3(256), 0, st, nrow, rs, m.cptr,
                       m.cidx, Ej, keys, kept);
    PYHIP_CHECK_HIP(hipGetLastError());
    // exclusive scan of kept[0..nrow] -> out_rs[0..nrow] (kept[nrow] is a zero pad); the
    // block totals of the scan follow the pad
    PYHIP_CHECK_HIP(hipMemsetAsync(kept + nrow, 0, sizeof(int32_t), st));
    int rc = exclusive_scan_i32(nrow + 1, kept, kept + nrow + 1, out_rs, st);
    if (rc) return rc;
    hipLaunchKernelGGL(export_place_kernel, dim3(blocks), dim3(256), 0, st, nrow, rs, col, keys,
                       vals, out_rs, out_col, out_val, sorted ? 1 : 0);
    PYHIP_CHECK_HIP(hipGetLastError());
    int32_t total = 0;
    PYHIP_CHECK_HIP(hipMemcpyAsync(&total, out_rs + nrow, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    PYHIP_CHECK_HIP(hipStreamSynchronize(st));
    *nnz_out = total;
    return PYHIP_OK;
}

}  // namespace pyhip
