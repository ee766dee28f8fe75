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
// Here the row's element rows ("occurrences", element order) are precomputed; a
// 32-lane group owns one row, keeps the row's partial sums in LDS, and adds the element
// rows in occurrence order (lane j adds slot j of the element row to its column
// position) -- the same summation order, with no atomics, so the result is bit-for-bit
// reproducible and independent of how rows are spread over waves.
#include "common.h"
#include "pattern.h"

#include <hipcub/hipcub.hpp>

namespace pyhip {

constexpr int kRowLanes = 32;
constexpr int kRowsPerBlock = 256 / kRowLanes;
constexpr int kMaxRow = 256;

// CSR values (column-sorted structural pattern) from element matrices.
template <typename TE, typename TV>
__global__ __launch_bounds__(256) void gather_values_kernel(
    int64_t nrow, const int32_t *__restrict__ row_start, const int64_t *__restrict__ occ_ptr,
    const int64_t *__restrict__ occ_eoff, const uint8_t *__restrict__ occ_pos,
    const TE *__restrict__ Ej, TV *__restrict__ vals) {
    __shared__ TV acc_s[kRowsPerBlock][kMaxRow];
    const int g = threadIdx.x / kRowLanes, lane = threadIdx.x % kRowLanes;
    const int64_t r = (int64_t)blockIdx.x * kRowsPerBlock + g;
    if (r >= nrow) return;  // whole 32-lane group leaves together; no block barrier below
    volatile TV *acc = acc_s[g];
    const int32_t s = row_start[r], len = row_start[r + 1] - s;
    for (int p = lane; p < len; p += kRowLanes) acc[p] = TV(0);
    const int64_t o0 = occ_ptr[r], o1 = occ_ptr[r + 1];
    for (int64_t o = o0; o < o1; o++) {
        if (lane < kMaxSlot) {
            const uint8_t pos = occ_pos[o * kMaxSlot + lane];
            if (pos != kNoPos) {
                const TV v = TV(Ej[occ_eoff[o] + lane]);
                // oomph skips contributions with !(fabs(v) > 0): zeros and NaN
                if (v > TV(0) || v < TV(0)) acc[pos] = acc[pos] + v;
            }
        }
    }
    for (int p = lane; p < len; p += kRowLanes) vals[s + p] = acc[p];
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

// Export, pass 1: key of every structural entry = (occurrence index within the row, slot)
// of its first non-zero contribution; INT32_MAX if it has none (dropped).
__global__ __launch_bounds__(256) void export_keys_kernel(
    int64_t nrow, const int32_t *__restrict__ row_start, const int64_t *__restrict__ occ_ptr,
    const int64_t *__restrict__ occ_eoff, const uint8_t *__restrict__ occ_pos,
    const double *__restrict__ Ej, int32_t *__restrict__ keys, int32_t *__restrict__ kept) {
    __shared__ int32_t key_s[kRowsPerBlock][kMaxRow];
    const int g = threadIdx.x / kRowLanes, lane = threadIdx.x % kRowLanes;
    const int64_t r = (int64_t)blockIdx.x * kRowsPerBlock + g;
    if (r >= nrow) return;
    volatile int32_t *key = key_s[g];
    const int32_t s = row_start[r], len = row_start[r + 1] - s;
    for (int p = lane; p < len; p += kRowLanes) key[p] = INT32_MAX;
    const int64_t o0 = occ_ptr[r], o1 = occ_ptr[r + 1];
    for (int64_t o = o0; o < o1; o++) {
        if (lane < kMaxSlot) {
            const uint8_t pos = occ_pos[o * kMaxSlot + lane];
            if (pos != kNoPos) {
                const double v = Ej[occ_eoff[o] + lane];
                if ((v > 0.0 || v < 0.0) && key[pos] == INT32_MAX)
                    key[pos] = (int32_t)((o - o0) * 32 + lane);
            }
        }
    }
    int n = 0;
    for (int p = lane; p < len; p += kRowLanes) {
        const int32_t k = key[p];
        keys[s + p] = k;
        n += (k != INT32_MAX);
    }
    // sum over the 32 lanes of the group
    for (int off = 16; off > 0; off >>= 1) n += __shfl_xor(n, off, kRowLanes);
    if (lane == 0) kept[r] = n;
}

// Export, pass 2: place every kept entry at its first-touch rank within the row.
__global__ __launch_bounds__(256) void export_place_kernel(
    int64_t nrow, const int32_t *__restrict__ row_start, const int32_t *__restrict__ col,
    const int32_t *__restrict__ keys, const double *__restrict__ vals,
    const int32_t *__restrict__ new_start, int32_t *__restrict__ out_col,
    double *__restrict__ out_val) {
    const int g = threadIdx.x / kRowLanes, lane = threadIdx.x % kRowLanes;
    const int64_t r = (int64_t)blockIdx.x * kRowsPerBlock + g;
    if (r >= nrow) return;
    const int32_t s = row_start[r], len = row_start[r + 1] - s;
    for (int p = lane; p < len; p += kRowLanes) {
        const int32_t k = keys[s + p];
        if (k == INT32_MAX) continue;
        int rank = 0;
        for (int q = 0; q < len; q++) rank += (keys[s + q] < k);
        out_col[new_start[r] + rank] = col[s + p];
        out_val[new_start[r] + rank] = vals[s + p];
    }
}

int launch_gather_values_f64(int64_t nrow, const int32_t *rs, const int64_t *occ_ptr,
                             const int64_t *occ_eoff, const uint8_t *occ_pos, const double *Ej,
                             double *vals, hipStream_t st) {
    if (nrow == 0) return PYHIP_OK;
    const int blocks = (int)((nrow + kRowsPerBlock - 1) / kRowsPerBlock);
    hipLaunchKernelGGL((gather_values_kernel<double, double>), dim3(blocks), dim3(256), 0, st,
                       nrow, rs, occ_ptr, occ_eoff, occ_pos, Ej, vals);
    PYHIP_CHECK_HIP(hipGetLastError());
    return PYHIP_OK;
}

int launch_gather_values_f32(int64_t nrow, const int32_t *rs, const int64_t *occ_ptr,
                             const int64_t *occ_eoff, const uint8_t *occ_pos, const float *Ej,
                             float *vals, hipStream_t st) {
    if (nrow == 0) return PYHIP_OK;
    const int blocks = (int)((nrow + kRowsPerBlock - 1) / kRowsPerBlock);
    hipLaunchKernelGGL((gather_values_kernel<float, float>), dim3(blocks), dim3(256), 0, st,
                       nrow, rs, occ_ptr, occ_eoff, occ_pos, Ej, vals);
    PYHIP_CHECK_HIP(hipGetLastError());
    return PYHIP_OK;
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

// Reference-form CSR export. Scratch: keys [nnz] int32, kept [nrow+1] int32.
int export_reference_csr(int64_t nrow, const int32_t *rs, const int32_t *col,
                         const int64_t *occ_ptr, const int64_t *occ_eoff,
                         const uint8_t *occ_pos, const double *Ej, const double *vals,
                         int32_t *keys, int32_t *kept, int32_t *out_rs, int32_t *out_col,
                         double *out_val, int64_t *nnz_out, hipStream_t st) {
    if (nrow == 0) {
        *nnz_out = 0;
        PYHIP_CHECK_HIP(hipMemsetAsync(out_rs, 0, sizeof(int32_t), st));
        return PYHIP_OK;
    }
    const int blocks = (int)((nrow + kRowsPerBlock - 1) / kRowsPerBlock);
    hipLaunchKernelGGL(export_keys_kernel, dim3(blocks), dim3(256), 0, st, nrow, rs, occ_ptr,
                       occ_eoff, occ_pos, Ej, keys, kept);
    PYHIP_CHECK_HIP(hipGetLastError());
    // exclusive scan of kept[0..nrow] -> out_rs[0..nrow] (kept[nrow] is a zero pad)
    PYHIP_CHECK_HIP(hipMemsetAsync(kept + nrow, 0, sizeof(int32_t), st));
    size_t tmp_bytes = 0;
    PYHIP_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, kept, out_rs,
                                                     (int)(nrow + 1), st));
    void *tmp = nullptr;
    PYHIP_CHECK_HIP(hipMallocAsync(&tmp, tmp_bytes, st));
    PYHIP_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, kept, out_rs,
                                                     (int)(nrow + 1), st));
    PYHIP_CHECK_HIP(hipFreeAsync(tmp, st));
    hipLaunchKernelGGL(export_place_kernel, dim3(blocks), dim3(256), 0, st, nrow, rs, col, keys,
                       vals, out_rs, out_col, out_val);
    PYHIP_CHECK_HIP(hipGetLastError());
    int32_t total = 0;
    PYHIP_CHECK_HIP(hipMemcpyAsync(&total, out_rs + nrow, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    PYHIP_CHECK_HIP(hipStreamSynchronize(st));
    *nnz_out = total;
    return PYHIP_OK;
}

}  // namespace pyhip
