// Host-side sparsity pattern and element-to-nnz map of a (batched) problem.
//
// The reference discovers the pattern while assembling: oomph's vectors_of_pairs loop
// (problem.cc:5459-5560) walks elements in order and, for each local row i and column j,
// searches the global row's pair vector. Here the structural pattern is built once per
// mesh and numbering, and the per-assembly work becomes a gather:
//   * for every global row r, its "occurrences" (element e, local row i) in element
//     order, i.e. the element rows that add into r (problem.cc:5519, residuals[eqn] +=),
//   * for every entry of the column-sorted pattern, the list of element-matrix entries
//     that add into it, in element order ("contributions", CSR over the nnz entries).
// Summing an entry's contributions in list order reproduces oomph's element-order
// summation. Within a row the element-matrix index of a contribution grows with
// (occurrence, element column), the reference's visiting order, so an entry's first
// non-zero contribution index is its first-touch key for the reference-order export.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace pyhip {

constexpr int kMaxSlot = 27;     // element dofs per element (Q2 hex)
constexpr int kMaxRowLen = 255;  // entries per row

struct ElementBlock {
    int ns;                      // slots per element
    int64_t nel;
    const int32_t *seqn;         // [nel][ns] global equation, -1 pinned (element-major)
    int64_t er_off;              // offset of element 0 in the element-vector buffer
    int64_t ej_off;              // offset of element 0 in the element-matrix buffer
};

struct HostPattern {
    int64_t ndof = 0, nnz = 0, nocc = 0, ncontrib = 0;
    std::vector<int32_t> row_start;   // [ndof+1] column-sorted pattern
    std::vector<int32_t> col;         // [nnz]
    std::vector<int64_t> occ_ptr;     // [ndof+1]
    std::vector<int64_t> occ_roff;    // [nocc] element-vector entry of the occurrence
    std::vector<int32_t> cptr;        // [nnz+1] contributions of each entry
    std::vector<uint32_t> cidx;       // [ncontrib] element-matrix index, element order
    std::vector<int32_t> diag_pos;    // [ndof] index of the diagonal entry (or -1)
};

// Device copies of the map above, as the gather kernels take them.
struct GatherMap {
    int64_t nrow, nnz;
    const int32_t *rs;
    const int64_t *occ_ptr, *occ_roff;
    const int32_t *cptr;
    const uint32_t *cidx;
};

// Blocks must be given in element order (the reference's mesh element order; several
// sub-problems are several meshes one after another). Returns 0 or a negative status.
int build_pattern(int64_t ndof, const std::vector<ElementBlock> &blocks, HostPattern &out,
                  std::string &err);

}  // namespace pyhip
