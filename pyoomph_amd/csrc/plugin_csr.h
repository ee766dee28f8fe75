// Host-side CSR preparation for the linear-solver plugin entry (pyhip_solve_serial).
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace pyhip {

// From the caller's compressed-column arrays (values, rowind, colptr):
//   transpose != 0: they are J^T in CSC, i.e. J in CSR (what oomph passes,
//                   linear_solver.cc:2253-2311);
//   transpose == 0: they are A in CSC.
// Builds the column-sorted CSR of the matrix to solve with (rs, cl), the caller's value
// index of every entry (src, -1 for added structural zeros), and the rows whose
// diagonal is absent or zero (last: ordered last by ILU(0)). Zero-diagonal rows get a
// structural zero diagonal and the two-hop pattern among themselves (p ~ q when both
// couple to a common unknown with a non-zero diagonal), the pattern the element
// assembly gives a saddle-point constraint block, so ILU(0) can approximate its Schur
// complement. Returns 0, or -1 with err set.
int plugin_build_csr(int n, int nnz, const double *values, const int32_t *rowind,
                     const int32_t *colptr, int transpose, std::vector<int32_t> &rs,
                     std::vector<int32_t> &cl, std::vector<int64_t> &src,
                     std::vector<char> &last, std::string &err);

}  // namespace pyhip
