// Row order, permuted pattern and level sets of the ILU(0) preconditioner (host, once per
// pattern; see ilu_plan.cpp). No reference counterpart: pyoomph solves with SuperLU
// (pyoomph/solvers/scipy.py:47-104); this preconditioner serves the batched GMRES/CG.
#pragma once
#include <cstdint>
#include <string>
#include <vector>

namespace pyhip {

struct IluHostPlan {
    std::vector<int32_t> perm;               // [n] new -> old
    std::vector<int32_t> prs, pcol, psrc;    // permuted column-sorted pattern, source entry
    std::vector<int32_t> pdiag;              // [n] diagonal position
    std::vector<int32_t> lo_rows, up_rows;   // rows grouped by lower / upper solve level
    std::vector<int64_t> lo_ptr, up_ptr;     // level boundaries
    bool multicolor = false;                 // some sub-problem takes the multicolour order
};

// order: 0 auto (multicolour for each sub-problem whose natural order has more than
// auto_levels lower-solve levels), 1 natural, 2 multicolour. is_p[i] = 1 puts row i in
// the second group of its sub-problem (Taylor-Hood pressures). Returns 0 or -1 (err).
int ilu_plan_host(int64_t n, const int32_t *rs, const int32_t *col, const std::vector<int64_t> &sub_off,
                  const std::vector<char> &is_p, int order, int auto_levels, IluHostPlan &out, std::string &err);

}  // namespace pyhip
