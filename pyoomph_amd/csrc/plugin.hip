// Linear-solver plugin entry with the arguments of pyoomph's
// GenericLinearSystemSolver.solve_serial (pyoomph/solvers/generic.py:77-78) and of the
// extern "C" superlu() wrapper (src/pybind/solver.cpp:105-108) that oomph calls
// (linear_solver.cc:2297-2309). oomph passes its CSR Jacobian as compressed-column
// arrays of J^T with transpose = 1 (linear_solver.cc:2253-2311): colptr is J's row
// start, rowind J's column index.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <numeric>

#include "plugin_csr.h"
#include "problem.h"

namespace pyhip {
int plugin_setup(pyhip_problem *p, const pyhip_linear_opts &o);
double *plugin_values(pyhip_problem *p);
int plugin_factor(pyhip_problem *p, const pyhip_linear_opts &o);
int plugin_solve(pyhip_problem *p, const pyhip_linear_opts &o, const double *b, double *x, int *iters);
}  // namespace pyhip
extern "C" int pyhip_problem_destroy(pyhip_problem *p);

using namespace pyhip;

namespace {
struct SolveHandle {
    pyhip_problem *p = nullptr;
    pyhip_linear_opts o;
    std::vector<int64_t> src;     // sorted entry -> index in the caller's values
    int n = 0;
    double *b = nullptr, *x = nullptr;
};
}  // namespace

extern "C" int pyhip_solve_serial(int op_flag, int n, int nnz, int nrhs, const double *values,
                                  const int32_t *rowind, const int32_t *colptr, double *b,
                                  int ldb, int transpose, void **handle) {
    PYHIP_REQUIRE(handle, "null handle");
    SolveHandle *H = (SolveHandle *)*handle;
    if (op_flag == 3) {
        if (H) {
            dfree(H->b);
            dfree(H->x);
            pyhip_problem_destroy(H->p);
            delete H;
        }
        *handle = nullptr;
        return PYHIP_OK;
    }
    if (op_flag == 1) {
        PYHIP_REQUIRE(n > 0 && nnz >= 0 && values && rowind && colptr, "invalid matrix");
        if (H) {
            dfree(H->b);
            dfree(H->x);
            pyhip_problem_destroy(H->p);
            delete H;
            *handle = nullptr;
        }
        // CSR of the matrix to solve with; rows are column-sorted and a structural zero
        // diagonal is added where the caller's CSR has none (the reference drops exact
        // zeros, e.g. the pressure-pressure block of Taylor-Hood), so ILU(0) can fill it
        std::vector<int32_t> rs, cl;
        std::vector<int64_t> src;
        std::vector<char> last;
        std::string err;
        if (plugin_build_csr(n, nnz, values, rowind, colptr, transpose, rs, cl, src, last, err)) {
            set_error(err);
            return PYHIP_ERR_INVALID;
        }
        nnz = rs[n];
        H = new SolveHandle();
        H->n = n;
        H->src = src;
        pyhip_linear_opts_default(&H->o);
        if (const char *pc = getenv("PYHIP_SOLVE_SERIAL_PRECOND"))
            H->o.precond = (std::string(pc) == "jacobi") ? PYHIP_PRECOND_JACOBI : PYHIP_PRECOND_ILU0;
        pyhip_problem *P = new pyhip_problem();
        H->p = P;
        P->nsub = 1;
        P->ndof = n;
        P->nnz = nnz;
        P->dof_off_h = {0, n};
        P->row_start_h = rs;
        P->col_h = cl;
        std::vector<int32_t> diag(n, -1), dsub(n, 0);
        for (int i = 0; i < n; i++)
            for (int32_t k = rs[i]; k < rs[i + 1]; k++)
                if (cl[k] == i) diag[i] = k;
        P->ilu_last_h = last;
        int64_t &B = P->device_bytes;
        int rc;
        std::vector<int64_t> off = {0, n};
        if ((rc = dev_alloc(&P->row_start, n + 1, B)) || (rc = dev_alloc(&P->col, nnz, B)) ||
            (rc = dev_alloc(&P->diag, n, B)) || (rc = dev_alloc(&P->dof_sub, n, B)) ||
            (rc = dev_alloc(&P->dof_off, 2, B)) || (rc = dev_alloc(&H->b, n, B)) ||
            (rc = dev_alloc(&H->x, n, B))) {
            pyhip_problem_destroy(P);
            delete H;
            return rc;
        }
        PYHIP_CHECK_HIP(hipMemcpy(P->row_start, rs.data(), (n + 1) * 4, hipMemcpyHostToDevice));
        PYHIP_CHECK_HIP(hipMemcpy(P->col, cl.data(), (size_t)nnz * 4, hipMemcpyHostToDevice));
        PYHIP_CHECK_HIP(hipMemcpy(P->diag, diag.data(), (size_t)n * 4, hipMemcpyHostToDevice));
        PYHIP_CHECK_HIP(hipMemcpy(P->dof_sub, dsub.data(), (size_t)n * 4, hipMemcpyHostToDevice));
        PYHIP_CHECK_HIP(hipMemcpy(P->dof_off, off.data(), 2 * 8, hipMemcpyHostToDevice));
        if ((rc = plugin_setup(P, H->o))) return rc;
        std::vector<double> v(nnz);
        for (int k = 0; k < nnz; k++) v[k] = src[k] >= 0 ? values[src[k]] : 0.0;
        PYHIP_CHECK_HIP(hipMemcpy(plugin_values(P), v.data(), (size_t)nnz * 8, hipMemcpyHostToDevice));
        *handle = H;
        return plugin_factor(P, H->o);
    }
    if (op_flag == 2) {
        PYHIP_REQUIRE(H && n == H->n && b && nrhs >= 1 && ldb >= n, "solve before factorise");
        for (int r = 0; r < nrhs; r++) {
            double *br = b + (size_t)r * ldb;
            PYHIP_CHECK_HIP(hipMemcpy(H->b, br, (size_t)n * 8, hipMemcpyHostToDevice));
            int its = 0;
            int rc = plugin_solve(H->p, H->o, H->b, H->x, &its);
            if (rc) return rc;
            PYHIP_CHECK_HIP(hipMemcpy(br, H->x, (size_t)n * 8, hipMemcpyDeviceToHost));
        }
        return PYHIP_OK;
    }
    set_error("op_flag must be 1 (factorise), 2 (solve) or 3 (free)");
    return PYHIP_ERR_INVALID;
}
