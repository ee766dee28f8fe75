// Multi-assembly: several quantities at the same dofs in one call -- the residual, the
// Jacobian, the mass matrix, their derivatives by the global parameter and Hessian-vector
// products (plain and transposed) -- numbered and returned as the reference's
// Problem::assemble_multiassembly returns them (src/problem.cpp:2326-2368) for the request
// list of CustomMultiAssembleHandler (src/bifurcation.cpp:5354-5480).
//
// Each quantity runs the batched element kernels and the deterministic gather on the
// problem's stream (one launch pair per group of requests that one element pass serves:
// residual + Jacobian, dR/dp + dJ/dp, one Hessian direction), and every matrix leaves in the
// reference's multi-assembly CSR: exact-zero element contributions dropped
// (src/problem.cpp:2195, Numerical_zero_for_sparse_assembly = 0.0) and the columns of a
// row in ascending order (one std::map per row, src/problem.cpp:2232-2262).
#include "common.h"
#include "problem.h"

#include <algorithm>
#include <string>
#include <vector>

namespace pyhip {

static bool ma_vector_kind(int w) { return w == PYHIP_MA_RESIDUALS || w == PYHIP_MA_DRESIDUALS_DPARAMETER; }
static bool ma_param_kind(int w) {
    return w == PYHIP_MA_DRESIDUALS_DPARAMETER || w == PYHIP_MA_DJACOBIAN_DPARAMETER ||
           w == PYHIP_MA_DMASS_MATRIX_DPARAMETER;
}
static bool ma_hessian_kind(int w) {
    return w == PYHIP_MA_HESSIAN_VECTOR_PRODUCT || w == PYHIP_MA_MASS_MATRIX_HESSIAN_VECTOR_PRODUCT;
}

static const char *ma_name(int w) {
    switch (w) {
    case PYHIP_MA_RESIDUALS: return "residuals";
    case PYHIP_MA_JACOBIAN: return "jacobian";
    case PYHIP_MA_MASS_MATRIX: return "mass_matrix";
    case PYHIP_MA_DRESIDUALS_DPARAMETER: return "dresiduals_dparameter";
    case PYHIP_MA_DJACOBIAN_DPARAMETER: return "djacobian_dparameter";
    case PYHIP_MA_DMASS_MATRIX_DPARAMETER: return "dmass_matrix_dparameter";
    case PYHIP_MA_HESSIAN_VECTOR_PRODUCT: return "hessian_vector_product";
    case PYHIP_MA_MASS_MATRIX_HESSIAN_VECTOR_PRODUCT: return "mass_matrix_hessian_vector_product";
    }
    return "?";
}

static std::string ma_contrib(const pyhip_multi_request &r) { return r.contribution ? r.contribution : ""; }

// the distinct contributions of the requests, in order of first appearance
static std::vector<std::string> ma_contributions(int nreq, const pyhip_multi_request *req) {
    std::vector<std::string> u;
    for (int i = 0; i < nreq; i++) {
        const std::string c = ma_contrib(req[i]);
        if (std::find(u.begin(), u.end(), c) == u.end()) u.push_back(c);
    }
    return u;
}

// Numbering of CustomMultiAssembleHandler (src/bifurcation.cpp:5393-5480): per contribution
// (residual table) in order of first appearance, its requests in request order; a vector
// request returns its index k >= 0, a matrix request -1 - m. The same quantity twice for
// one contribution is an error, and so is a mix of transposed and plain Hessian products,
// as in the reference.
static int ma_plan(int nreq, const pyhip_multi_request *req, int nhvec, int *ret, int *nvec, int *nmat) {
    PYHIP_REQUIRE(nreq >= 0 && (nreq == 0 || (req && ret)) && nvec && nmat, "null argument");
    bool plain = false, transposed = false;
    for (int i = 0; i < nreq; i++) {
        const pyhip_multi_request &r = req[i];
        PYHIP_REQUIRE(r.what >= PYHIP_MA_RESIDUALS && r.what <= PYHIP_MA_MASS_MATRIX_HESSIAN_VECTOR_PRODUCT,
                      "unknown multi-assembly request kind " + std::to_string(r.what));
        if (ma_param_kind(r.what))
            PYHIP_REQUIRE(r.param >= 0, std::string("no parameter given for '") + ma_name(r.what) + "'");
        if (ma_hessian_kind(r.what)) {
            PYHIP_REQUIRE(r.hvec >= 0, std::string("no Hessian vector given for '") + ma_name(r.what) + "'");
            PYHIP_REQUIRE(r.hvec < nhvec, "Hessian vector index out of bounds");
            if (r.transposed) transposed = true;
            else plain = true;
            PYHIP_REQUIRE(!(plain && transposed),
                          "cannot assemble transposed and non-transposed Hessian vector products simultaneously");
        }
    }
    int nv = 0, nm = 0;
    for (const std::string &c : ma_contributions(nreq, req)) {
        for (int i = 0; i < nreq; i++) {
            const pyhip_multi_request &r = req[i];
            if (ma_contrib(r) != c) continue;
            for (int j = 0; j < i; j++) {
                const pyhip_multi_request &q = req[j];
                bool same = q.what == r.what && ma_contrib(q) == c;
                if (same && ma_param_kind(r.what)) same = q.param == r.param;
                if (same && ma_hessian_kind(r.what)) same = q.hvec == r.hvec && (q.transposed != 0) == (r.transposed != 0);
                PYHIP_REQUIRE(!same, std::string("multiple '") + ma_name(r.what) + "' requests for the same contribution '" +
                                         c + "'");
            }
            ret[i] = ma_vector_kind(r.what) ? nv++ : -1 - nm++;
        }
    }
    *nvec = nv;
    *nmat = nm;
    return PYHIP_OK;
}

static int ma_scratch(pyhip_problem *p) {
    int rc;
    if (!p->ma_vals && (rc = dev_alloc(&p->ma_vals, std::max<int64_t>(p->nnz, 1), p->device_bytes))) return rc;
    if (!p->ma_res && (rc = dev_alloc(&p->ma_res, std::max<int64_t>(p->ndof, 1), p->device_bytes))) return rc;
    if (!p->keys) {
        if ((rc = dev_alloc(&p->keys, std::max<int64_t>(p->nnz, 1), p->device_bytes))) return rc;
        if ((rc = dev_alloc(&p->kept, export_kept_ints(p->ndof), p->device_bytes))) return rc;
    }
    return PYHIP_OK;
}

}  // namespace pyhip

using namespace pyhip;

extern "C" {

int pyhip_multi_assembly_plan(int nreq, const pyhip_multi_request *req, int nhvec, int *return_indices,
                              int *nvector, int *nmatrix) {
    return ma_plan(nreq, req, nhvec, return_indices, nvector, nmatrix);
}

int pyhip_assemble_multi(pyhip_problem *p, const double *dofs, int nreq, const pyhip_multi_request *req,
                         const double *hvecs, int nhvec, double *const *vectors, int32_t *const *row_start,
                         int32_t *const *col, double *const *val, int64_t *nnz_out) {
    PYHIP_REQUIRE(p && dofs, "null argument");
    std::vector<int> ret(std::max(nreq, 1));
    int nv = 0, nm = 0, rc;
    if ((rc = ma_plan(nreq, req, nhvec, ret.data(), &nv, &nm))) return rc;
    PYHIP_REQUIRE(nv == 0 || vectors, "vector outputs required");
    PYHIP_REQUIRE(nm == 0 || (row_start && col && val && nnz_out), "matrix outputs required");
    PYHIP_REQUIRE(nhvec == 0 || hvecs, "Hessian vectors required");
    for (int k = 0; k < nv; k++) PYHIP_REQUIRE(vectors[k], "null vector output");
    for (int m = 0; m < nm; m++) PYHIP_REQUIRE(row_start[m] && col[m] && val[m], "null matrix output");
    bool no_hv = false, generated_th = false, all_generated_th = true;
    for (const auto &G : p->groups) {
        if (G.form >= PYHIP_GENERATED_FORM_BASE) {
            no_hv = no_hv || !gen_form_has_hv(G.form);
            generated_th = generated_th || th_layout(G.form);
        }
        if (!face_group_form(G.form)) all_generated_th = all_generated_th && G.form >= PYHIP_GENERATED_FORM_BASE &&
                                                         th_layout(G.form);
    }
    for (int i = 0; i < nreq; i++) {
        if (ma_hessian_kind(req[i].what))
            PYHIP_REQUIRE(!no_hv, "Hessian-vector products need every generated form's _f64_hv / _f64_hvt kernels");
        // the mass coefficient of a generated Taylor-Hood form may depend on the parameter
        // and on the dofs: dM/dp comes from its kernels (every bulk group generated
        // Taylor-Hood), the mass Hessian is not generated
        if (req[i].what == PYHIP_MA_DMASS_MATRIX_DPARAMETER)
            PYHIP_REQUIRE(!generated_th || all_generated_th,
                          "d(mass matrix)/d(parameter) of run-time Taylor-Hood forms mixed with other forms");
        if (req[i].what == PYHIP_MA_MASS_MATRIX_HESSIAN_VECTOR_PRODUCT)
            PYHIP_REQUIRE(!generated_th, "mass-matrix Hessian products are not generated for run-time Taylor-Hood forms");
    }
    if ((rc = ma_scratch(p))) return rc;

    std::string cur;   // the contribution being assembled
    auto mine = [&](int i) { return ma_contrib(req[i]) == cur; };
    auto find = [&](int what, int param, int hvec, int tr) -> int {
        for (int i = 0; i < nreq; i++) {
            const pyhip_multi_request &r = req[i];
            if (r.what != what || !mine(i)) continue;
            if (ma_param_kind(what) && r.param != param) continue;
            if (ma_hessian_kind(what) && (r.hvec != hvec || (r.transposed != 0) != (tr != 0))) continue;
            return i;
        }
        return -1;
    };
    auto vec_out = [&](int i) -> double * { return i < 0 ? p->ma_res : vectors[ret[i]]; };
    auto export_to = [&](int i) -> int {
        const int m = -1 - ret[i];
        return export_reference_csr(p->gmap(), p->col, p->Ej64, p->ma_vals, p->keys, p->kept, row_start[m], col[m],
                                    val[m], &nnz_out[m], p->st, true);
    };
    // a matrix without any non-zero element contribution: every row empty
    auto zero_to = [&](int i) -> int {
        const int m = -1 - ret[i];
        PYHIP_CHECK_HIP(hipMemsetAsync(row_start[m], 0, (size_t)(p->ndof + 1) * sizeof(int32_t), p->st));
        nnz_out[m] = 0;
        return PYHIP_OK;
    };

    // one contribution: its residual table is the solved residual meanwhile
    auto run = [&]() -> int {
        int rc;
        // residual and Jacobian: one element pass
        const int iR = find(PYHIP_MA_RESIDUALS, 0, 0, 0), iJ = find(PYHIP_MA_JACOBIAN, 0, 0, 0);
        if (iR >= 0 || iJ >= 0) {
            if ((rc = assemble_subset(p, dofs, iJ >= 0 ? 1 : 0, -1, PYHIP_F64, vec_out(iR), p->ma_vals, nullptr)))
                return rc;
            if (iJ >= 0 && (rc = export_to(iJ))) return rc;
        }
        // mass matrix dR/d(du/dt): the flag-2 pass with the time derivative on (BDF1 weights
        // over the current dofs when the problem is steady; the matrix does not depend on them)
        const int iM = find(PYHIP_MA_MASS_MATRIX, 0, 0, 0);
        if (iM >= 0) {
            const bool was = p->transient;
            const double tw[3] = {p->tw[0], p->tw[1], p->tw[2]};
            const double *h1 = p->hist1, *h2 = p->hist2;
            if (!was) {
                p->transient = true;
                p->tw[0] = 1.0, p->tw[1] = -1.0, p->tw[2] = 0.0;
                p->hist1 = p->hist2 = dofs;
            }
            rc = assemble_subset(p, dofs, 2, -1, PYHIP_F64, p->ma_res, p->ma_vals, nullptr);
            p->transient = was;
            p->tw[0] = tw[0], p->tw[1] = tw[1], p->tw[2] = tw[2];
            p->hist1 = h1, p->hist2 = h2;
            if (rc || (rc = export_to(iM))) return rc;
        }
        // parameter derivatives, one element pass per parameter
        for (int i = 0; i < nreq; i++) {
            if (!ma_param_kind(req[i].what) || !mine(i)) continue;
            const int prm = req[i].param;
            bool first = true;
            for (int j = 0; j < i; j++) first = first && !(mine(j) && ma_param_kind(req[j].what) && req[j].param == prm);
            if (!first) continue;
            const int iDR = find(PYHIP_MA_DRESIDUALS_DPARAMETER, prm, 0, 0);
            const int iDJ = find(PYHIP_MA_DJACOBIAN_DPARAMETER, prm, 0, 0);
            const int iDM = find(PYHIP_MA_DMASS_MATRIX_DPARAMETER, prm, 0, 0);
            // the same argument checks as a plain assembly (the parameter must exist)
            if ((rc = assemble_subset(p, dofs, iDJ >= 0 ? 1 : 0, prm, PYHIP_F64, vec_out(iDR), p->ma_vals, nullptr)))
                return rc;
            if (iDJ >= 0 && (rc = export_to(iDJ))) return rc;
            // the mass coefficients of the compiled and generated scalar forms do not depend on
            // a parameter: dM/dp has no non-zero element contribution; a generated
            // Taylor-Hood form's is the flag-2 pass in derivative mode (its kernels put dm/dp
            // in the mass coefficient's place)
            if (iDM >= 0 && generated_th) {
                const bool was = p->transient;
                const double tw[3] = {p->tw[0], p->tw[1], p->tw[2]};
                const double *h1 = p->hist1, *h2 = p->hist2;
                if (!was) {
                    p->transient = true;
                    p->tw[0] = 1.0, p->tw[1] = -1.0, p->tw[2] = 0.0;
                    p->hist1 = p->hist2 = dofs;
                }
                rc = assemble_subset(p, dofs, 2, prm, PYHIP_F64, p->ma_res, p->ma_vals, nullptr);
                p->transient = was;
                p->tw[0] = tw[0], p->tw[1] = tw[1], p->tw[2] = tw[2];
                p->hist1 = h1, p->hist2 = h2;
                if (rc || (rc = export_to(iDM))) return rc;
            } else if (iDM >= 0 && (rc = zero_to(iDM))) {
                return rc;
            }
        }
        // Hessian-vector products, one element pass per direction
        for (int i = 0; i < nreq; i++) {
            if (!ma_hessian_kind(req[i].what) || !mine(i)) continue;
            const int hv = req[i].hvec, tr = req[i].transposed != 0;
            bool first = true;
            for (int j = 0; j < i; j++)
                first = first && !(mine(j) && ma_hessian_kind(req[j].what) && req[j].hvec == hv &&
                                   (req[j].transposed != 0) == tr);
            if (!first) continue;
            const int iH = find(PYHIP_MA_HESSIAN_VECTOR_PRODUCT, 0, hv, tr);
            const int iMH = find(PYHIP_MA_MASS_MATRIX_HESSIAN_VECTOR_PRODUCT, 0, hv, tr);
            if (iH >= 0) {
                p->hvec = hvecs + (int64_t)hv * p->ndof;
                p->htr = tr;
                rc = assemble_subset(p, dofs, 1, -1, PYHIP_F64, p->ma_res, p->ma_vals, nullptr);
                p->hvec = nullptr;
                p->htr = 0;
                if (rc || (rc = export_to(iH))) return rc;
            }
            // the mass matrices of these forms (compiled ones and generated scalar forms, whose
            // mass coefficient is 1) do not depend on the dofs
            if (iMH >= 0 && (rc = zero_to(iMH))) return rc;
        }
        return PYHIP_OK;
    };
    const std::string saved = p->solved_residual;
    for (const std::string &c : ma_contributions(nreq, req)) {
        cur = c;
        int ng = 0;
        if ((rc = pyhip_problem_set_solved_residual(p, c.c_str(), &ng))) return rc;
        if (!c.empty() && ng == 0) {
            pyhip_problem_set_solved_residual(p, saved.c_str(), nullptr);
            set_error("no equations define the residual contribution '" + c + "'");
            return PYHIP_ERR_INVALID;
        }
        rc = run();
        if (rc) {
            pyhip_problem_set_solved_residual(p, saved.c_str(), nullptr);
            return rc;
        }
    }
    if ((rc = pyhip_problem_set_solved_residual(p, saved.c_str(), nullptr))) return rc;
    PYHIP_CHECK_HIP(hipStreamSynchronize(p->st));
    return PYHIP_OK;
}

}  // extern "C"
