// The reference's JIT element table (include/pyoomph_jit_table.h): init, per-element slots
// and the superlu() entry.
//
// A generated JIT_ELEMENT_init (src/codegen.cpp:6398-7090) describes one element's fields
// and fills function slots that the host calls per element with the element's nodal data
// (JITElementInfo_t, src/jitbridge.h:88-120) and its shape buffer. Here the slots run this
// library's device element kernel on that one element: the element's values, coordinates
// and local equations are staged into one small device block, the kernel (the same
// instantiation pyhip_assemble launches, so the same bits) writes the element vector and
// matrix in slot order, and the host adds them into R / J / M at the element's local
// equations, as the generated code's accumulation macros do (src/jitbridge_hang.h:30-184,
// row-major J with the shape info's jacobian_size as stride).
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "../../include/pyoomph_jit_table.h"
#include "problem.h"

namespace pyhip {

namespace {

constexpr int kMaxNodes = 27;
constexpr int kMaxSlots = 27;

// (node, value) of slot s; value = the field index in nodal_data
void jit_slot(int form, int s, int &node, int &val) {
    if (th_layout(form)) {
        static const int n[22] = {0, 0, 0, 1, 1, 2, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 6, 7, 7, 8, 8, 8};
        static const int v[22] = {0, 1, 2, 0, 1, 0, 1, 2, 0, 1, 0, 1, 0, 1, 0, 1, 2, 0, 1, 0, 1, 2};
        node = n[s];
        val = v[s];
    } else {
        node = s;
        val = 0;
    }
}

// Device block of one element: everything the kernel reads, in one allocation.
struct ElemBlock {
    int32_t conn[kMaxNodes];
    int32_t seqn[kMaxSlots];
    int32_t sub[1];
    int32_t pad;
    double coords[kMaxNodes * 3];
    double dofs[kMaxSlots];
    double pinned[kMaxNodes * 3];
    double aux[kMaxNodes];
    double params[kParamStride];
    double hist1[kMaxSlots], hist2[kMaxSlots];
    double Er[kMaxSlots];
    double Ej[kMaxSlots * kMaxSlots];
};

// State of one (form, family) instantiation: the table it was filled into (the
// reference's generated units keep theirs in a static pointer too), the form parameters,
// device buffers (made on the first slot call).
struct Instance {
    int form = 0, kind = 0, nn = 0, ns = 0, nq = 0;
    // run-time generated forms (include/pyoomph_codegen.h): kernels from the form's module,
    // the auxiliary field in nodal field 1 when the form reads one, the global parameter's
    // index among the form parameters
    bool generated = false, has_aux = false, pderiv = false;
    int gparam = 0;
    JITFuncSpec_Table_FiniteElement_t *table = nullptr;
    double form_params[kParamStride] = {};
    int nform_params = 0;
    ElemBlock *dev = nullptr;
    DevTables<double> *tab = nullptr;
    std::mutex mu;
};

Instance g_inst[5][5];   // [form][kind]

// generated forms: a fixed set of instances, each with its own slot functions (the slot
// signature carries no user pointer, so every table needs distinct function addresses)
constexpr int kGenInstances = 16;
Instance g_gen[kGenInstances];

int ensure_device(Instance &I) {
    if (I.dev) return PYHIP_OK;
    int64_t bytes = 0;
    int rc = dev_alloc(&I.dev, 1, bytes);
    if (rc) return rc;
    RefTables t;
    if (fill_ref_tables(I.kind, &t) != 0) {
        set_error("no tables for element family");
        return PYHIP_ERR_INVALID;
    }
    DevTables<double> h;
    memset(&h, 0, sizeof(h));
    for (int i = 0; i < 27; i++) h.w[i] = t.w[i];
    for (int i = 0; i < 27 * 27; i++) h.psi[i] = t.psi[i];
    for (int i = 0; i < 27 * 27 * 3; i++) h.dpsi[i] = t.dpsi[i];
    for (int i = 0; i < 27 * 4; i++) h.psi1[i] = t.psi1[i];
    if ((rc = dev_alloc(&I.tab, 1, bytes))) return rc;
    PYHIP_CHECK_HIP(hipMemcpy(I.tab, &h, sizeof(h), hipMemcpyHostToDevice));
    return PYHIP_OK;
}

// One element through the device kernel. mode: 0 steady R (+J with flag), 1 parameter
// derivative, 2 time-dependent (Taylor-Hood with the shape info's BDF2 weights).
// out_R [ns], out_J [ns*ns] in slot order.
int run_element(Instance &I, const JITElementInfo_t *ei, const JITShapeInfo_t *si, int flag, int deriv,
                bool transient, double *out_R, double *out_J) {
    int rc = ensure_device(I);
    if (rc) return rc;
    ElemBlock h;
    memset(&h, 0, sizeof(h));
    const int dim = elem_dim(I.kind);
    for (int l = 0; l < I.nn; l++) {
        h.conn[l] = l;
        for (int d = 0; d < dim; d++) h.coords[l * 3 + d] = ei->nodal_coords[l][d][0];
    }
    for (int s = 0; s < I.ns; s++) {
        int l, v;
        jit_slot(I.form, s, l, v);
        const double val = ei->nodal_data[l][v][0];
        h.pinned[l * 3 + v] = val;
        const bool free_ = ei->nodal_local_eqn[l][v] >= 0;
        h.seqn[s] = free_ ? s : -1;
        h.dofs[s] = val;
        if (transient) {
            h.hist1[s] = ei->nodal_data[l][v][1];
            h.hist2[s] = ei->nodal_data[l][v][2];
        }
    }
    if (I.form == PYHIP_FORM_REACTION_DIFFUSION || I.has_aux)
        for (int l = 0; l < I.nn; l++) h.aux[l] = ei->nodal_data[l][1][0];   // the source field f
    const JITFuncSpec_Table_FiniteElement_t *T = I.table;
    const double gp = (T && T->global_parameters && T->global_parameters[0]) ? *T->global_parameters[0] : 0.0;
    if (I.form == PYHIP_FORM_BRATU) {
        h.params[0] = gp;
    } else {
        for (int k = 0; k < I.nform_params; k++) h.params[k] = I.form_params[k];
        if (I.generated && I.pderiv) h.params[I.gparam] = gp;
    }
    PYHIP_CHECK_HIP(hipMemcpy(I.dev, &h, offsetof(ElemBlock, Er), hipMemcpyHostToDevice));
    ElemArgs<double> a{};
    a.nel = 1;
    a.e0 = 0;
    a.e1 = 1;
    a.conn = I.dev->conn;
    a.seqn = I.dev->seqn;
    a.sub = I.dev->sub;
    a.coords = I.dev->coords;
    a.dofs = I.dev->dofs;
    a.pinned = I.dev->pinned;
    a.aux = I.dev->aux;
    a.params = I.dev->params;
    a.pstride = kParamStride;
    a.nvalue = 3;
    a.tab = I.tab;
    a.Er = I.dev->Er;
    a.Ej = flag ? I.dev->Ej : nullptr;
    a.flag = flag;
    a.deriv = deriv;
    a.transient = transient ? 1 : 0;
    if (transient) {
        a.tw0 = si->timestepper_weights_dt_BDF2[0];
        a.tw1 = si->timestepper_weights_dt_BDF2[1];
        a.tw2 = si->timestepper_weights_dt_BDF2[2];
        a.hist1 = I.dev->hist1;
        a.hist2 = I.dev->hist2;
    }
    if ((rc = I.generated ? launch_generated_f64(I.form, a, nullptr) : launch_elements_f64(I.kind, I.form, a, nullptr)))
        return rc;
    PYHIP_CHECK_HIP(hipMemcpy(out_R, I.dev->Er, I.ns * sizeof(double), hipMemcpyDeviceToHost));
    if (flag) PYHIP_CHECK_HIP(hipMemcpy(out_J, I.dev->Ej, (size_t)I.ns * I.ns * sizeof(double), hipMemcpyDeviceToHost));
    return PYHIP_OK;
}

// Adds the slot-ordered element outputs into the caller's buffers at the local equations
// (pinned slots have none), J and M row-major with stride jacobian_size.
void scatter(const Instance &I, const JITElementInfo_t *ei, const JITShapeInfo_t *si, const double *er,
             const double *ej, double *R, double *J) {
    int leq[kMaxSlots];
    for (int s = 0; s < I.ns; s++) {
        int l, v;
        jit_slot(I.form, s, l, v);
        leq[s] = ei->nodal_local_eqn[l][v];
    }
    const unsigned stride = si ? si->jacobian_size : ei->ndof;
    for (int i = 0; i < I.ns; i++) {
        if (leq[i] < 0) continue;
        if (R) R[leq[i]] += er[i];
        if (!J || !ej) continue;
        for (int j = 0; j < I.ns; j++) {
            if (leq[j] < 0) continue;
            J[(size_t)leq[i] * stride + leq[j]] += ej[i * I.ns + j];
        }
    }
}

void report(int rc) {
    if (rc) fprintf(stderr, "[pyhip] JIT table slot failed: %s\n", pyhip_last_error());
}

// The reference signature: (eleminfo, shapeinfo, residuals, jacobian, mass_matrix, flag)
// with flag 0: R, 1: R and J, 2: R, J and M.
void res_jac(Instance &I, bool steady, const JITElementInfo_t *ei, const JITShapeInfo_t *si, double *R, double *J,
             double *M, unsigned flag) {
    std::lock_guard<std::mutex> lock(I.mu);
    // time derivatives exist in the Taylor-Hood momentum only (the other forms are steady)
    const bool tr = !steady && th_layout(I.form) && si && si->timestepper_weights_dt_BDF2;
    double er[kMaxSlots], ej[kMaxSlots * kMaxSlots];
    int rc = run_element(I, ei, si, flag >= 1 ? 1 : 0, 0, tr, er, ej);
    if (rc) return report(rc);
    scatter(I, ei, si, er, flag >= 1 ? ej : nullptr, R, flag >= 1 ? J : nullptr);
    if (flag >= 2 && M && tr) {
        double mr[kMaxSlots], mm[kMaxSlots * kMaxSlots];
        if ((rc = run_element(I, ei, si, 2, 0, true, mr, mm))) return report(rc);
        scatter(I, ei, si, mr, mm, nullptr, M);
    }
}

void dparam(Instance &I, const JITElementInfo_t *ei, const JITShapeInfo_t *si, double *R, double *J,
            unsigned flag) {
    std::lock_guard<std::mutex> lock(I.mu);
    double er[kMaxSlots], ej[kMaxSlots * kMaxSlots];
    int rc = run_element(I, ei, si, flag >= 1 ? 1 : 0, 1, false, er, ej);
    if (rc) return report(rc);
    scatter(I, ei, si, er, flag >= 1 ? ej : nullptr, R, flag >= 1 ? J : nullptr);
}

template <int FORM, int KIND, bool STEADY>
void slot_res_jac(const JITElementInfo_t *ei, const JITShapeInfo_t *si, double *R, double *J, double *M,
                  unsigned flag) {
    res_jac(g_inst[FORM][KIND], STEADY, ei, si, R, J, M, flag);
}

template <int FORM, int KIND>
void slot_dparam(const JITElementInfo_t *ei, const JITShapeInfo_t *si, double *R, double *J, double *M,
                 unsigned flag) {
    (void)M;
    dparam(g_inst[FORM][KIND], ei, si, R, J, flag);
}

template <int IDX, bool STEADY>
void gen_slot_res_jac(const JITElementInfo_t *ei, const JITShapeInfo_t *si, double *R, double *J, double *M,
                      unsigned flag) {
    res_jac(g_gen[IDX], STEADY, ei, si, R, J, M, flag);
}

template <int IDX>
void gen_slot_dparam(const JITElementInfo_t *ei, const JITShapeInfo_t *si, double *R, double *J, double *M,
                     unsigned flag) {
    (void)M;
    dparam(g_gen[IDX], ei, si, R, J, flag);
}

using Slot = JITFuncSpec_ResidualAndJacobian_FiniteElement;

template <int FORM, int KIND>
void slots_of(Slot &unsteady, Slot &steady, Slot &dparam) {
    unsteady = &slot_res_jac<FORM, KIND, false>;
    steady = &slot_res_jac<FORM, KIND, true>;
    dparam = &slot_dparam<FORM, KIND>;
}

bool pick_slots(int form, int kind, Slot &u, Slot &s, Slot &d) {
#define PYHIP_SLOT_CASE(F, K)                                                               \
    if (form == F && kind == K) {                                                           \
        slots_of<F, K>(u, s, d);                                                            \
        return true;                                                                        \
    }
    PYHIP_SLOT_CASE(PYHIP_FORM_POISSON, PYHIP_ELEM_TRI_P1)
    PYHIP_SLOT_CASE(PYHIP_FORM_POISSON, PYHIP_ELEM_TRI_P2)
    PYHIP_SLOT_CASE(PYHIP_FORM_POISSON, PYHIP_ELEM_QUAD_Q2)
    PYHIP_SLOT_CASE(PYHIP_FORM_POISSON, PYHIP_ELEM_HEX_Q2)
    PYHIP_SLOT_CASE(PYHIP_FORM_BRATU, PYHIP_ELEM_TRI_P1)
    PYHIP_SLOT_CASE(PYHIP_FORM_BRATU, PYHIP_ELEM_TRI_P2)
    PYHIP_SLOT_CASE(PYHIP_FORM_BRATU, PYHIP_ELEM_QUAD_Q2)
    PYHIP_SLOT_CASE(PYHIP_FORM_BRATU, PYHIP_ELEM_HEX_Q2)
    PYHIP_SLOT_CASE(PYHIP_FORM_REACTION_DIFFUSION, PYHIP_ELEM_TRI_P1)
    PYHIP_SLOT_CASE(PYHIP_FORM_REACTION_DIFFUSION, PYHIP_ELEM_TRI_P2)
    PYHIP_SLOT_CASE(PYHIP_FORM_REACTION_DIFFUSION, PYHIP_ELEM_QUAD_Q2)
    PYHIP_SLOT_CASE(PYHIP_FORM_REACTION_DIFFUSION, PYHIP_ELEM_HEX_Q2)
    PYHIP_SLOT_CASE(PYHIP_FORM_NAVIER_STOKES_TH, PYHIP_ELEM_QUAD_Q2)
#undef PYHIP_SLOT_CASE
    return false;
}

template <int... IDX>
void gen_slots_table(std::integer_sequence<int, IDX...>, int k, Slot &u, Slot &st, Slot &d) {
    static const Slot su[] = {&gen_slot_res_jac<IDX, false>...};
    static const Slot ss[] = {&gen_slot_res_jac<IDX, true>...};
    static const Slot sd[] = {&gen_slot_dparam<IDX>...};
    u = su[k];
    st = ss[k];
    d = sd[k];
}

void pick_gen_slots(int k, Slot &u, Slot &s, Slot &d) {
    gen_slots_table(std::make_integer_sequence<int, kGenInstances>{}, k, u, s, d);
}

char *dupstr(const char *s) {
    char *d = (char *)malloc(strlen(s) + 1);
    if (d) strcpy(d, s);
    return d;
}

char **names(std::initializer_list<const char *> l) {
    char **a = (char **)calloc(l.size() + 1, sizeof(char *));
    int i = 0;
    for (const char *s : l) a[i++] = dupstr(s);
    return a;
}

void free_names(char **a, unsigned n) {
    if (!a) return;
    for (unsigned i = 0; i < n; i++) free(a[i]);
    free(a);
}

// frees what pyhip_jit_element_init allocated (the host then deletes the table,
// src/problem.cpp:144-165)
void clean_up(JITFuncSpec_Table_FiniteElement_t *t) {
    if (!t) return;
    free_names(t->fieldnames_C1, t->numfields_C1);
    free_names(t->fieldnames_C2, t->numfields_C2);
    free_names(t->res_jac_names, t->num_res_jacs);
    free(t->ResidualAndJacobian);
    free(t->ResidualAndJacobianSteady);
    free(t->ResidualAndJacobian_NoHang);
    free(t->missing_residual_assembly);
    free(t->shapes_required_ResJac);
    if (t->ParameterDerivative) {
        for (unsigned r = 0; r < t->num_res_jacs; r++) free(t->ParameterDerivative[r]);
        free(t->ParameterDerivative);
    }
    free(t->global_paramindices);
    free(t->global_parameters);
    free(t->dominant_space);
    t->fieldnames_C1 = t->fieldnames_C2 = t->res_jac_names = nullptr;
    t->ResidualAndJacobian = t->ResidualAndJacobianSteady = t->ResidualAndJacobian_NoHang = nullptr;
    t->missing_residual_assembly = nullptr;
    t->shapes_required_ResJac = nullptr;
    t->ParameterDerivative = nullptr;
    t->global_paramindices = nullptr;
    t->global_parameters = nullptr;
    t->dominant_space = nullptr;
    for (auto &row : g_inst)
        for (auto &I : row)
            if (I.table == t) I.table = nullptr;
    for (auto &I : g_gen)
        if (I.table == t) I.table = nullptr;
}

}  // namespace

}  // namespace pyhip

using namespace pyhip;

extern "C" {

int pyhip_jit_element_init(int form, int elem_kind, JITFuncSpec_Table_FiniteElement_t *t) {
    PYHIP_REQUIRE(t, "null table");
    const bool gen = gen_form_known(form);
    PYHIP_REQUIRE(((form >= 1 && form <= 4) || gen) && elem_nnode(elem_kind) > 0,
                  "unknown weak form or element family");
    Slot su = nullptr, ss = nullptr, sd = nullptr;
    std::string field = "u", aux_field;
    int gparam = 0;
    Instance *ip = nullptr;
    if (gen) {
        PYHIP_REQUIRE(gen_form_kind(form) == elem_kind, "generated form compiled for another element family");
        (void)gen_form_fields(form, field, aux_field, gparam);
        // the instance of this form, else a free one
        int k = -1;
        for (int i = 0; i < kGenInstances && k < 0; i++)
            if (g_gen[i].generated && g_gen[i].form == form) k = i;
        for (int i = 0; i < kGenInstances && k < 0; i++)
            if (!g_gen[i].generated) k = i;
        PYHIP_REQUIRE(k >= 0, "JIT tables of at most 16 generated forms");
        pick_gen_slots(k, su, ss, sd);
        ip = &g_gen[k];
    } else {
        PYHIP_REQUIRE(pick_slots(form, elem_kind, su, ss, sd), "weak form not available on this element family");
        ip = &g_inst[form][elem_kind];
    }
    Instance &I = *ip;
    std::lock_guard<std::mutex> lock(I.mu);
    I.form = form;
    I.kind = elem_kind;
    I.nn = elem_nnode(elem_kind);
    I.ns = th_layout(form) ? 22 : I.nn;
    I.table = t;
    I.generated = gen;
    I.has_aux = gen && !aux_field.empty();
    I.pderiv = gen && gen_form_has_pderiv(form);
    I.gparam = gparam;
    const bool th = th_layout(form);
    const bool p1 = elem_kind == PYHIP_ELEM_TRI_P1;
    t->nodal_dim = (unsigned)elem_dim(elem_kind);
    t->lagr_dim = 0;
    // fields and their nodal-data order: C2 fields first, then C1 (Taylor-Hood: velocity in
    // C2 at offset 0, pressure in C1 at offset 2; scalar forms: "u" in the family's space,
    // reaction-diffusion adds its source field "f", pinned, after it; generated forms their
    // field and auxiliary field names)
    const bool rd = form == PYHIP_FORM_REACTION_DIFFUSION || I.has_aux;
    if (gen) {
        if (aux_field.empty()) aux_field = "f";
    } else {
        aux_field = "f";
    }
    if (th) {
        t->numfields_C2 = t->numfields_C2_bulk = t->numfields_C2_basebulk = t->numfields_C2_new = 2;
        t->fieldnames_C2 = names({"velocity_x", "velocity_y"});
        t->numfields_C1 = t->numfields_C1_bulk = t->numfields_C1_basebulk = t->numfields_C1_new = 1;
        t->fieldnames_C1 = names({"pressure"});
        t->buffer_offset_C2_basebulk = 0;
        t->buffer_offset_C1_basebulk = 2;
        t->nodal_offset_C2_basebulk = 0;
        t->nodal_offset_C1_basebulk = 2;
    } else if (p1) {
        const unsigned nf = rd ? 2 : 1;
        t->numfields_C1 = t->numfields_C1_bulk = t->numfields_C1_basebulk = t->numfields_C1_new = nf;
        t->fieldnames_C1 = rd ? names({field.c_str(), aux_field.c_str()}) : names({field.c_str()});
    } else {
        const unsigned nf = rd ? 2 : 1;
        t->numfields_C2 = t->numfields_C2_bulk = t->numfields_C2_basebulk = t->numfields_C2_new = nf;
        t->fieldnames_C2 = rd ? names({field.c_str(), aux_field.c_str()}) : names({field.c_str()});
    }
    t->num_res_jacs = 1;
    t->current_res_jac = 0;
    t->res_jac_names = names({""});
    t->ResidualAndJacobian = (Slot *)calloc(1, sizeof(Slot));
    t->ResidualAndJacobianSteady = (Slot *)calloc(1, sizeof(Slot));
    t->ResidualAndJacobian_NoHang = (Slot *)calloc(1, sizeof(Slot));   // no hanging nodes here
    t->missing_residual_assembly = (bool *)calloc(1, sizeof(bool));
    t->ResidualAndJacobian[0] = su;
    t->ResidualAndJacobianSteady[0] = ss;
    t->ResidualAndJacobian_NoHang[0] = su;
    // required shapes: the family's space with Eulerian derivatives (Taylor-Hood also C1)
    t->shapes_required_ResJac =
        (JITFuncSpec_RequiredShapes_FiniteElement_t *)calloc(1, sizeof(JITFuncSpec_RequiredShapes_FiniteElement_t));
    JITFuncSpec_RequiredShapes_FiniteElement_t &rs = *t->shapes_required_ResJac;
    if (p1) rs.psi_C1 = rs.dx_psi_C1 = true;
    else rs.psi_C2 = rs.dx_psi_C2 = true;
    if (th) rs.psi_C1 = true;
    rs.psi_Pos = rs.dx_psi_Pos = true;
    t->merged_required_shapes = rs;
    if (form == PYHIP_FORM_BRATU || I.pderiv) {
        t->numglobal_params = 1;
        t->global_paramindices = (unsigned *)calloc(1, sizeof(unsigned));
        t->global_parameters = (double **)calloc(1, sizeof(double *));
        t->ParameterDerivative = (Slot **)calloc(1, sizeof(Slot *));
        t->ParameterDerivative[0] = (Slot *)calloc(1, sizeof(Slot));
        t->ParameterDerivative[0][0] = sd;
    }
    t->dominant_space = dupstr(p1 ? "C1" : "C2");
    t->max_dt_order = th ? 1 : 0;
    t->integration_order = 0;
    t->moving_nodes = false;
    t->fd_jacobian = false;
    t->hessian_generated = false;
    t->clean_up = &clean_up;
    return PYHIP_OK;
}

int pyhip_jit_set_form_params(JITFuncSpec_Table_FiniteElement_t *t, const double *params, int nparams) {
    PYHIP_REQUIRE(t && (params || nparams == 0), "null argument");
    PYHIP_REQUIRE(nparams >= 0 && nparams <= kParamStride, "too many parameters");
    auto set = [&](Instance &I) {
        std::lock_guard<std::mutex> lock(I.mu);
        for (int k = 0; k < nparams; k++) I.form_params[k] = params[k];
        I.nform_params = nparams;
    };
    for (auto &row : g_inst)
        for (auto &I : row)
            if (I.table == t) {
                set(I);
                return PYHIP_OK;
            }
    for (auto &I : g_gen)
        if (I.table == t) {
            set(I);
            return PYHIP_OK;
        }
    set_error("table not filled by pyhip_jit_element_init (or cleaned up)");
    return PYHIP_ERR_INVALID;
}

int pyhip_superlu(int *op_flag, int *n, int *nnz, int *nrhs, double *values, int *rowind, int *colptr, double *b,
                  int *ldb, int *transpose, int *doc, void **f_factors, int *info) {
    (void)doc;
    if (info) *info = 0;
    PYHIP_REQUIRE(op_flag && f_factors && info, "null argument");
    const int op = *op_flag;
    const int nn = n ? *n : 0;
    int rc;
    if (op == 1) {
        PYHIP_REQUIRE(n && nnz, "null size");
        rc = pyhip_solve_serial(1, nn, *nnz, 0, values, rowind, colptr, nullptr, nn, transpose ? 1 : 0,
                                f_factors);
    } else if (op == 2) {
        const int nr = nrhs ? *nrhs : 1;
        const int ld = (ldb && *ldb >= nn) ? *ldb : nn;
        rc = pyhip_solve_serial(2, nn, 0, nr, nullptr, nullptr, nullptr, b, ld, transpose ? 1 : 0, f_factors);
    } else if (op == 3) {
        rc = pyhip_solve_serial(3, nn, 0, 0, nullptr, nullptr, nullptr, nullptr, nn, 0, f_factors);
    } else {
        set_error("op_flag must be 1 (factorise), 2 (solve) or 3 (free)");
        rc = PYHIP_ERR_INVALID;
    }
    if (rc != PYHIP_OK) {
        *info = -rc;
        return rc;
    }
    return 0;   // sign-of-determinant slot of op 1: not known to a Krylov solve
}

}  // extern "C"
