// C ABI: element table, problem set-up, assembly and reference-CSR export.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "problem.h"

namespace pyhip {

static thread_local std::string g_last_error;
void set_error(const std::string &msg) { g_last_error = msg; }

static int form_nslot(int form, int kind) {
    if (th_layout(form)) return 22;
    return elem_nnode(kind);
}
static int form_nvalue(int form) { return th_layout(form) ? 3 : 1; }

// (node, value) of each slot (oomph local dof order). Taylor-Hood: see element_kernels.h.
static void slot_map(int form, int kind, int s, int &node, int &val) {
    if (th_layout(form)) {
        static const int n[22] = {0, 0, 0, 1, 1, 2, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 6, 7, 7, 8, 8, 8};
        static const int v[22] = {0, 1, 2, 0, 1, 0, 1, 2, 0, 1, 0, 1, 0, 1, 0, 1, 2, 0, 1, 0, 1, 2};
        node = n[s];
        val = v[s];
    } else {
        node = s;
        val = 0;
    }
    (void)kind;
}

template <typename T>
static int upload_tables(int kind, DevTables<T> **out, int64_t &bytes) {
    RefTables t;
    if (fill_ref_tables(kind, &t) != 0) {
        set_error("no tables for element family");
        return PYHIP_ERR_INVALID;
    }
    DevTables<T> h;
    memset(&h, 0, sizeof(h));
    for (int i = 0; i < 27; i++) h.w[i] = T(t.w[i]);
    for (int i = 0; i < 27 * 27; i++) h.psi[i] = T(t.psi[i]);
    for (int i = 0; i < 27 * 27 * 3; i++) h.dpsi[i] = T(t.dpsi[i]);
    for (int i = 0; i < 27 * 4; i++) h.psi1[i] = T(t.psi1[i]);
    int rc = dev_alloc(out, 1, bytes);
    if (rc) return rc;
    PYHIP_CHECK_HIP(hipMemcpy(*out, &h, sizeof(h), hipMemcpyHostToDevice));
    return PYHIP_OK;
}

}  // namespace pyhip

using namespace pyhip;

extern "C" {

const char *pyhip_last_error(void) { return g_last_error.c_str(); }
int pyhip_version(void) { return 1; }
int pyhip_device_count(int *count) {
    PYHIP_CHECK_HIP(hipGetDeviceCount(count));
    return PYHIP_OK;
}

int pyhip_element_init(int form, int elem_kind, pyhip_element_table *t) {
    PYHIP_REQUIRE(t, "null table");
    PYHIP_REQUIRE(elem_nnode(elem_kind) > 0, "unknown element family");
    PYHIP_REQUIRE(form >= 1 && form <= 4, "unknown weak form");
    PYHIP_REQUIRE(form != PYHIP_FORM_NAVIER_STOKES_TH || elem_kind == PYHIP_ELEM_QUAD_Q2,
                  "Taylor-Hood needs Q2 quadrilaterals");
    memset(t, 0, sizeof(*t));
    t->form = form;
    t->elem_kind = elem_kind;
    t->num_res_jacs = 1;
    t->res_jac_names[0] = "";
    t->nodal_values = form_nvalue(form);
    t->nslot = form_nslot(form, elem_kind);
    t->nnode = elem_nnode(elem_kind);
    t->dim = elem_dim(elem_kind);
    t->integration_order = 0;
    RefTables rt;
    fill_ref_tables(elem_kind, &rt);
    t->nqp = rt.nqp;
    if (form == PYHIP_FORM_BRATU) {
        t->numglobal_params = 1;
        t->global_param_names[0] = "lambda";
        t->has_param_derivative[0] = 1;
    }
    t->moving_nodes = 0;
    return PYHIP_OK;
}

// face element kind of a bulk element kind
static int face_kind(int elem_kind, int form) {
    if (th_layout(form)) return kKindLine3TH;
    switch (elem_kind) {
    case PYHIP_ELEM_TRI_P1: return kKindLine2;
    case PYHIP_ELEM_QUAD_Q2: return kKindLine3Q;
    case PYHIP_ELEM_HEX_Q2: return kKindQuad9;
    default: return kKindLine3;
    }
}

// The multigrid's element level of one sub-problem (amg_element_level_rows): scalar forms
// without EnforcedBC multipliers; with `velocity`, also the two velocity components of the
// Taylor-Hood layout (the block preconditioner's velocity V-cycle)
static void element_level_rows(const pyhip_subproblem_desc &d, AmgFixedInterp &F, bool velocity = false) {
    if (velocity && th_layout(d.form)) {
        amg_element_level_rows(d.elem_kind, d.nnode, d.nelem, d.conn, d.eqn, d.nvalue, 2, d.ndof, true, F);
        return;
    }
    const bool scalar = d.nvalue == 1 && !th_layout(d.form) && d.nenforced == 0;
    amg_element_level_rows(d.elem_kind, d.nnode, d.nelem, d.conn, d.eqn, 1, 1, d.ndof, scalar, F);
}

// face element kind of the EnforcedBC faces of a bulk element kind (-1: not built)
static int enforced_face_kind(int elem_kind) {
    switch (elem_kind) {
    case PYHIP_ELEM_TRI_P1: return kKindLine2E;
    case PYHIP_ELEM_TRI_P2: return kKindLine3E;
    case PYHIP_ELEM_QUAD_Q2: return kKindLine3QE;
    }
    return -1;
}

int pyhip_problem_create(const pyhip_subproblem_desc *subs, int nsub, void *stream,
                         pyhip_problem **out) {
    PYHIP_REQUIRE(subs && nsub > 0 && out, "invalid arguments");
    *out = nullptr;
    std::unique_ptr<pyhip_problem> P(new pyhip_problem());
    P->st = (hipStream_t)stream;
    P->nsub = nsub;
    // validate, count, group by (form, kind) in first-appearance order
    std::vector<std::pair<int, int>> gkeys;
    for (int s = 0; s < nsub; s++) {
        const auto &d = subs[s];
        PYHIP_REQUIRE(elem_nnode(d.elem_kind) > 0, "unknown element family");
        PYHIP_REQUIRE((d.form >= 1 && d.form <= 4) ||
                          (gen_form_known(d.form) && !gen_face_form(d.form) && !gen_constraint_form(d.form)),
                      "unknown weak form");
        PYHIP_REQUIRE(d.enforced_form == 0 || (gen_constraint_form(d.enforced_form) &&
                                               gen_form_kind(d.enforced_form) == d.elem_kind),
                      "enforced_form: a generated constraint form compiled for this element family");
        PYHIP_REQUIRE(!gen_form_known(d.form) || gen_form_kind(d.form) == d.elem_kind,
                      "generated form compiled for another element family");
        PYHIP_REQUIRE(d.form != PYHIP_FORM_NAVIER_STOKES_TH || d.elem_kind == PYHIP_ELEM_QUAD_Q2,
                      "Taylor-Hood needs Q2 quadrilaterals");
        PYHIP_REQUIRE(d.nenforced >= 0 && (d.nenforced == 0 || (d.enforced_conn && (d.enforced_coef ||
                                                                                   (d.enforced_form != 0 &&
                                                                                    gen_form_ncoef(d.enforced_form) == 0)))),
                      "EnforcedBC faces without connectivity or coefficients");
        PYHIP_REQUIRE(d.nenforced == 0 || (!th_layout(d.form) && enforced_face_kind(d.elem_kind) > 0),
                      "EnforcedBC multipliers: scalar forms on P1 / P2 triangles and Q2 quadrilaterals");
        PYHIP_REQUIRE(d.nvalue == form_nvalue(d.form) + (d.nenforced > 0 ? 1 : 0),
                      "values per node do not match the form (EnforcedBC adds the multiplier value)");
        PYHIP_REQUIRE(d.nnode > 0 && d.nelem > 0 && d.coords && d.conn && d.eqn && d.pinned_values,
                      "empty mesh or missing array");
        PYHIP_REQUIRE(d.nparams >= 0 && d.nparams <= kParamStride, "too many parameters");
        PYHIP_REQUIRE(d.form != PYHIP_FORM_REACTION_DIFFUSION || d.aux,
                      "reaction-diffusion needs the nodal source field (aux)");
        PYHIP_REQUIRE(d.ndof >= 0, "negative ndof");
        PYHIP_REQUIRE(d.face_form == 0 || (gen_face_form(d.face_form) && gen_form_kind(d.face_form) == d.elem_kind &&
                                           !th_layout(d.form)),
                      "face_form: a generated boundary-face form compiled for this element family, on a scalar form");
        PYHIP_REQUIRE(d.nface >= 0 && (d.nface == 0 || (d.face_conn && (d.face_coef || (d.face_form != 0 &&
                                                                                       gen_form_ncoef(d.face_form) == 0)))),
                      "faces without connectivity or flux coefficients");
        PYHIP_REQUIRE(d.nface == 0 || !th_layout(d.form) || d.nvalue == 3,
                      "boundary-face flux elements: scalar forms and Taylor-Hood tractions");
        std::pair<int, int> k(d.form, d.elem_kind);
        if (std::find(gkeys.begin(), gkeys.end(), k) == gkeys.end()) gkeys.push_back(k);
    }
    // face groups after the bulk groups, one per face kind (flux faces, then EnforcedBC faces)
    for (int s = 0; s < nsub; s++) {
        if (subs[s].nface == 0) continue;
        std::pair<int, int> k(subs[s].face_form ? subs[s].face_form : kFormFlux,
                              face_kind(subs[s].elem_kind, subs[s].form));
        if (std::find(gkeys.begin(), gkeys.end(), k) == gkeys.end()) gkeys.push_back(k);
    }
    for (int s = 0; s < nsub; s++) {
        if (subs[s].nenforced == 0) continue;
        std::pair<int, int> k(subs[s].enforced_form ? subs[s].enforced_form : kFormFlux,
                              enforced_face_kind(subs[s].elem_kind));
        if (std::find(gkeys.begin(), gkeys.end(), k) == gkeys.end()) gkeys.push_back(k);
    }
    P->groups.resize(gkeys.size());
    for (size_t g = 0; g < gkeys.size(); g++) {
        Group &G = P->groups[g];
        G.form = gkeys[g].first;
        G.kind = gkeys[g].second;
        if (face_group_form(G.form)) {
            G.nn = (G.kind == kKindLine2 || G.kind == kKindLine2E) ? 2 : (G.kind == kKindQuad9 ? 9 : 3);
            G.ns = (G.kind == kKindLine3TH || enforced_kind(G.kind)) ? 2 * G.nn : G.nn;
            G.dim = G.kind == kKindQuad9 ? 3 : 2;
            continue;
        }
        G.ns = form_nslot(G.form, G.kind);
        G.nn = elem_nnode(G.kind);
        G.dim = elem_dim(G.kind);
    }
    for (auto &G : P->groups) G.active = G.form;
    P->subs.resize(nsub);
    P->dof_off_h.assign(nsub + 1, 0);
    int64_t node_off = 0;
    for (int s = 0; s < nsub; s++) {
        const auto &d = subs[s];
        SubInfo &S = P->subs[s];
        S.form = d.form;
        S.kind = d.elem_kind;
        S.nvalue = d.nvalue;
        S.dof_off = P->dof_off_h[s];
        S.ndof = d.ndof;
        S.node_off = node_off;
        S.nnode = d.nnode;
        S.group = (int)(std::find(gkeys.begin(), gkeys.end(), std::make_pair(d.form, d.elem_kind)) -
                        gkeys.begin());
        S.el_off = P->groups[S.group].nel;
        S.nel = d.nelem;
        P->groups[S.group].nel += d.nelem;
        if (d.nface > 0) {
            const std::pair<int, int> fk(d.face_form ? d.face_form : kFormFlux, face_kind(d.elem_kind, d.form));
            S.fgroup = (int)(std::find(gkeys.begin(), gkeys.end(), fk) - gkeys.begin());
            S.fel_off = P->groups[S.fgroup].nel;
            S.nface = d.nface;
            P->groups[S.fgroup].nel += d.nface;
        }
        if (d.nenforced > 0) {
            const std::pair<int, int> ek(d.enforced_form ? d.enforced_form : kFormFlux, enforced_face_kind(d.elem_kind));
            S.egroup = (int)(std::find(gkeys.begin(), gkeys.end(), ek) - gkeys.begin());
            S.eel_off = P->groups[S.egroup].nel;
            S.nenf = d.nenforced;
            P->groups[S.egroup].nel += d.nenforced;
        }
        P->dof_off_h[s + 1] = S.dof_off + d.ndof;
        node_off += d.nnode;
    }
    P->ndof = P->dof_off_h[nsub];
    P->nnode = node_off;
    PYHIP_REQUIRE(P->ndof < INT32_MAX, "too many dofs");
    int64_t er = 0, ej = 0;
    for (auto &G : P->groups) {
        G.er_off = er;
        G.ej_off = ej;
        er += G.nel * G.ns;
        ej += G.nel * G.ns * G.ns;
    }
    P->n_er = er;
    P->n_ej = ej;

    // host staging: element-major global equations per group, slot-major copies
    std::vector<std::vector<int32_t>> seqn_em(P->groups.size()), conn_sm(P->groups.size()),
        seqn_sm(P->groups.size()), sub_g(P->groups.size());
    std::vector<std::vector<double>> coef_g(P->groups.size());
    for (size_t g = 0; g < P->groups.size(); g++) {
        const Group &G = P->groups[g];
        seqn_em[g].resize(G.nel * G.ns);
        conn_sm[g].resize(G.nn * G.nel);
        seqn_sm[g].resize(G.ns * G.nel);
        sub_g[g].resize(G.nel);
        if (face_group_form(G.form)) coef_g[g].resize((size_t)group_ncoef(G.form, G.kind) * G.nel);
    }
    std::vector<double> coords_h(P->nnode * 3, 0.0), pinned_h(P->nnode * 3, 0.0),
        aux_h(P->nnode, 0.0);
    P->params_h.assign((size_t)nsub * kParamStride, 0.0);
    P->dof_val_h.assign(P->ndof, 0);
    P->pmass_h.assign(P->ndof, 0.0);
    for (int s = 0; s < nsub; s++) {
        const auto &d = subs[s];
        const SubInfo &S = P->subs[s];
        const Group &G = P->groups[S.group];
        const int dim = G.dim;
        for (int64_t n = 0; n < d.nnode; n++) {
            for (int k = 0; k < dim; k++) coords_h[(S.node_off + n) * 3 + k] = d.coords[n * dim + k];
            for (int v = 0; v < d.nvalue; v++) {
                pinned_h[(S.node_off + n) * 3 + v] = d.pinned_values[n * d.nvalue + v];
                const int32_t q = d.eqn[n * d.nvalue + v];
                PYHIP_REQUIRE(q >= -1 && q < d.ndof, "equation number out of range");
                if (q >= 0) P->dof_val_h[S.dof_off + q] = (int8_t)v;
            }
            if (d.aux) aux_h[S.node_off + n] = d.aux[n];
        }
        if (th_layout(d.form)) {
            // lumped Q1 pressure mass: a quarter of each element's area (corners 0, 2, 8, 6
            // of the 3x3 node block, counter-clockwise) to each corner's pressure unknown
            static const int corner[4] = {0, 2, 8, 6};
            for (int64_t e = 0; e < d.nelem; e++) {
                double area2 = 0.0;
                for (int c = 0; c < 4; c++) {
                    const int32_t a = d.conn[e * G.nn + corner[c]], b = d.conn[e * G.nn + corner[(c + 1) % 4]];
                    area2 += d.coords[a * dim] * d.coords[b * dim + 1] - d.coords[b * dim] * d.coords[a * dim + 1];
                }
                const double quarter = 0.125 * std::fabs(area2);
                for (int c = 0; c < 4; c++) {
                    const int32_t q = d.eqn[(int64_t)d.conn[e * G.nn + corner[c]] * d.nvalue + 2];
                    if (q >= 0) P->pmass_h[S.dof_off + q] += quarter;
                }
            }
        }
        for (int k = 0; k < d.nparams; k++) P->params_h[(size_t)s * kParamStride + k] = d.params[k];
        // face sets: the flux faces, then the EnforcedBC faces (slots: the node's values 0
        // and 1 per face node, u and the multiplier)
        for (int fs = 0; fs < 2; fs++) {
            const int fg = fs == 0 ? S.fgroup : S.egroup;
            const int64_t nf = fs == 0 ? S.nface : S.nenf, foff = fs == 0 ? S.fel_off : S.eel_off;
            const int32_t *fconn = fs == 0 ? d.face_conn : d.enforced_conn;
            const double *fcoef = fs == 0 ? d.face_coef : d.enforced_coef;
            if (nf == 0) continue;
            const Group &F = P->groups[fg];
            const int nc = group_ncoef(F.form, F.kind), vps = F.ns / F.nn;   // values per face node
            for (int64_t f = 0; f < nf; f++) {
                const int64_t ge = foff + f;
                sub_g[fg][ge] = s;
                for (int k = 0; k < nc; k++) coef_g[fg][nc * ge + k] = fcoef[nc * f + k];
                for (int l = 0; l < F.nn; l++) {
                    const int32_t n = fconn[f * F.nn + l];
                    PYHIP_REQUIRE(n >= 0 && n < d.nnode, "face connectivity out of range");
                    conn_sm[fg][(int64_t)l * F.nel + ge] = (int32_t)(S.node_off + n);
                    for (int c = 0; c < vps; c++) {
                        const int32_t q = d.eqn[(int64_t)n * d.nvalue + c];
                        const int32_t gq = q >= 0 ? (int32_t)(q + S.dof_off) : -1;
                        const int sl = l * vps + c;
                        seqn_em[fg][ge * F.ns + sl] = gq;
                        seqn_sm[fg][(int64_t)sl * F.nel + ge] = gq;
                    }
                }
            }
        }
        std::vector<int32_t> &sem = seqn_em[S.group];
        for (int64_t e = 0; e < d.nelem; e++) {
            const int64_t ge = S.el_off + e;
            sub_g[S.group][ge] = s;
            for (int l = 0; l < G.nn; l++) {
                const int32_t n = d.conn[e * G.nn + l];
                PYHIP_REQUIRE(n >= 0 && n < d.nnode, "connectivity out of range");
                conn_sm[S.group][(int64_t)l * G.nel + ge] = (int32_t)(S.node_off + n);
            }
            for (int sl = 0; sl < G.ns; sl++) {
                int ln, lv;
                slot_map(G.form, G.kind, sl, ln, lv);
                const int32_t n = d.conn[e * G.nn + ln];
                const int32_t q = d.eqn[(int64_t)n * d.nvalue + lv];
                const int32_t gq = q >= 0 ? (int32_t)(q + S.dof_off) : -1;
                sem[ge * G.ns + sl] = gq;
                seqn_sm[S.group][(int64_t)sl * G.nel + ge] = gq;
            }
        }
    }
    // the multigrid's element level, rows over all unknowns; with Taylor-Hood sub-problems
    // also the block preconditioner's (velocity components interpolated too)
    P->amg_fix = AmgFixedInterp();
    P->amg_fix_vel = AmgFixedInterp();
    bool any_th = false;
    for (int s = 0; s < nsub; s++) {
        element_level_rows(subs[s], P->amg_fix);
        any_th = any_th || th_layout(subs[s].form);
    }
    P->amg_fix.rs.push_back((int32_t)P->amg_fix.col.size());
    if (any_th) {
        for (int s = 0; s < nsub; s++) element_level_rows(subs[s], P->amg_fix_vel, true);
        P->amg_fix_vel.rs.push_back((int32_t)P->amg_fix_vel.col.size());
    }
    // pattern: element blocks in element order of each sub-problem
    std::vector<ElementBlock> blocks;
    for (int s = 0; s < nsub; s++) {
        const SubInfo &S = P->subs[s];
        const Group &G = P->groups[S.group];
        ElementBlock B;
        B.ns = G.ns;
        B.nel = S.nel;
        B.seqn = seqn_em[S.group].data() + S.el_off * G.ns;
        B.er_off = G.er_off + S.el_off * G.ns;
        B.ej_off = G.ej_off + S.el_off * G.ns * G.ns;
        blocks.push_back(B);
        // the interface elements follow the bulk elements of their mesh: flux faces, then
        // EnforcedBC faces
        for (int fs = 0; fs < 2; fs++) {
            const int fg = fs == 0 ? S.fgroup : S.egroup;
            const int64_t nf = fs == 0 ? S.nface : S.nenf, foff = fs == 0 ? S.fel_off : S.eel_off;
            if (nf == 0) continue;
            const Group &F = P->groups[fg];
            ElementBlock FB;
            FB.ns = F.ns;
            FB.nel = nf;
            FB.seqn = seqn_em[fg].data() + foff * F.ns;
            FB.er_off = F.er_off + foff * F.ns;
            FB.ej_off = F.ej_off + foff * F.ns * F.ns;
            blocks.push_back(FB);
        }
    }
    HostPattern hp;
    std::string err;
    if (build_pattern(P->ndof, blocks, hp, err) != 0) {
        set_error("pattern: " + err);
        return PYHIP_ERR_INVALID;
    }
    P->nnz = hp.nnz;
    P->nocc = hp.nocc;
    P->ncontrib = hp.ncontrib;
    std::vector<int32_t> dof_sub_h(P->ndof);
    for (int s = 0; s < nsub; s++)
        for (int64_t r = P->dof_off_h[s]; r < P->dof_off_h[s + 1]; r++) dof_sub_h[r] = s;

    // uploads
    int64_t &B = P->device_bytes;
    int rc;
#define UP(dst, src, n)                                                                   \
    do {                                                                                  \
        rc = dev_alloc(&(dst), (n), B);                                                   \
        if (rc) return rc;                                                                \
        if ((n) > 0)                                                                      \
            PYHIP_CHECK_HIP(hipMemcpy((dst), (src), (size_t)(n) * sizeof(*(dst)),         \
                                      hipMemcpyHostToDevice));                            \
    } while (0)
    UP(P->coords, coords_h.data(), coords_h.size());
    UP(P->pinned, pinned_h.data(), pinned_h.size());
    UP(P->aux, aux_h.data(), aux_h.size());
    UP(P->params, P->params_h.data(), P->params_h.size());
    for (size_t g = 0; g < P->groups.size(); g++) {
        Group &G = P->groups[g];
        UP(G.conn, conn_sm[g].data(), conn_sm[g].size());
        UP(G.seqn, seqn_sm[g].data(), seqn_sm[g].size());
        UP(G.sub, sub_g[g].data(), sub_g[g].size());
        if (face_group_form(G.form)) UP(G.coef, coef_g[g].data(), coef_g[g].size());
        // the faces of Q2 bricks use the Q2 quadrilateral tables
        const int tk = !face_group_form(G.form) ? G.kind : (G.kind == kKindQuad9 ? (int)ELEM_QUAD_Q2 : -1);
        if (tk >= 0 && !P->tab64[tk]) {
            rc = upload_tables<double>(tk, &P->tab64[tk], B);
            if (rc) return rc;
            rc = upload_tables<float>(tk, &P->tab32[tk], B);
            if (rc) return rc;
        }
    }
    UP(P->row_start, hp.row_start.data(), hp.row_start.size());
    UP(P->col, hp.col.data(), hp.col.size());
    UP(P->diag, hp.diag_pos.data(), hp.diag_pos.size());
    UP(P->occ_ptr, hp.occ_ptr.data(), hp.occ_ptr.size());
    UP(P->occ_roff, hp.occ_roff.data(), hp.occ_roff.size());
    UP(P->cptr, hp.cptr.data(), hp.cptr.size());
    UP(P->cidx, hp.cidx.data(), hp.cidx.size());
    UP(P->dof_sub, dof_sub_h.data(), dof_sub_h.size());
    UP(P->dof_off, P->dof_off_h.data(), P->dof_off_h.size());
#undef UP
    P->row_start_h = std::move(hp.row_start);
    P->col_h = std::move(hp.col);
    *out = P.release();
    return PYHIP_OK;
}

void pyhip_krylov_free(pyhip_problem *p);
int64_t pyhip_krylov_bytes(const pyhip_problem *p);

int pyhip_problem_destroy(pyhip_problem *p) {
    if (!p) return PYHIP_OK;
    pyhip_krylov_free(p);
    for (auto &G : p->groups) {
        dfree(G.conn);
        dfree(G.seqn);
        dfree(G.sub);
        dfree(G.coef);
    }
    void *ptrs[] = {p->coords, p->pinned, p->aux, p->params, p->row_start, p->col, p->diag,
                    p->occ_ptr, p->occ_roff, p->cptr, p->cidx,
                    p->dof_sub, p->dof_off,
                    p->Er64, p->Ej64, p->Er32, p->Ej32, p->keys, p->kept, p->ma_vals, p->ma_res};
    for (void *q : ptrs)
        dfree(q);
    for (int k = 0; k < 5; k++) {
        dfree(p->tab64[k]);
        dfree(p->tab32[k]);
    }
    delete p;
    return PYHIP_OK;
}

int pyhip_problem_info(const pyhip_problem *p, int64_t *ndof, int64_t *nnz, int64_t *ncontrib,
                       int64_t *device_bytes) {
    PYHIP_REQUIRE(p, "null problem");
    if (ndof) *ndof = p->ndof;
    if (nnz) *nnz = p->nnz;
    if (ncontrib) *ncontrib = p->ncontrib;
    if (device_bytes) *device_bytes = p->device_bytes + pyhip_krylov_bytes(p);
    return PYHIP_OK;
}

int pyhip_problem_dof_offsets(const pyhip_problem *p, int64_t *offsets) {
    PYHIP_REQUIRE(p && offsets, "null argument");
    for (int s = 0; s <= p->nsub; s++) offsets[s] = p->dof_off_h[s];
    return PYHIP_OK;
}

int pyhip_problem_set_params(pyhip_problem *p, int sub, const double *params, int nparams) {
    PYHIP_REQUIRE(p && sub >= 0 && sub < p->nsub, "bad sub-problem index");
    PYHIP_REQUIRE(nparams >= 0 && nparams <= kParamStride && (params || nparams == 0),
                  "bad parameters");
    for (int k = 0; k < nparams; k++) p->params_h[(size_t)sub * kParamStride + k] = params[k];
    PYHIP_CHECK_HIP(hipMemcpyAsync(p->params + (size_t)sub * kParamStride,
                                   p->params_h.data() + (size_t)sub * kParamStride,
                                   kParamStride * sizeof(double), hipMemcpyHostToDevice, p->st));
    PYHIP_CHECK_HIP(hipStreamSynchronize(p->st));
    return PYHIP_OK;
}

int pyhip_problem_set_stream(pyhip_problem *p, void *stream) {
    PYHIP_REQUIRE(p, "null problem");
    // the work queued so far finishes on the old stream before the new one is used
    PYHIP_CHECK_HIP(hipStreamSynchronize(p->st));
    p->st = (hipStream_t)stream;
    return PYHIP_OK;
}

int pyhip_problem_set_solved_residual(pyhip_problem *p, const char *name, int *ngroups) {
    PYHIP_REQUIRE(p && name, "null argument");
    const std::string n(name);
    int found = 0;
    for (auto &G : p->groups) {
        if (n.empty()) {
            G.active = G.form;
        } else if (!face_group_form(G.form) && G.form >= PYHIP_GENERATED_FORM_BASE) {
            G.active = gen_form_table(G.form, n);
        } else {
            G.active = -1;   // the compiled forms and the faces have the default residual only
        }
        if (G.active >= 0 && !face_group_form(G.form)) found++;
    }
    p->solved_residual = n;
    if (ngroups) *ngroups = found;
    return PYHIP_OK;
}

int pyhip_problem_set_aux(pyhip_problem *p, int sub, const double *aux, int64_t nnode) {
    PYHIP_REQUIRE(p && sub >= 0 && sub < p->nsub && aux, "bad arguments");
    const SubInfo &S = p->subs[sub];
    PYHIP_REQUIRE(nnode == S.nnode, "aux needs one value per node of the sub-problem's mesh");
    PYHIP_CHECK_HIP(hipMemcpyAsync(p->aux + S.node_off, aux, (size_t)nnode * sizeof(double), hipMemcpyHostToDevice,
                                   p->st));
    PYHIP_CHECK_HIP(hipStreamSynchronize(p->st));
    return PYHIP_OK;
}

int pyhip_pattern_copy(const pyhip_problem *p, int32_t *row_start, int32_t *col_index) {
    PYHIP_REQUIRE(p && row_start && col_index, "null argument");
    memcpy(row_start, p->row_start_h.data(), p->row_start_h.size() * sizeof(int32_t));
    memcpy(col_index, p->col_h.data(), p->col_h.size() * sizeof(int32_t));
    return PYHIP_OK;
}

}  // extern "C"

namespace pyhip {

static int check_assemble_args(pyhip_problem *p, const double *dofs, int flag, int deriv_param, int precision,
                               double *residual, void *values) {
    PYHIP_REQUIRE(p && dofs && residual, "null argument");
    PYHIP_REQUIRE(flag == 0 || flag == 1 || flag == 2, "flag must be 0, 1 or 2");
    PYHIP_REQUIRE(precision == PYHIP_F64 || precision == PYHIP_F32, "bad precision");
    PYHIP_REQUIRE(flag == 0 || values, "values buffer required for the Jacobian / mass matrix");
    // dM/dp: defined by the generated Taylor-Hood forms (their mass coefficient may hold the
    // parameter); the other forms' mass matrices do not depend on it
    if (flag == 2 && deriv_param >= 0)
        for (auto &G : p->groups)
            PYHIP_REQUIRE(face_group_form(G.form) || (G.form >= PYHIP_GENERATED_FORM_BASE && th_layout(G.form)),
                          "no mass matrix of a parameter derivative (defined for generated Taylor-Hood forms only)");
    PYHIP_REQUIRE(deriv_param == -1 || deriv_param == 0, "unknown global parameter");
    for (auto &G : p->groups)
        PYHIP_REQUIRE(deriv_param < 0 || G.form == PYHIP_FORM_BRATU || face_group_form(G.form) ||
                          gen_form_has_pderiv(G.form),
                      "parameter derivative only defined for forms with a global parameter");
    return PYHIP_OK;
}

// Element ranges [e0, e1) of group g to compute: all of it, or (sub_on_h, host [nsub]) the
// elements of the sub-problems with sub_on_h[s] != 0, adjacent ones merged
static std::vector<std::pair<int64_t, int64_t>> element_runs(const pyhip_problem *p, int g,
                                                             const int32_t *sub_on_h) {
    std::vector<std::pair<int64_t, int64_t>> runs;
    if (!sub_on_h) {
        runs.push_back({0, p->groups[g].nel});
        return runs;
    }
    for (int s = 0; s < p->nsub; s++) {
        if (!sub_on_h[s]) continue;
        const SubInfo &S = p->subs[s];
        int64_t b = -1, n = 0;
        if (S.group == g) b = S.el_off, n = S.nel;
        else if (S.fgroup == g) b = S.fel_off, n = S.nface;
        else if (S.egroup == g) b = S.eel_off, n = S.nenf;
        if (b < 0 || n == 0) continue;
        if (!runs.empty() && runs.back().second == b) runs.back().second = b + n;
        else runs.push_back({b, b + n});
    }
    return runs;
}

// Element kernels into the element buffers Er / Ej (FP64 or FP32 by precision; Ej may be
// null when flag == 0), then the deterministic gather into residual / values.
static int assemble_into(pyhip_problem *p, const double *dofs, int flag, int deriv_param, int precision,
                         double *residual, void *values, void *Er, void *Ej, const int32_t *sub_on_h = nullptr) {
    int rc;
    for (size_t g = 0; g < p->groups.size(); g++) {
        const Group &G = p->groups[g];
        if (G.active < 0) {
            // no table of the solved residual on this group: its elements contribute nothing
            const size_t es = precision == PYHIP_F64 ? sizeof(double) : sizeof(float);
            for (const auto &run : element_runs(p, (int)g, sub_on_h)) {
                const size_t n = (size_t)(run.second - run.first);
                PYHIP_CHECK_HIP(hipMemsetAsync((char *)Er + (G.er_off + run.first * G.ns) * es, 0, n * G.ns * es, p->st));
                if (flag)
                    PYHIP_CHECK_HIP(hipMemsetAsync((char *)Ej + (G.ej_off + run.first * G.ns * G.ns) * es, 0,
                                                   n * G.ns * G.ns * es, p->st));
            }
            continue;
        }
        const int form = G.active;
        for (const auto &run : element_runs(p, (int)g, sub_on_h)) {
            if (precision == PYHIP_F64) {
                ElemArgs<double> a{};
                a.nel = (int)G.nel;
                a.e0 = (int)run.first;
                a.e1 = (int)run.second;
                a.conn = G.conn; a.seqn = G.seqn; a.sub = G.sub;
                a.coords = p->coords; a.dofs = dofs; a.pinned = p->pinned; a.aux = p->aux;
                a.params = p->params; a.pstride = kParamStride; a.nvalue = 3;
                a.tab = !face_group_form(G.form) ? p->tab64[G.kind]
                                                 : (G.kind == kKindQuad9 ? p->tab64[ELEM_QUAD_Q2] : nullptr);
                a.Er = (double *)Er + G.er_off;
                a.Ej = flag ? (double *)Ej + G.ej_off : nullptr;
                a.flag = flag; a.deriv = deriv_param >= 0;
                a.transient = p->transient ? 1 : 0;
                a.tw0 = p->tw[0]; a.tw1 = p->tw[1]; a.tw2 = p->tw[2];
                a.hist1 = p->hist1; a.hist2 = p->hist2;
                a.hvec = p->hvec;
                a.htr = p->htr;
                a.fcoef = G.coef;
                TimeScope ts(&p->timer, "element", p->st);
                if (form == kFormFlux) {
                    if ((rc = launch_faces_f64(G.kind, a, p->st))) return rc;
                } else if (form >= PYHIP_GENERATED_FORM_BASE) {
                    if ((rc = launch_generated_f64(form, a, p->st))) return rc;
                } else if ((rc = launch_elements_f64(G.kind, form, a, p->st)))
                    return rc;
            } else {
                ElemArgs<float> a{};
                a.nel = (int)G.nel;
                a.e0 = (int)run.first;
                a.e1 = (int)run.second;
                a.conn = G.conn; a.seqn = G.seqn; a.sub = G.sub;
                a.coords = p->coords; a.dofs = dofs; a.pinned = p->pinned; a.aux = p->aux;
                a.params = p->params; a.pstride = kParamStride; a.nvalue = 3;
                a.tab = !face_group_form(G.form) ? p->tab32[G.kind]
                                                 : (G.kind == kKindQuad9 ? p->tab32[ELEM_QUAD_Q2] : nullptr);
                a.Er = (float *)Er + G.er_off;
                a.Ej = flag ? (float *)Ej + G.ej_off : nullptr;
                a.flag = flag; a.deriv = deriv_param >= 0;
                a.transient = p->transient ? 1 : 0;
                a.tw0 = p->tw[0]; a.tw1 = p->tw[1]; a.tw2 = p->tw[2];
                a.hist1 = p->hist1; a.hist2 = p->hist2;
                a.hvec = p->hvec;
                a.htr = p->htr;
                a.fcoef = G.coef;
                TimeScope ts(&p->timer, "element", p->st);
                if (form == kFormFlux) {
                    if ((rc = launch_faces_f32(G.kind, a, p->st))) return rc;
                } else if (form >= PYHIP_GENERATED_FORM_BASE) {
                    if ((rc = launch_generated_f32(form, a, p->st))) return rc;
                } else if ((rc = launch_elements_f32(G.kind, form, a, p->st)))
                    return rc;
            }
        }
    }
    const GatherMap m = p->gmap();
    if (precision == PYHIP_F64) {
        if (flag) {
            TimeScope ts(&p->timer, "gather_values", p->st);
            if ((rc = launch_gather_values_f64(m, (const double *)Ej, (const double *)Er, (double *)values,
                                               residual, p->st)))
                return rc;
        } else {
            TimeScope ts(&p->timer, "gather_residual", p->st);
            if ((rc = launch_gather_residual_f64(p->ndof, p->occ_ptr, p->occ_roff, (const double *)Er, residual,
                                                 p->st)))
                return rc;
        }
    } else {
        if (flag) {
            TimeScope ts(&p->timer, "gather_values", p->st);
            if ((rc = launch_gather_values_f32(m, (const float *)Ej, (const float *)Er, (float *)values, residual,
                                               p->st)))
                return rc;
        } else {
            TimeScope ts(&p->timer, "gather_residual", p->st);
            if ((rc = launch_gather_residual_f32(p->ndof, p->occ_ptr, p->occ_roff, (const float *)Er, residual,
                                                 p->st)))
                return rc;
        }
    }
    return PYHIP_OK;
}

// ---- the allocation-free assembly ------------------------------------------------------
// Element buffers of pyhip_assemble_noalloc: one array per device, alive until the process
// ends. Host code never uses the array's name as an address (on the host it is only a
// stand-in for the symbol); the device address comes from hipGetSymbolAddress, once per
// device, and is kept below.
static __device__ __align__(256) unsigned char scratch[PYHIP_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void *g_scratch_ptr[kMaxDevices] = {};

static int scratch_address(void **out) {
    int dev = -1;
    PYHIP_CHECK_HIP(hipGetDevice(&dev));
    PYHIP_REQUIRE(dev >= 0 && dev < kMaxDevices, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        PYHIP_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        PYHIP_REQUIRE(ptr != nullptr, "no device address for the scratch array");
        g_scratch_ptr[dev] = ptr;
    }
    *out = g_scratch_ptr[dev];
    return PYHIP_OK;
}

}  // namespace pyhip

extern "C" {

int pyhip_assemble(pyhip_problem *p, const double *dofs, int flag, int deriv_param,
                   int precision, double *residual, void *values) {
    return pyhip::assemble_subset(p, dofs, flag, deriv_param, precision, residual, values, nullptr);
}

}  // extern "C"

namespace pyhip {
// pyhip_assemble over the sub-problems with sub_on_h[s] != 0 (host array [nsub], or null for
// all): the element kernels run over those sub-problems' elements only; the residual and
// CSR rows of the others are then unspecified (gathered from element rows left by earlier
// calls).
int assemble_subset(pyhip_problem *p, const double *dofs, int flag, int deriv_param, int precision,
                    double *residual, void *values, const int32_t *sub_on_h) {
    int rc = check_assemble_args(p, dofs, flag, deriv_param, precision, residual, values);
    if (rc) return rc;
    void *Er, *Ej;
    if (precision == PYHIP_F64) {
        if (!p->Er64) {
            if ((rc = dev_alloc(&p->Er64, p->n_er, p->device_bytes))) return rc;
        }
        if (flag && !p->Ej64) {
            if ((rc = dev_alloc(&p->Ej64, p->n_ej, p->device_bytes))) return rc;
        }
        Er = p->Er64;
        Ej = p->Ej64;
    } else {
        if (!p->Er32) {
            if ((rc = dev_alloc(&p->Er32, p->n_er, p->device_bytes))) return rc;
        }
        if (flag && !p->Ej32) {
            if ((rc = dev_alloc(&p->Ej32, p->n_ej, p->device_bytes))) return rc;
        }
        Er = p->Er32;
        Ej = p->Ej32;
    }
    if ((rc = assemble_into(p, dofs, flag, deriv_param, precision, residual, values, Er, Ej, sub_on_h))) return rc;
    // the export reads the element matrices of the last complete FP64 matrix assembly
    if (precision == PYHIP_F64 && flag) p->have_ej64 = !sub_on_h;
    return PYHIP_OK;
}
}  // namespace pyhip

extern "C" {

int pyhip_assemble_noalloc(pyhip_problem *p, const double *dofs, int flag, int deriv_param, int precision,
                           double *residual, void *values) {
    int rc = check_assemble_args(p, dofs, flag, deriv_param, precision, residual, values);
    if (rc) return rc;
    const size_t esz = precision == PYHIP_F64 ? sizeof(double) : sizeof(float);
    const size_t er_bytes = ((size_t)p->n_er * esz + 255) / 256 * 256;
    const size_t need = er_bytes + (flag ? (size_t)p->n_ej * esz : 0);
    if (need > (size_t)PYHIP_NOALLOC_SCRATCH_BYTES) {
        set_error("pyhip_assemble_noalloc: the element buffers of this problem need " + std::to_string(need) +
                  " bytes, more than the " + std::to_string((long long)PYHIP_NOALLOC_SCRATCH_BYTES) +
                  "-byte static scratch; use pyhip_assemble");
        return PYHIP_ERR_INVALID;
    }
    void *base = nullptr;
    if ((rc = scratch_address(&base))) return rc;
    void *Er = base;
    void *Ej = flag ? (void *)((unsigned char *)base + er_bytes) : nullptr;
    if ((rc = assemble_into(p, dofs, flag, deriv_param, precision, residual, values, Er, Ej))) return rc;
    // the last FP64 element matrices are not in the problem's own buffer
    if (precision == PYHIP_F64 && flag) p->have_ej64 = false;
    return PYHIP_OK;
}

int pyhip_assemble_hessian_vector(pyhip_problem *p, const double *dofs, const double *y, double *residual,
                                  double *values) {
    PYHIP_REQUIRE(p && dofs && y && residual && values, "null argument");
    p->hvec = y;
    const int rc = pyhip_assemble(p, dofs, 1, -1, PYHIP_F64, residual, values);
    p->hvec = nullptr;
    return rc;
}

int pyhip_set_time_stepping(pyhip_problem *p, int enabled, const double *weights, const double *hist1,
                            const double *hist2) {
    PYHIP_REQUIRE(p, "null argument");
    if (!enabled) {
        p->transient = false;
        p->tw[0] = p->tw[1] = p->tw[2] = 0.0;
        p->hist1 = p->hist2 = nullptr;
        return PYHIP_OK;
    }
    PYHIP_REQUIRE(weights && hist1 && hist2, "time stepping needs weights and two history vectors");
    p->transient = true;
    for (int t = 0; t < 3; t++) p->tw[t] = weights[t];
    p->hist1 = hist1;
    p->hist2 = hist2;
    return PYHIP_OK;
}

int pyhip_export_reference_csr(pyhip_problem *p, const double *values, int32_t *rs_out,
                               int32_t *col_out, double *val_out, int64_t *nnz_out) {
    PYHIP_REQUIRE(p && values && rs_out && col_out && val_out && nnz_out, "null argument");
    PYHIP_REQUIRE(p->have_ej64 && p->Ej64, "no complete FP64 matrix assembled yet");
    int rc;
    if (!p->keys) {
        if ((rc = dev_alloc(&p->keys, p->nnz, p->device_bytes))) return rc;
        if ((rc = dev_alloc(&p->kept, export_kept_ints(p->ndof), p->device_bytes))) return rc;
    }
    return export_reference_csr(p->gmap(), p->col, p->Ej64, values, p->keys, p->kept, rs_out,
                                col_out, val_out, nnz_out, p->st);
}

int pyhip_timing_enable(pyhip_problem *p, int enable) {
    PYHIP_REQUIRE(p, "null problem");
    PYHIP_CHECK_HIP(hipStreamSynchronize(p->st));
    p->timer.flush();
    p->timer.total.clear();
    p->timer.work.clear();
    p->timer.on = enable != 0;
    return PYHIP_OK;
}

int pyhip_timing_read(pyhip_problem *p, const char *family, int64_t *launches, double *total_ms) {
    PYHIP_REQUIRE(p && family && launches && total_ms, "null argument");
    PYHIP_CHECK_HIP(hipStreamSynchronize(p->st));
    p->timer.flush();
    auto it = p->timer.total.find(family);
    *launches = it == p->timer.total.end() ? 0 : it->second.first;
    *total_ms = it == p->timer.total.end() ? 0.0 : it->second.second;
    return PYHIP_OK;
}

int pyhip_timing_bytes(pyhip_problem *p, const char *family, int64_t *bytes) {
    PYHIP_REQUIRE(p && family && bytes, "null argument");
    auto it = p->timer.work.find(family);
    *bytes = it == p->timer.work.end() ? 0 : it->second;
    return PYHIP_OK;
}

}  // extern "C"

// Host only (tests on the CPU): the multigrid element-level rows of one sub-problem
// descriptor (Taylor-Hood: the block preconditioner's, velocity components): nc coarse unknowns, rs [ndof+1], and (with col / w non-null) the entries,
// nnz_out of them. No device work.
extern "C" int pyhip_element_level_rows(const pyhip_subproblem_desc *d, int64_t *nc, int32_t *rs, int32_t *col,
                                        double *w, int64_t *nnz_out) {
    PYHIP_REQUIRE(d && nc && nnz_out, "null argument");
    PYHIP_REQUIRE(elem_nnode(d->elem_kind) > 0 && d->nnode > 0 && d->nelem > 0 && d->conn && d->eqn && d->ndof >= 0,
                  "invalid descriptor");
    const int nn = elem_nnode(d->elem_kind);
    for (int64_t k = 0; k < d->nelem * nn; k++)
        PYHIP_REQUIRE(d->conn[k] >= 0 && d->conn[k] < d->nnode, "connectivity out of range");
    for (int64_t k = 0; k < d->nnode * d->nvalue; k++)
        PYHIP_REQUIRE(d->eqn[k] >= -1 && d->eqn[k] < d->ndof, "equation number out of range");
    AmgFixedInterp F;
    element_level_rows(*d, F, true);
    F.rs.push_back((int32_t)F.col.size());
    *nc = F.on[0] ? F.nc[0] : 0;
    *nnz_out = (int64_t)F.col.size();
    if (rs) std::memcpy(rs, F.rs.data(), F.rs.size() * sizeof(int32_t));
    if (col && !F.col.empty()) std::memcpy(col, F.col.data(), F.col.size() * sizeof(int32_t));
    if (w && !F.w.empty()) std::memcpy(w, F.w.data(), F.w.size() * sizeof(double));
    return PYHIP_OK;
}
