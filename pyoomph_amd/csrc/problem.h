// Internal definition of pyhip_problem (device-resident batched problem).
#pragma once
#include <map>
#include <memory>
#include <string>
#include <vector>

#include "../../include/pyoomph_codegen.h"
#include "amg_plan.h"
#include "common.h"
#include "element_kernels.h"
#include "pattern.h"

namespace pyhip {

constexpr int kParamStride = 64;  // doubles of parameters per sub-problem

// Boundary-face (interface) elements: internal form and element kinds of their groups
// (each face carries its own flux coefficients, Group::coef).
constexpr int kFormFlux = 16;          // flux = alpha*u + beta on the faces
constexpr int kKindLine2 = 16;         // 2-node faces of P1 triangles (TGauss<1,2>)
constexpr int kKindLine3 = 17;         // 3-node faces of P2 triangles (TGauss<1,3>)
constexpr int kKindLine3Q = 18;        // 3-node faces of Q2 quadrilaterals (Gauss<1,3>)
constexpr int kKindQuad9 = 19;         // 9-node faces of Q2 bricks (Gauss<2,3>)
constexpr int kKindLine3TH = 20;       // traction faces of Taylor-Hood Q2 quads (2 slots per node)
// faces of EnforcedBC constraints: slots (u, multiplier) per face node
constexpr int kKindLine2E = 21;        // on P1 triangles
constexpr int kKindLine3E = 22;        // on P2 triangles
constexpr int kKindLine3QE = 23;       // on Q2 quadrilaterals
inline bool enforced_kind(int face_kind) { return face_kind >= kKindLine2E && face_kind <= kKindLine3QE; }
// flux coefficients per face: alpha, beta (scalar forms) or alpha, beta_x, beta_y (Taylor-Hood)
inline int face_ncoef(int face_kind) { return face_kind == kKindLine3TH ? 3 : 2; }

struct SubInfo {
    int form, kind, nvalue;
    int64_t dof_off, ndof, node_off, nnode;
    int group;               // element group holding its elements
    int64_t el_off;          // first element inside that group
    int64_t nel;
    int fgroup = -1;         // face group (-1: no faces)
    int64_t fel_off = 0, nface = 0;
    int egroup = -1;         // group of the EnforcedBC faces (-1: none)
    int64_t eel_off = 0, nenf = 0;
};

struct Group {
    int kind, form, ns, nn, dim;
    // the form whose kernels the group runs for the residual being solved (the reference's
    // current_res_jac of the element code, src/problem.cpp:186-200): form itself for the
    // default residual "", a generated form's named residual table, or -1 when the group
    // has no such table (its elements then contribute nothing)
    int active = 0;
    int64_t nel = 0;
    int64_t er_off = 0, ej_off = 0;   // offsets into the element vector / matrix buffers
    int32_t *conn = nullptr;          // [nn][nel]
    int32_t *seqn = nullptr;          // [ns][nel]
    int32_t *sub = nullptr;           // [nel]
    double *coef = nullptr;           // face groups: [nel][face_ncoef] flux coefficients of each face
};

struct KrylovWork;   // krylov.hip


// HIP-event timing of kernel families on the problem's stream (measurement aid).
struct KernelTimer {
    bool on = false;
    std::vector<hipEvent_t> pool;
    size_t used = 0;
    std::map<std::string, std::vector<std::pair<hipEvent_t, hipEvent_t>>> pending;
    std::map<std::string, std::pair<int64_t, double>> total;
    std::map<std::string, int64_t> work;   // algorithmic bytes moved by the family's kernels
    // While deferred, launches add only their per-launch part; the caller adds the bytes
    // that depend on the rows visited once it knows them exactly (per-sub-problem counts).
    bool defer = false;
    void add_work(const char *fam, int64_t bytes, int64_t per_launch = 0) {
        if (on) work[fam] += defer ? per_launch : bytes;
    }
    hipEvent_t take() {
        if (used == pool.size()) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return nullptr;
            pool.push_back(e);
        }
        return pool[used++];
    }
    void flush() {
        for (auto &kv : pending) {
            auto &t = total[kv.first];
            for (auto &pr : kv.second) {
                float ms = 0.f;
                if (hipEventSynchronize(pr.second) == hipSuccess &&
                    hipEventElapsedTime(&ms, pr.first, pr.second) == hipSuccess) {
                    t.first++;
                    t.second += ms;
                }
            }
        }
        pending.clear();
        used = 0;
    }
    ~KernelTimer() {
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// Records start/end events around the launches made while it is alive.
struct TimeScope {
    KernelTimer *t;
    const char *fam;
    hipStream_t st;
    hipEvent_t a = nullptr;
    TimeScope(KernelTimer *timer, const char *family, hipStream_t s) : t(timer), fam(family), st(s) {
        if (t && t->on) {
            if (t->used + 2 > 4096) t->flush();
            a = t->take();
            if (a) (void)hipEventRecord(a, st);
        }
    }
    ~TimeScope() {
        if (a) {
            hipEvent_t b = t->take();
            if (b) {
                (void)hipEventRecord(b, st);
                t->pending[fam].push_back({a, b});
            }
        }
    }
};

// pyhip_assemble restricted to the sub-problems with sub_on_h[s] != 0 (host [nsub]; null:
// all of them); api.hip
int assemble_subset(pyhip_problem *p, const double *dofs, int flag, int deriv_param, int precision,
                    double *residual, void *values, const int32_t *sub_on_h);
int launch_elements_f64(int kind, int form, const ElemArgs<double> &a, hipStream_t s);
int launch_elements_f32(int kind, int form, const ElemArgs<float> &a, hipStream_t s);
// run-time generated forms (codegen_rt.hip)
bool gen_form_known(int form);
int gen_form_kind(int form);
bool gen_form_has_pderiv(int form);
bool gen_form_fields(int form, std::string &field, std::string &aux_field, int &global_param);
int gen_form_layout(int form);
// the Taylor-Hood value layout (22 slots: ux, uy at the nodes, p at the corners)
inline bool th_layout(int form) {
    return form == PYHIP_FORM_NAVIER_STOKES_TH || (form >= PYHIP_GENERATED_FORM_BASE && gen_form_layout(form) == 1);
}
// per-face coefficients of a generated boundary-face or constraint form (layouts 2, 3), -1
// for other forms
int gen_form_ncoef(int form);
// the generated form has Hessian-vector product kernels
bool gen_form_has_hv(int form);
// id of the residual table `name` of a generated form (-1: none)
int gen_form_table(int form, const std::string &name);
inline bool gen_face_form(int form) { return form >= PYHIP_GENERATED_FORM_BASE && gen_form_layout(form) == 2; }
// a generated EnforcedBC constraint form (layout 3): its faces carry (u, multiplier) slots
inline bool gen_constraint_form(int form) { return form >= PYHIP_GENERATED_FORM_BASE && gen_form_layout(form) == 3; }
// groups of boundary-face elements: the built-in flux / traction / EnforcedBC faces and the
// generated boundary-face forms
inline bool face_group_form(int form) { return form == kFormFlux || gen_face_form(form) || gen_constraint_form(form); }
inline int group_ncoef(int form, int kind) { return form == kFormFlux ? face_ncoef(kind) : gen_form_ncoef(form); }
int launch_generated_f64(int form, const ElemArgs<double> &a, hipStream_t s);
int launch_generated_f32(int form, const ElemArgs<float> &a, hipStream_t s);
int launch_faces_f64(int kind, const ElemArgs<double> &a, hipStream_t s);
int launch_faces_f32(int kind, const ElemArgs<float> &a, hipStream_t s);
int launch_gather_values_f64(const GatherMap &m, const double *Ej, const double *Er, double *vals,
                             double *res, hipStream_t st);
int launch_gather_values_f32(const GatherMap &m, const float *Ej, const float *Er, float *vals,
                             double *res, hipStream_t st);
int launch_gather_residual_f64(int64_t nrow, const int64_t *occ_ptr, const int64_t *occ_roff,
                               const double *Er, double *res, hipStream_t st);
int launch_gather_residual_f32(int64_t nrow, const int64_t *occ_ptr, const int64_t *occ_roff,
                               const float *Er, double *res, hipStream_t st);
// rows in first-touch column order, or (sorted) in ascending column order; kept holds
// export_kept_ints(nrow) ints
int export_reference_csr(const GatherMap &m, const int32_t *col, const double *Ej,
                         const double *vals, uint32_t *keys, int32_t *kept, int32_t *out_rs,
                         int32_t *out_col, double *out_val, int64_t *nnz_out, hipStream_t st,
                         bool sorted = false);
int64_t export_kept_ints(int64_t nrow);


}  // namespace pyhip

struct pyhip_problem {
    hipStream_t st = nullptr;
    int nsub = 0;
    std::vector<pyhip::SubInfo> subs;
    std::vector<pyhip::Group> groups;
    std::string solved_residual;             // name of the residual the groups run ("": default)
    std::vector<int64_t> dof_off_h;          // [nsub+1]
    int64_t ndof = 0, nnode = 0, nnz = 0, nocc = 0, ncontrib = 0, n_er = 0, n_ej = 0;
    int64_t device_bytes = 0;
    double *coords = nullptr;   // [nnode][3]
    double *pinned = nullptr;   // [nnode][3]
    double *aux = nullptr;      // [nnode]
    double *params = nullptr;   // [nsub][kParamStride]
    std::vector<double> params_h;
    pyhip::DevTables<double> *tab64[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    pyhip::DevTables<float> *tab32[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    // pattern / map
    int32_t *row_start = nullptr, *col = nullptr, *diag = nullptr;
    int64_t *occ_ptr = nullptr, *occ_roff = nullptr;
    int32_t *cptr = nullptr;
    uint32_t *cidx = nullptr;
    int32_t *dof_sub = nullptr;             // [ndof] sub-problem of each row
    int64_t *dof_off = nullptr;             // [nsub+1]
    std::vector<int32_t> row_start_h, col_h;
    std::vector<int8_t> dof_val_h;           // [ndof] nodal value index of each unknown (Taylor-Hood: 0 ux, 1 uy, 2 p)
    std::vector<double> pmass_h;             // [ndof] lumped Q1 pressure mass of Taylor-Hood pressure unknowns
    pyhip::AmgFixedInterp amg_fix;           // the multigrid's element level (quadratic elements of scalar forms)
    pyhip::AmgFixedInterp amg_fix_vel;       // ... and with the Taylor-Hood velocity components (block preconditioner)
    // element buffers (lazy)
    double *Er64 = nullptr, *Ej64 = nullptr;
    float *Er32 = nullptr, *Ej32 = nullptr;
    bool have_ej64 = false;                  // Ej64 holds the element matrices of the last complete FP64 matrix assembly
    // export scratch (lazy)
    uint32_t *keys = nullptr;
    int32_t *kept = nullptr;
    // multi-assembly scratch (lazy): the structural values and the residual of a pass whose
    // output was not requested
    double *ma_vals = nullptr, *ma_res = nullptr;
    // Newton / Krylov workspace (lazy)
    pyhip::KrylovWork *kw = nullptr;
    // ILU(0) ordering: rows listed here go after all other rows of their sub-problem
    // (set for matrices that arrive without element information)
    std::vector<char> ilu_last_h;
    pyhip::KernelTimer timer;
    // time stepping of the scalar forms (pyhip_set_time_stepping)
    bool transient = false;
    double tw[3] = {0.0, 0.0, 0.0};
    const double *hist1 = nullptr, *hist2 = nullptr;
    const double *hvec = nullptr;            // set while a Hessian-vector product assembles
    int htr = 0;                             // ... and 1 for the transposed product
    pyhip::GatherMap gmap() const {
        return pyhip::GatherMap{ndof, nnz, row_start, occ_ptr, occ_roff, cptr, cidx};
    }
};
