// Run-time compiled element kernels (include/pyoomph_codegen.h): hiprtc compilation,
// module loading, the registry of generated forms and their launches.
#include <hip/hiprtc.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/pyoomph_codegen.h"
#include "problem.h"

namespace pyhip {

namespace {

struct GenForm {
    std::string name;
    int kind = 0;
    bool symmetric = false, has_pderiv = false;
    std::string field = "u", aux_field;   // nodal field names (JIT table)
    int global_param = 0;                 // the parameter dR/dp is taken in
    int layout = 0;                       // 0 one value per node, 1 Taylor-Hood (22 slots), 2 boundary
                                          // faces, 3 EnforcedBC constraint faces
    int ncoef = 0;                        // layouts 2, 3: coefficients per face
    std::vector<std::pair<std::string, int>> tables;   // named residual tables: (name, form id)
    hipModule_t mod = nullptr;
    hipFunction_t f64 = nullptr, f64_tr = nullptr, f32 = nullptr, f32_tr = nullptr;
    // Hessian-vector products (FP64; scalar bulk forms): sum_j y_j d2R_i/du_j du_k and the
    // transposed product, or null where the source does not define them
    hipFunction_t f64_hv = nullptr, f64_hvt = nullptr;
};

std::mutex g_mu;
std::vector<GenForm> g_forms;

// elements per 256-thread workgroup of the scalar element kernels (element_common.h
// scalar_epb): 256 / max(nodes, points per phase-A chunk)
int gen_epb(int kind) {
    switch (kind) {
    case ELEM_TRI_P1: return 256 / 3;
    case ELEM_TRI_P2: return 256 / 7;
    case ELEM_QUAD_Q2: return 256 / 9;
    case ELEM_HEX_Q2: return 256 / 27;
    }
    return 0;
}

#define PYHIP_CHECK_RTC(expr)                                                         \
    do {                                                                              \
        hiprtcResult _r = (expr);                                                     \
        if (_r != HIPRTC_SUCCESS) {                                                   \
            set_error(std::string(#expr) + ": " + hiprtcGetErrorString(_r));          \
            return PYHIP_ERR_INVALID;                                                 \
        }                                                                             \
    } while (0)

int compile(const char *source, const char *name, int nh, const char *const *hs, const char *const *hn,
            std::vector<char> &code) {
    hiprtcProgram prog;
    PYHIP_CHECK_RTC(hiprtcCreateProgram(&prog, source, (std::string(name) + ".hip").c_str(), nh, hs, hn));
    // FP64 results must follow the written operation order (no contraction), and packed
    // FP32 instructions are not generated
    const char *opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-ffp-contract=off",
                          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    const hiprtcResult r = hiprtcCompileProgram(prog, (int)(sizeof(opts) / sizeof(opts[0])), opts);
    if (r != HIPRTC_SUCCESS) {
        size_t ls = 0;
        std::string log;
        if (hiprtcGetProgramLogSize(prog, &ls) == HIPRTC_SUCCESS && ls > 1) {
            log.resize(ls);
            (void)hiprtcGetProgramLog(prog, &log[0]);
        }
        (void)hiprtcDestroyProgram(&prog);
        set_error(std::string("hiprtc: ") + hiprtcGetErrorString(r) + "\n" + log);
        return PYHIP_ERR_INVALID;
    }
    size_t cs = 0;
    PYHIP_CHECK_RTC(hiprtcGetCodeSize(prog, &cs));
    code.resize(cs);
    PYHIP_CHECK_RTC(hiprtcGetCode(prog, code.data()));
    (void)hiprtcDestroyProgram(&prog);
    return PYHIP_OK;
}

}  // namespace

bool gen_form_known(int form) {
    std::lock_guard<std::mutex> lock(g_mu);
    const int k = form - PYHIP_GENERATED_FORM_BASE;
    return k >= 0 && k < (int)g_forms.size();
}

int gen_form_kind(int form) {
    std::lock_guard<std::mutex> lock(g_mu);
    return g_forms[form - PYHIP_GENERATED_FORM_BASE].kind;
}

bool gen_form_has_pderiv(int form) {
    std::lock_guard<std::mutex> lock(g_mu);
    const int k = form - PYHIP_GENERATED_FORM_BASE;
    return k >= 0 && k < (int)g_forms.size() && g_forms[k].has_pderiv;
}

int gen_form_layout(int form) {
    std::lock_guard<std::mutex> lock(g_mu);
    const int k = form - PYHIP_GENERATED_FORM_BASE;
    return (k >= 0 && k < (int)g_forms.size()) ? g_forms[k].layout : 0;
}

bool gen_form_has_hv(int form) {
    std::lock_guard<std::mutex> lock(g_mu);
    const int k = form - PYHIP_GENERATED_FORM_BASE;
    return k >= 0 && k < (int)g_forms.size() && g_forms[k].f64_hv && g_forms[k].f64_hvt;
}

int gen_form_table(int form, const std::string &name) {
    std::lock_guard<std::mutex> lock(g_mu);
    const int k = form - PYHIP_GENERATED_FORM_BASE;
    if (k < 0 || k >= (int)g_forms.size()) return -1;
    for (const auto &t : g_forms[k].tables)
        if (t.first == name) return t.second;
    return -1;
}

int gen_form_ncoef(int form) {
    std::lock_guard<std::mutex> lock(g_mu);
    const int k = form - PYHIP_GENERATED_FORM_BASE;
    return (k >= 0 && k < (int)g_forms.size() && (g_forms[k].layout == 2 || g_forms[k].layout == 3)) ? g_forms[k].ncoef
                                                                                                      : -1;
}

bool gen_form_fields(int form, std::string &field, std::string &aux_field, int &global_param) {
    std::lock_guard<std::mutex> lock(g_mu);
    const int k = form - PYHIP_GENERATED_FORM_BASE;
    if (k < 0 || k >= (int)g_forms.size()) return false;
    field = g_forms[k].field;
    aux_field = g_forms[k].aux_field;
    global_param = g_forms[k].global_param;
    return true;
}

template <typename T>
static int launch_gen(int form, const ElemArgs<T> &a, hipStream_t st) {
    hipFunction_t f;
    int kind, layout;
    {
        std::lock_guard<std::mutex> lock(g_mu);
        const GenForm &G = g_forms[form - PYHIP_GENERATED_FORM_BASE];
        const bool f64 = std::is_same<T, double>::value;
        f = f64 ? (a.transient ? G.f64_tr : G.f64) : (a.transient ? G.f32_tr : G.f32);
        if (a.hvec) f = f64 ? (a.htr ? G.f64_hvt : G.f64_hv) : nullptr;
        kind = G.kind;
        layout = G.layout;
    }
    PYHIP_REQUIRE(f != nullptr, "Hessian-vector products need the form's _f64_hv / _f64_hvt kernels (FP64)");
    // elements per workgroup: one thread per slot (Taylor-Hood: 22) or per node / point;
    // boundary-face and constraint forms: one thread per face
    const int epb = (layout == 2 || layout == 3) ? 256 : (layout == 1 ? 256 / 22 : gen_epb(kind));
    const unsigned blocks = (unsigned)((a.e1 - a.e0 + epb - 1) / epb);
    if (a.e1 <= a.e0) return PYHIP_OK;
    ElemArgs<T> arg = a;
    void *params[] = {&arg};
    PYHIP_CHECK_HIP(hipModuleLaunchKernel(f, blocks, 1, 1, 256, 1, 1, 0, st, params, nullptr));
    return PYHIP_OK;
}

int launch_generated_f64(int form, const ElemArgs<double> &a, hipStream_t st) { return launch_gen(form, a, st); }
int launch_generated_f32(int form, const ElemArgs<float> &a, hipStream_t st) { return launch_gen(form, a, st); }

}  // namespace pyhip

using namespace pyhip;

extern "C" {

int pyhip_codegen_compile(const char *name, const char *source, int nheaders, const char *const *header_sources,
                          const char *const *header_names, int elem_kind, int symmetric, int has_param_derivative,
                          int load, int *form_id, size_t *code_size) {
    PYHIP_REQUIRE(name && source && nheaders >= 0 && (nheaders == 0 || (header_sources && header_names)),
                  "null argument");
    PYHIP_REQUIRE(gen_epb(elem_kind) > 0, "unknown element family");
    std::vector<char> code;
    int rc = compile(source, name, nheaders, header_sources, header_names, code);
    if (rc) return rc;
    if (code_size) *code_size = code.size();
    if (!load) return PYHIP_OK;
    PYHIP_REQUIRE(form_id, "null form id");
    GenForm G;
    G.name = name;
    G.kind = elem_kind;
    G.symmetric = symmetric != 0;
    G.has_pderiv = has_param_derivative != 0;
    PYHIP_CHECK_HIP(hipModuleLoadData(&G.mod, code.data()));
    const std::string n(name);
    PYHIP_CHECK_HIP(hipModuleGetFunction(&G.f64, G.mod, (n + "_f64").c_str()));
    PYHIP_CHECK_HIP(hipModuleGetFunction(&G.f64_tr, G.mod, (n + "_f64_tr").c_str()));
    PYHIP_CHECK_HIP(hipModuleGetFunction(&G.f32, G.mod, (n + "_f32").c_str()));
    PYHIP_CHECK_HIP(hipModuleGetFunction(&G.f32_tr, G.mod, (n + "_f32_tr").c_str()));
    // the Hessian-vector product kernels are optional: looked up only when the source defines
    // them (a failed lookup would leave an error for the next HIP call to report)
    const std::string src(source);
    if (src.find(" " + n + "_f64_hv(") != std::string::npos)
        PYHIP_CHECK_HIP(hipModuleGetFunction(&G.f64_hv, G.mod, (n + "_f64_hv").c_str()));
    if (src.find(" " + n + "_f64_hvt(") != std::string::npos)
        PYHIP_CHECK_HIP(hipModuleGetFunction(&G.f64_hvt, G.mod, (n + "_f64_hvt").c_str()));
    std::lock_guard<std::mutex> lock(g_mu);
    g_forms.push_back(G);
    *form_id = PYHIP_GENERATED_FORM_BASE + (int)g_forms.size() - 1;
    return PYHIP_OK;
}

int pyhip_codegen_set_fields(int form_id, const char *field, const char *aux_field, int global_param) {
    PYHIP_REQUIRE(field && field[0], "null field name");
    PYHIP_REQUIRE(global_param >= 0 && global_param < kParamStride, "parameter index out of range");
    std::lock_guard<std::mutex> lock(g_mu);
    const int k = form_id - PYHIP_GENERATED_FORM_BASE;
    PYHIP_REQUIRE(k >= 0 && k < (int)g_forms.size(), "unknown generated form");
    g_forms[k].field = field;
    g_forms[k].aux_field = aux_field ? aux_field : "";
    g_forms[k].global_param = global_param;
    return PYHIP_OK;
}

int pyhip_codegen_set_layout(int form_id, int layout) {
    PYHIP_REQUIRE(layout == 0 || layout == 1, "layout must be 0 (one value per node) or 1 (Taylor-Hood)");
    std::lock_guard<std::mutex> lock(g_mu);
    const int k = form_id - PYHIP_GENERATED_FORM_BASE;
    PYHIP_REQUIRE(k >= 0 && k < (int)g_forms.size(), "unknown generated form");
    PYHIP_REQUIRE(layout == 0 || g_forms[k].kind == ELEM_QUAD_Q2, "the Taylor-Hood layout is for Q2 quadrilaterals");
    g_forms[k].layout = layout;
    return PYHIP_OK;
}

int pyhip_codegen_set_face_form(int form_id, int ncoef) {
    PYHIP_REQUIRE(ncoef >= 0 && ncoef <= 16, "coefficients per face: 0..16");
    std::lock_guard<std::mutex> lock(g_mu);
    const int k = form_id - PYHIP_GENERATED_FORM_BASE;
    PYHIP_REQUIRE(k >= 0 && k < (int)g_forms.size(), "unknown generated form");
    const int kind = g_forms[k].kind;
    PYHIP_REQUIRE(kind == ELEM_TRI_P1 || kind == ELEM_TRI_P2 || kind == ELEM_QUAD_Q2,
                  "boundary-face forms are generated for the line faces of P1 / P2 triangles and Q2 quadrilaterals");
    g_forms[k].layout = 2;
    g_forms[k].ncoef = ncoef;
    return PYHIP_OK;
}

int pyhip_codegen_set_constraint_form(int form_id, int ncoef) {
    PYHIP_REQUIRE(ncoef >= 0 && ncoef <= 16, "coefficients per face: 0..16");
    std::lock_guard<std::mutex> lock(g_mu);
    const int k = form_id - PYHIP_GENERATED_FORM_BASE;
    PYHIP_REQUIRE(k >= 0 && k < (int)g_forms.size(), "unknown generated form");
    const int kind = g_forms[k].kind;
    PYHIP_REQUIRE(kind == ELEM_TRI_P1 || kind == ELEM_TRI_P2 || kind == ELEM_QUAD_Q2,
                  "constraint forms are generated for the line faces of P1 / P2 triangles and Q2 quadrilaterals");
    g_forms[k].layout = 3;
    g_forms[k].ncoef = ncoef;
    return PYHIP_OK;
}

int pyhip_codegen_add_residual_table(int form_id, const char *name, int table_form_id) {
    PYHIP_REQUIRE(name && name[0], "a residual table needs a name (\"\" is the form itself)");
    std::lock_guard<std::mutex> lock(g_mu);
    const int k = form_id - PYHIP_GENERATED_FORM_BASE, t = table_form_id - PYHIP_GENERATED_FORM_BASE;
    PYHIP_REQUIRE(k >= 0 && k < (int)g_forms.size() && t >= 0 && t < (int)g_forms.size(), "unknown generated form");
    PYHIP_REQUIRE(g_forms[k].kind == g_forms[t].kind && g_forms[k].layout == g_forms[t].layout && g_forms[k].layout != 2,
                  "a residual table is a bulk form of the same element family and value layout");
    for (auto &e : g_forms[k].tables)
        if (e.first == name) {
            e.second = table_form_id;
            return PYHIP_OK;
        }
    g_forms[k].tables.push_back({name, table_form_id});
    return PYHIP_OK;
}

int pyhip_codegen_count(int *count) {
    PYHIP_REQUIRE(count, "null argument");
    std::lock_guard<std::mutex> lock(g_mu);
    *count = (int)g_forms.size();
    return PYHIP_OK;
}

}  // extern "C"
