"""Scalar weak forms written in the expression IR, and their HIP element kernels.

A form is the reference's generic residual (src/codegen.cpp:2015-2257: per integration
point, per test function i, the contributions of a weak form, and per shape function j its
derivative)

    R_i = int  F(u, grad u, x, p) . grad psi_i  +  S(u, grad u, x, p, f) psi_i  dx

with F the flux (one expression per dimension), S the source, p the form's parameters (one
of them may be a global parameter with dR/dp, as ``dResidual{i}dParameter_{p}``,
src/codegen.cpp:4733-4752) and f an auxiliary nodal field. The Jacobian is derived on the
expression tree:

    J_ij = int  (dF_a/d(d_b u) d_b psi_j) d_a psi_i + (dF_a/du psi_j) d_a psi_i
              + (dS/d(d_b u) d_b psi_j) psi_i + (dS/du psi_j) psi_i  dx

``emit_kernels`` prints the batched element kernel of the form for one element family in
the shape of the compiled kernels (element_kernels.h): per element the nodal data in LDS,
per integration point (phase A) the geometry, u, grad u, x, f and the form's point
quantities, then per test row (phase B) the residual entry and the Jacobian row, in the
order the expressions are written. With the time derivative (the transient instantiation)
the residual gains du/dt psi_i and the Jacobian w0 psi_j psi_i, as in the compiled kernels;
flag 2 gives the mass matrix. Four extern "C" kernels: <name>_f64, <name>_f64_tr,
<name>_f32, <name>_f32_tr.
"""
from . import expr as E

KIND_IDS = {"tri_p1": 1, "tri_p2": 2, "quad_q2": 3, "hex_q2": 4}
KIND_DIM = {1: 2, 2: 2, 3: 2, 4: 3}
KIND_NN = {1: 3, 2: 6, 3: 9, 4: 27}


class FormSymbols:
    """The symbols a form's expressions are written in (their C names in phase A)."""

    def __init__(self, dim, nparams):
        self.dim = dim
        self.u = E.sym("u")
        self.grad_u = [E.sym("g%d" % a) for a in range(dim)]
        self.x = [E.sym("x%d" % a) for a in range(dim)]
        self.f = E.sym("f")
        self.p = [E.sym("T(par[%d])" % k) for k in range(nparams)]


class ScalarWeakForm:
    """flux(s) -> [dim Exprs], source(s) -> Expr, with s a FormSymbols.

    nparams: parameters per sub-problem (params array of the problem); global_param: index
    of the parameter with a derivative entry (or None) and its name; aux: the form reads the
    auxiliary nodal field f; tables: named residual tables {name: (flux, source)} of the same
    unknown (selected with Problem._set_solved_residual / assemble_jacobian(which_one=name))."""

    def __init__(self, name, flux, source, *, nparams=0, global_param=None, global_param_name="lambda",
                 aux=False, tables=None):
        if not name.isidentifier():
            raise ValueError("form name must be a C identifier")
        self.name = name
        self.flux_fn, self.source_fn = flux, source
        # named residual tables: {name: (flux, source)}, the residuals pyoomph collects with
        # add_residual(..., destination=name) (pyoomph/generic/codegen.py:777-800)
        self.tables = dict(tables or {})
        self.nparams = int(nparams)
        self.global_param = global_param
        self.global_param_name = global_param_name if global_param is not None else None
        self.aux = bool(aux)

    def build(self, dim):
        s = FormSymbols(dim, self.nparams)
        F = list(self.flux_fn(s))
        S = E._wrap(self.source_fn(s))
        if len(F) != dim:
            raise ValueError("the flux needs %d components" % dim)
        F = [E._wrap(v) for v in F]
        uses_x = any(_uses(e, s.x) for e in F + [S])
        uses_f = any(_uses(e, [s.f]) for e in F + [S])
        if uses_f and not self.aux:
            raise ValueError("the form reads the auxiliary field f: pass aux=True")
        return s, F, S, uses_x

    def symmetric(self, dim):
        """Jacobian symmetric for every state: dF/d(grad u) symmetric, no dF/du or dS/d(grad u)
        coupling (then conjugate gradients apply where it is also positive definite)."""
        s, F, S, _ = self.build(dim)
        K = [[E.diff(F[a], s.grad_u[b]) for b in range(dim)] for a in range(dim)]
        if any(K[a][b] is not K[b][a] for a in range(dim) for b in range(dim)):
            return False
        if any(not E.is_zero(E.diff(F[a], s.u)) for a in range(dim)):
            return False
        if any(not E.is_zero(E.diff(S, s.grad_u[b])) for b in range(dim)):
            return False
        return True


def _uses(e, syms):
    seen = set()
    stack = [e]
    syms = set(syms)
    while stack:
        x = stack.pop()
        if x in seen:
            continue
        seen.add(x)
        if x in syms:
            return True
        stack.extend(x.args)
    return False


class _PhaseB:
    """Phase-B symbols: the point quantities stored in LDS and the test / shape data."""

    def __init__(self, dim):
        self.W = E.sym("W")
        self.Pi = E.sym("P[row]")
        self.Pj = E.sym("P[j]")
        self.dxi = [E.sym("dxi[%d]" % a) for a in range(dim)]
        self.dxj = [E.sym("dxj[%d]" % a) for a in range(dim)]


def _entries(form, dim, s, F, S, param=None):
    """(residual integrand pieces, Jacobian coefficient pieces) of the form, or of its
    derivative in the parameter symbol ``param``."""
    if param is not None:
        F = [E.diff(v, param) for v in F]
        S = E.diff(S, param)
    K = [[E.diff(F[a], s.grad_u[b]) for b in range(dim)] for a in range(dim)]
    G = [E.diff(F[a], s.u) for a in range(dim)]
    H = [E.diff(S, s.grad_u[b]) for b in range(dim)]
    D = E.diff(S, s.u)
    return F, S, K, G, H, D


def _store(quantities, e, stored):
    """Phase-B symbol of point quantity e: symbols and constants print as themselves in
    phase B only if they are constants; everything else is stored per point."""
    if e.op == "const":
        return e
    k = stored.get(e)
    if k is None:
        k = len(quantities)
        stored[e] = k
        quantities.append(e)
    return E.sym("Q[%d]" % k)


def _dot(a, b):
    acc = E.ZERO
    for x, y in zip(a, b):
        acc = acc + x * y
    return acc


def _residual_jacobian_exprs(dim, pieces, quantities, stored, B):
    """Phase-B expressions (residual integrand without W, Jacobian integrand without W)."""
    F, S, K, G, H, D = pieces
    q = lambda e: _store(quantities, e, stored)  # noqa: E731
    Fq = [q(v) for v in F]
    flux = _dot(Fq, B.dxi)
    res = flux + q(S) * B.Pi
    # Jacobian: flux part (sum_a (sum_b K_ab dxj_b) dxi_a), dF/du, dS/d(grad u), dS/du
    kj = [_dot([q(K[a][b]) for b in range(dim)], B.dxj) for a in range(dim)]
    jac = _dot(kj, B.dxi)
    g = _dot([q(v) for v in G], B.dxi)
    if not E.is_zero(g):
        jac = jac + g * B.Pj
    h = _dot([q(v) for v in H], B.dxj)
    if not E.is_zero(h):
        jac = jac + h * B.Pi
    dq = q(D)
    if not E.is_zero(dq):
        jac = jac + (dq * B.Pj) * B.Pi
    return res, jac


KERNEL = r'''
#include "element_common.h"
namespace pyhip_gen_{name} {{
using namespace pyhip;
constexpr int KIND = {kind};
// HV: 0 the residual / Jacobian / mass matrix kernels; 1 the Hessian-vector product
// sum_j y_j d2R_i/du_j du_k (the Jacobian of the form's derivative in the direction y),
// 2 the transposed product sum_j y_j d2R_j/du_i du_k (the Jacobian of the first variation of
// int F . grad y + S y); the residual output stays the form's residual
template <typename T, bool TR, int HV>
__device__ __forceinline__ void body(const ElemArgs<T> &a) {{
    using Tr = ElemTraits<KIND>;
    constexpr int NN = Tr::NN, DIM = Tr::DIM, NQ = Tr::NQ, NS = NN;
    constexpr int EPB = scalar_epb<KIND>();
    constexpr int QC = ScalarChunk<KIND>::QC;
    constexpr int NQV = {nqv};     // point quantities
    constexpr int NQH = HV ? {nqh} : 0;   // [+ Hessian point quantities]
    constexpr int NQS = NQV + NQH + (TR ? 1 : 0);   // [+ du/dt]
    __shared__ T sX[EPB][NN][DIM];
    __shared__ T sU[EPB][NN];
    __shared__ T sH[TR ? EPB : 1][NN];
    __shared__ T sF[{aux} ? EPB : 1][NN];
    __shared__ int sEq[EPB][NS];
    __shared__ T sW[EPB][QC];
    __shared__ T sDx[EPB][QC][NN][DIM];
    __shared__ T sQ[EPB][QC][NQS];
    __shared__ T sY[HV ? EPB : 1][NN];
    const int le = threadIdx.x / NS, row = threadIdx.x % NS;
    const int e = a.e0 + (int)blockIdx.x * EPB + le;
    const bool active = (le < EPB) && (e < a.e1);
    if (active) {{
        const int node = a.conn[(int64_t)row * a.nel + e];
        const int eq = a.seqn[(int64_t)row * a.nel + e];
#pragma unroll
        for (int d = 0; d < DIM; d++) sX[le][row][d] = T(a.coords[(int64_t)node * 3 + d]);
        sU[le][row] = T(eq >= 0 ? a.dofs[eq] : a.pinned[(int64_t)node * a.nvalue]);
        if constexpr ({aux}) sF[le][row] = T(a.aux[node]);
        if constexpr (TR) {{
            const double pin = a.pinned[(int64_t)node * a.nvalue];
            const double u1 = eq >= 0 ? a.hist1[eq] : pin, u2 = eq >= 0 ? a.hist2[eq] : pin;
            sH[le][row] = T(a.tw1 * u1 + a.tw2 * u2);
        }}
        if constexpr (HV > 0) sY[le][row] = T(eq >= 0 ? a.hvec[eq] : 0.0);
        sEq[le][row] = eq;
    }}
    __syncthreads();
    const int la = threadIdx.x / QC, qa = threadIdx.x % QC;
    const int ea = a.e0 + (int)blockIdx.x * EPB + la;
    const bool active_a = (la < EPB) && (ea < a.e1);
    const double *par = nullptr;
    if (active_a) par = a.params + (int64_t)a.sub[ea] * a.pstride;
    const bool myrow = active && sEq[le][row] >= 0;
    T Racc = T(0);
    T Jacc[NS];
#pragma unroll
    for (int j = 0; j < NS; j++) Jacc[j] = T(0);
    for (int q0 = 0; q0 < NQ; q0 += QC) {{
        if (active_a && q0 + qa < NQ) {{
            const int le = la, qq = qa;
            const int q = q0 + qq;
            const T *D = a.tab->dpsi + q * NN * DIM;
            const T *P = a.tab->psi + q * NN;
            const T W = point_geometry<T, NN, DIM>(sX[le], D, a.tab->w[q], sDx[le][qq]);
            sW[le][qq] = W;
            T u = T(0);
#pragma unroll
            for (int l = 0; l < NN; l++) u = u + sU[le][l] * P[l];
            T gv[DIM];
#pragma unroll
            for (int i = 0; i < DIM; i++) {{
                T sacc = T(0);
#pragma unroll
                for (int l = 0; l < NN; l++) sacc = sacc + sU[le][l] * sDx[le][qq][l][i];
                gv[i] = sacc;
            }}
{grad_names}
{x_block}
{f_block}
            T *Q = sQ[le][qq];
            if (a.deriv) {{
{phase_a_deriv}
            }} else {{
{phase_a}
            }}
            if constexpr (HV > 0) {{
                // the direction y and its gradient at the point
                T yv = T(0);
#pragma unroll
                for (int l = 0; l < NN; l++) yv = yv + sY[le][l] * P[l];
                T gyv[DIM];
#pragma unroll
                for (int i = 0; i < DIM; i++) {{
                    T sacc = T(0);
#pragma unroll
                    for (int l = 0; l < NN; l++) sacc = sacc + sY[le][l] * sDx[le][qq][l][i];
                    gyv[i] = sacc;
                }}
{gy_names}
                T *QH = Q + NQV;
                (void)QH;
                (void)yv;
                if constexpr (HV == 1) {{
{phase_a_hv}
                }} else {{
{phase_a_hvt}
                }}
            }}
            if constexpr (TR) {{
                T hq = T(0);
#pragma unroll
                for (int l = 0; l < NN; l++) hq = hq + sH[le][l] * P[l];
                sQ[le][qq][NQV + NQH] = T(a.tw0) * u + hq;
            }}
        }}
        __syncthreads();
        if (myrow) {{
            for (int qq = 0; qq < QC && q0 + qq < NQ; qq++) {{
                const int q = q0 + qq;
                const T *P = a.tab->psi + q * NN;
                const T W = sW[le][qq];
                const T *dxi = sDx[le][qq][row];
                const T *Q = sQ[le][qq];
                const bool tr = TR && !a.deriv;
                if (a.deriv)
                    Racc = Racc + W * ({res_deriv});
                else if (tr)
                    Racc = Racc + W * (({res}) + Q[NQV] * P[row]);
                else
                    Racc = Racc + W * ({res});
                if (a.flag == 2 && !TR) continue;
                const T *QH = Q + NQV;
                (void)QH;
                if (a.flag) {{
                    const T tw0 = T(a.tw0);
#pragma unroll
                    for (int j = 0; j < NS; j++) {{
                        if (sEq[le][j] < 0) continue;
                        const T *dxj = sDx[le][qq][j];
                        T v;
                        if constexpr (HV == 1)
                            v = W * ({jac_hv});
                        else if constexpr (HV == 2)
                            v = W * ({jac_hvt});
                        else if (a.flag == 2)
                            v = W * (P[j] * P[row]);
                        else if (a.deriv)
                            v = W * ({jac_deriv});
                        else
                            v = tr ? W * (({jac}) + (tw0 * P[j]) * P[row]) : W * ({jac});
                        Jacc[j] = Jacc[j] + v;
                    }}
                }}
            }}
        }}
        if (q0 + QC < NQ) __syncthreads();
    }}
    if (myrow) {{
        a.Er[(int64_t)e * NS + row] = Racc;
        if (a.flag) {{
            T *dst = a.Ej + ((int64_t)e * NS + row) * NS;
#pragma unroll
            for (int j = 0; j < NS; j++) dst[j] = Jacc[j];
        }}
    }}
}}
}}  // namespace pyhip_gen_{name}
extern "C" __global__ __launch_bounds__(256) void {name}_f64(pyhip::ElemArgs<double> a) {{
    pyhip_gen_{name}::body<double, false, 0>(a);
}}
extern "C" __global__ __launch_bounds__(256) void {name}_f64_tr(pyhip::ElemArgs<double> a) {{
    pyhip_gen_{name}::body<double, true, 0>(a);
}}
extern "C" __global__ __launch_bounds__(256) void {name}_f32(pyhip::ElemArgs<float> a) {{
    pyhip_gen_{name}::body<float, false, 0>(a);
}}
extern "C" __global__ __launch_bounds__(256) void {name}_f32_tr(pyhip::ElemArgs<float> a) {{
    pyhip_gen_{name}::body<float, true, 0>(a);
}}
extern "C" __global__ __launch_bounds__(256) void {name}_f64_hv(pyhip::ElemArgs<double> a) {{
    pyhip_gen_{name}::body<double, false, 1>(a);
}}
extern "C" __global__ __launch_bounds__(256) void {name}_f64_hvt(pyhip::ElemArgs<double> a) {{
    pyhip_gen_{name}::body<double, false, 2>(a);
}}
'''


def _indent(lines, n):
    return "\n".join(" " * n + l for l in lines)


def emit_kernels(form, kind):
    """HIP source of the form's element kernels for element family ``kind`` ("tri_p1",
    "tri_p2", "quad_q2", "hex_q2" or its id). Returns (source, kernel base name, info)."""
    kid = KIND_IDS[kind] if isinstance(kind, str) else int(kind)
    dim = KIND_DIM[kid]
    s, F, S, uses_x = form.build(dim)
    B = _PhaseB(dim)
    quantities, stored = [], {}
    pieces = _entries(form, dim, s, F, S)
    res, jac = _residual_jacobian_exprs(dim, pieces, quantities, stored, B)
    nq_main = len(quantities)
    main_q = list(quantities)
    if form.global_param is not None:
        pd = _entries(form, dim, s, F, S, param=s.p[form.global_param])
        q2, st2 = [], {}
        res_d, jac_d = _residual_jacobian_exprs(dim, pd, q2, st2, B)
    else:
        q2, res_d, jac_d = [], E.ZERO, E.ZERO
    nqv = max(nq_main, len(q2), 1)
    # Hessian-vector products: the Jacobians of two derived forms in the direction y
    yv = E.sym("yv")
    gy = [E.sym("gy%d" % a_) for a_ in range(dim)]
    dirv = lambda e: E.diff(e, s.u) * yv + _dot([E.diff(e, s.grad_u[b]) for b in range(dim)], gy)  # noqa: E731
    Fy, Sy = [dirv(v) for v in F], dirv(S)
    phi = _dot(F, gy) + S * yv
    Ft, St = [E.diff(phi, s.grad_u[a_]) for a_ in range(dim)], E.diff(phi, s.u)
    BH = _PhaseB(dim)
    qh, sth = [], {}
    _, jac_hv = _residual_jacobian_exprs(dim, _entries(form, dim, s, Fy, Sy), qh, sth, BH)
    qht, stht = [], {}
    _, jac_hvt = _residual_jacobian_exprs(dim, _entries(form, dim, s, Ft, St), qht, stht, BH)
    nqh = max(len(qh), len(qht), 1)

    def phase_b_text(e):
        return E.CPrinter("T", prefix="b").expr_text(e)

    # phase A: compute the quantities into Q[k]; gradient symbols g_a are the phase-A
    # gradient values
    def phase_a_lines(qs, prefix):
        p = E.CPrinter("T", prefix=prefix)
        return p.assign([("Q[%d]" % k, e) for k, e in enumerate(qs)])

    x_block = ""
    if uses_x:
        xl = []
        for a_ in range(dim):
            xl.append("T x%d = T(0);" % a_)
            xl.append("#pragma unroll")
            xl.append("for (int l = 0; l < NN; l++) x%d = x%d + sX[le][l][%d] * P[l];" % (a_, a_, a_))
        x_block = _indent(xl, 12)
    f_block = ""
    if form.aux:
        f_block = _indent(["T f = T(0);", "#pragma unroll", "for (int l = 0; l < NN; l++) f = f + sF[le][l] * P[l];"],
                          12)
    grad_names = _indent(["const T g%d = gv[%d];" % (a_, a_) for a_ in range(dim)] +
                         ["(void)g%d;" % a_ for a_ in range(dim)], 12)
    gy_names = _indent(["const T gy%d = gyv[%d];" % (a_, a_) for a_ in range(dim)] +
                       ["(void)gy%d;" % a_ for a_ in range(dim)], 16)

    def hv_lines(qs, prefix):
        p = E.CPrinter("T", prefix=prefix)
        return p.assign([("QH[%d]" % k, e) for k, e in enumerate(qs)])

    def hv_text(e):
        # phase B reads the Hessian quantities through QH = Q + NQV
        return phase_b_text(e).replace("Q[", "QH[")
    src = KERNEL.format(
        name=form.name, kind=kid, nqv=nqv, aux="true" if form.aux else "false",
        grad_names=grad_names, x_block=x_block, f_block=f_block,
        phase_a=_indent(phase_a_lines(main_q, "a") or ["(void)Q;"], 16),
        phase_a_deriv=_indent(phase_a_lines(q2, "d") or ["(void)Q;"], 16),
        res=phase_b_text(res), jac=phase_b_text(jac),
        res_deriv=phase_b_text(res_d), jac_deriv=phase_b_text(jac_d),
        nqh=nqh, gy_names=gy_names,
        phase_a_hv=_indent(hv_lines(qh, "h") or ["(void)QH;"], 20),
        phase_a_hvt=_indent(hv_lines(qht, "k") or ["(void)QH;"], 20),
        jac_hv=hv_text(jac_hv), jac_hvt=hv_text(jac_hvt))
    info = {"kind": kid, "dim": dim, "nslot": KIND_NN[kid], "nquantities": nqv,
            "symmetric": form.symmetric(dim), "has_param_derivative": form.global_param is not None}
    return src, form.name, info
