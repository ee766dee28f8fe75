"""Boundary-face forms written in the expression IR, and their HIP face kernels.

A face form is the flux of pyoomph's NeumannBC(u=flux) @ boundary (pyoomph/meshes/bcs.py:
71-89: an InterfaceEquations residual weak(flux, testfunction(u)) whose element code the
reference generates like any other residual, src/codegen.cpp:3912-4109) with a flux that is
an expression in

    u       the field on the face,
    x0, x1  the position,
    c[k]    coefficients of each face (one row of the faces' coefficient array),
    p[k]    the sub-problem's parameters (one of them may be the global parameter, with
            dR/dp as ``dResidual{i}dParameter_{p}``, src/codegen.cpp:4733-4752),
    n0, n1  the outward unit normal (pyoomph's var("normal") on an interface; the
            reference's line-element normal is the tangent dx/ds turned by a quarter over
            its length, src/elements.cpp:1724-1750): here at every point the tangent turned
            clockwise, times the face's side (an extra coefficient after the form's own,
            +-1, which MeshProblem sets so that the normal leaves the bulk element),

so that, on the line faces of P1 / P2 triangles and Q2 quadrilaterals,

    R_i = int flux psi_i ds,    J_ij = int (dflux/du psi_j) psi_i ds.

The kernel has the shape of the library's flux-face kernel (faces.hip): one thread per face,
its nodal coordinates and values in registers, per point the line's tangent, W = w |t|, u
and x, then the point quantities (flux and its u-derivative, printed from the expression
tree by ``expr.CPrinter``), and every residual and Jacobian entry accumulated in point order.
The flux alpha*u + beta written as ``c[0] * u + c[1]`` reproduces the built-in faces bit for
bit; a flux that reads the normal gets it per point from the tangent. The _f64_hv /
_f64_hvt kernels give the Hessian-vector products (the flux's second u-derivative). Faces have no time derivative (the mass matrix is zero).
"""
from . import expr as E

# line faces of an element family: (nodes per face, C++ rule type of element_common.h)
FACE_RULES = {1: (2, "FaceRule<2>"), 2: (3, "FaceRule<3>"), 3: (3, "FaceRuleQ3")}


class FaceSymbols:
    """The symbols a face form's flux is written in (their C names in the kernel)."""

    def __init__(self, ncoef, nparams):
        self.u = E.sym("u")
        self.x = [E.sym("x0"), E.sym("x1")]
        self.c = [E.sym("T(cf[%d])" % k) for k in range(ncoef)]
        self.p = [E.sym("T(par[%d])" % k) for k in range(nparams)]
        self.n = [E.sym("n0"), E.sym("n1")]


class FaceFluxForm:
    """flux(s) -> Expr with s a FaceSymbols: the flux of NeumannBC(u=flux) on the faces.

    ncoef: coefficients per face (the MeshProblem's flux rows); nparams: parameters of the
    sub-problem the flux reads; global_param: index of the parameter with a derivative
    entry (or None)."""

    def __init__(self, name, flux, *, ncoef=0, nparams=0, global_param=None):
        if not name.isidentifier():
            raise ValueError("form name must be a C identifier")
        self.name = name
        self.flux_fn = flux
        self.ncoef = int(ncoef)
        self.nparams = int(nparams)
        self.global_param = global_param

    def build(self):
        s = FaceSymbols(self.ncoef, self.nparams)
        F = E._wrap(self.flux_fn(s))
        return s, F

    def uses_normal(self):
        """The flux reads the normal: each face then carries its side as coefficient ncoef."""
        s, F = self.build()
        return _uses(F, s.n)

    def device_ncoef(self):
        """Coefficients per face on the device (the side of each face after the form's own)."""
        return self.ncoef + (1 if self.uses_normal() else 0)


KERNEL = r'''
#include "element_common.h"
namespace pyhip_gen_{name} {{
using namespace pyhip;
template <typename T, int HV>
__device__ __forceinline__ void body(const ElemArgs<T> &a) {{
    constexpr int NN = {nn};
    using RULE = {rule};
    const int e = a.e0 + (int)(blockIdx.x * 256 + threadIdx.x);
    if (e >= a.e1) return;
    const double *cf = a.fcoef + (int64_t)e * {dncoef};
    const double *par = a.params + (int64_t)a.sub[e] * a.pstride;
    (void)cf;
    (void)par;
    T X[NN][2], U[NN], Y[NN];
#pragma unroll
    for (int l = 0; l < NN; l++) {{
        const int32_t n = a.conn[(int64_t)l * a.nel + e];
        X[l][0] = (T)a.coords[(int64_t)n * 3];
        X[l][1] = (T)a.coords[(int64_t)n * 3 + 1];
        const int32_t q = a.seqn[(int64_t)l * a.nel + e];
        U[l] = (T)(q >= 0 ? a.dofs[q] : a.pinned[(int64_t)n * a.nvalue]);
        Y[l] = (HV && q >= 0) ? (T)a.hvec[q] : T(0);   // the Hessian direction (0 on pinned values)
    }}
    T R[NN], J[NN][NN];
#pragma unroll
    for (int i = 0; i < NN; i++) {{
        R[i] = T(0);
#pragma unroll
        for (int j = 0; j < NN; j++) J[i][j] = T(0);
    }}
#pragma unroll
    for (int g = 0; g < NN; g++) {{
        double sd, wd;
        RULE::point(g, sd, wd);
        const T sq = (T)sd, wq = (T)wd;
        T psi[NN], dpsi[NN];
        RULE::shape(sq, psi, dpsi);
        T tx = T(0), ty = T(0), u = T(0), yh = T(0);
#pragma unroll
        for (int l = 0; l < NN; l++) {{
            tx = tx + X[l][0] * dpsi[l];
            ty = ty + X[l][1] * dpsi[l];
            u = u + U[l] * psi[l];
            if (HV) yh = yh + Y[l] * psi[l];
        }}
        (void)u;
        (void)yh;
{x_block}
        const T W = wq * sqrt(tx * tx + ty * ty);
{n_block}
        T Q[{nqv}];
        T QH = T(0);   // d2 flux / du2 (Hessian-vector products)
        if (HV) {{
{phase_h}
        }}
        if (a.deriv) {{
{phase_a_deriv}
        }} else {{
{phase_a}
        }}
#pragma unroll
        for (int i = 0; i < NN; i++) {{
            R[i] = R[i] + W * ({res});
            if (a.flag == 1) {{
#pragma unroll
                for (int j = 0; j < NN; j++)
                    J[i][j] = J[i][j] + W * (HV ? (((QH * yh) * psi[j]) * psi[i]) : ({jac}));
            }}
        }}
    }}
    // no time derivative on a face (the mass matrix is zero)
    const bool zero_j = a.flag == 2;
#pragma unroll
    for (int i = 0; i < NN; i++) {{
        a.Er[(int64_t)e * NN + i] = R[i];
        if (a.Ej) {{
#pragma unroll
            for (int j = 0; j < NN; j++) a.Ej[((int64_t)e * NN + i) * NN + j] = zero_j ? T(0) : J[i][j];
        }}
    }}
}}
}}  // namespace pyhip_gen_{name}
extern "C" __global__ __launch_bounds__(256) void {name}_f64(pyhip::ElemArgs<double> a) {{
    pyhip_gen_{name}::body<double, 0>(a);
}}
extern "C" __global__ __launch_bounds__(256) void {name}_f64_tr(pyhip::ElemArgs<double> a) {{
    pyhip_gen_{name}::body<double, 0>(a);
}}
extern "C" __global__ __launch_bounds__(256) void {name}_f32(pyhip::ElemArgs<float> a) {{
    pyhip_gen_{name}::body<float, 0>(a);
}}
extern "C" __global__ __launch_bounds__(256) void {name}_f32_tr(pyhip::ElemArgs<float> a) {{
    pyhip_gen_{name}::body<float, 0>(a);
}}
// Hessian-vector products: sum_k dJ_ij/du_k y_k = int (F_uu y) psi_j psi_i, symmetric in
// i and j, so the transposed product is the same matrix
extern "C" __global__ __launch_bounds__(256) void {name}_f64_hv(pyhip::ElemArgs<double> a) {{
    pyhip_gen_{name}::body<double, 1>(a);
}}
extern "C" __global__ __launch_bounds__(256) void {name}_f64_hvt(pyhip::ElemArgs<double> a) {{
    pyhip_gen_{name}::body<double, 1>(a);
}}
'''


def _pieces(s, F, param=None):
    """(flux, dflux/du), or their derivatives in the parameter symbol ``param``."""
    if param is not None:
        F = E.diff(F, param)
    return F, E.diff(F, s.u)


def _phase_b(F, D, quantities, stored):
    """Residual and Jacobian integrands (without W) over the stored point quantities."""
    def q(e):
        if e.op == "const":
            return e
        k = stored.get(e)
        if k is None:
            k = len(quantities)
            stored[e] = k
            quantities.append(e)
        return E.sym("Q[%d]" % k)

    Pi, Pj = E.sym("psi[i]"), E.sym("psi[j]")
    res = q(F) * Pi
    dq = q(D)
    jac = (dq * Pj) * Pi if not E.is_zero(dq) else E.ZERO
    return res, jac


def _indent(lines, n):
    return "\n".join(" " * n + l for l in lines)


def emit_face_kernels(form, kind):
    """HIP source of the face form's kernels on the line faces of element family ``kind``
    ("tri_p1", "tri_p2", "quad_q2" or its id). Returns (source, kernel base name, info)."""
    from .scalar import KIND_IDS
    kid = KIND_IDS[kind] if isinstance(kind, str) else int(kind)
    if kid not in FACE_RULES:
        raise ValueError("boundary-face forms are generated for P1 / P2 triangles and Q2 quadrilaterals")
    nn, rule = FACE_RULES[kid]
    s, F = form.build()
    quantities, stored = [], {}
    res, jac = _phase_b(*_pieces(s, F), quantities, stored)
    if form.global_param is not None:
        q2, st2 = [], {}
        res_d, jac_d = _phase_b(*_pieces(s, F, s.p[form.global_param]), q2, st2)
    else:
        q2, res_d, jac_d = [], E.ZERO, E.ZERO
    nqv = max(len(quantities), len(q2), 1)

    def lines(qs, prefix):
        p = E.CPrinter("T", prefix=prefix)
        return p.assign([("Q[%d]" % k, e) for k, e in enumerate(qs)])

    def b_text(e):
        return E.CPrinter("T", prefix="b").expr_text(e)

    uses_x = any(_uses(e, s.x) for e in [F])
    x_block = ""
    if uses_x:
        xl = []
        for a_ in range(2):
            xl.append("T x%d = T(0);" % a_)
            xl.append("#pragma unroll")
            xl.append("for (int l = 0; l < NN; l++) x%d = x%d + X[l][%d] * psi[l];" % (a_, a_, a_))
        x_block = _indent(xl, 8)
    n_block = ""
    normal = _uses(F, s.n)
    if normal:
        # the tangent turned clockwise, oriented by the face's side, over its length
        n_block = _indent(["const T tl = sqrt(tx * tx + ty * ty), side = T(cf[%d]);" % form.ncoef,
                           "const T n0 = side * ty / tl, n1 = -(side * tx) / tl;",
                           "(void)n0;", "(void)n1;"], 8)
    # the derivative entries of a form without a global parameter are zero: Q stays unset
    # and the integrands are the constant 0
    fuu = E.diff(E.diff(F, s.u), s.u)
    src = KERNEL.format(
        name=form.name, nn=nn, rule=rule, dncoef=form.ncoef + (1 if normal else 0), nqv=nqv,
        x_block=x_block, n_block=n_block,
        phase_h=_indent(E.CPrinter("T", prefix="h").assign([("QH", fuu)]), 12),
        phase_a=_indent(lines(quantities, "a") or ["(void)Q;"], 12),
        phase_a_deriv=_indent(lines(q2, "d") or ["(void)Q;"], 12),
        res="a.deriv ? (%s) : (%s)" % (b_text(res_d), b_text(res)),
        jac="a.deriv ? (%s) : (%s)" % (b_text(jac_d), b_text(jac)))
    info = {"kind": kid, "face_nodes": nn, "ncoef": form.ncoef, "nquantities": nqv, "symmetric": False,
            "has_param_derivative": form.global_param is not None, "normal": normal}
    return src, form.name, info


def _uses(e, syms):
    seen, stack, syms = set(), [e], set(syms)
    while stack:
        x = stack.pop()
        if x in seen:
            continue
        seen.add(x)
        if x in syms:
            return True
        stack.extend(x.args)
    return False


class ConstraintForm:
    """constraint(s) -> Expr with s a FaceSymbols: the constraint c of pyoomph's
    EnforcedBC(u=c) @ boundary (pyoomph/meshes/bcs.py:129-253) on the faces, enforced by a
    Lagrange multiplier l (value 1 of the face nodes): R[u_i] += int l psi_i ds,
    R[l_i] += int c psi_i ds, J[u_i][l_j] = int psi_j psi_i ds, J[l_i][u_j] = int (dc/du psi_j)
    psi_i ds (the residuals weak(l, u_test) and weak(c, l_test), bcs.py:185-189). ncoef,
    nparams, global_param as for FaceFluxForm; c[0] * u + c[1] reproduces the built-in
    constraint faces bit for bit."""

    def __init__(self, name, constraint, *, ncoef=0, nparams=0, global_param=None):
        if not name.isidentifier():
            raise ValueError("form name must be a C identifier")
        self.name = name
        self.constraint_fn = constraint
        self.ncoef = int(ncoef)
        self.nparams = int(nparams)
        self.global_param = global_param

    def build(self):
        s = FaceSymbols(self.ncoef, self.nparams)
        return s, E._wrap(self.constraint_fn(s))

    def uses_normal(self):
        s, F = self.build()
        return _uses(F, s.n)


CONSTRAINT_KERNEL = r'''
#include "element_common.h"
namespace pyhip_gen_{name} {{
using namespace pyhip;
template <typename T, int HV>
__device__ __forceinline__ void body(const ElemArgs<T> &a) {{
    constexpr int NN = {nn};
    constexpr int NS = 2 * NN;   // slots (u_l, l_l) per face node l
    using RULE = {rule};
    const int e = a.e0 + (int)(blockIdx.x * 256 + threadIdx.x);
    if (e >= a.e1) return;
    const double *cf = a.fcoef + (int64_t)e * {dncoef};
    const double *par = a.params + (int64_t)a.sub[e] * a.pstride;
    (void)cf;
    (void)par;
    T X[NN][2], U[NN], Lm[NN], Yu[NN], Yl[NN];
#pragma unroll
    for (int l = 0; l < NN; l++) {{
        const int32_t n = a.conn[(int64_t)l * a.nel + e];
        X[l][0] = (T)a.coords[(int64_t)n * 3];
        X[l][1] = (T)a.coords[(int64_t)n * 3 + 1];
        const int32_t qu = a.seqn[(int64_t)(2 * l) * a.nel + e];
        const int32_t ql = a.seqn[(int64_t)(2 * l + 1) * a.nel + e];
        U[l] = (T)(qu >= 0 ? a.dofs[qu] : a.pinned[(int64_t)n * a.nvalue]);
        Lm[l] = (T)(ql >= 0 ? a.dofs[ql] : a.pinned[(int64_t)n * a.nvalue + 1]);
        // the Hessian direction on the u and multiplier slots (0 on pinned values)
        Yu[l] = (HV && qu >= 0) ? (T)a.hvec[qu] : T(0);
        Yl[l] = (HV && ql >= 0) ? (T)a.hvec[ql] : T(0);
    }}
    T R[NS], J[NS][NS];
#pragma unroll
    for (int i = 0; i < NS; i++) {{
        R[i] = T(0);
#pragma unroll
        for (int j = 0; j < NS; j++) J[i][j] = T(0);
    }}
#pragma unroll
    for (int g = 0; g < NN; g++) {{
        double sd, wd;
        RULE::point(g, sd, wd);
        const T sq = (T)sd, wq = (T)wd;
        T psi[NN], dpsi[NN];
        RULE::shape(sq, psi, dpsi);
        T tx = T(0), ty = T(0), u = T(0), lm = T(0), yu = T(0), yl = T(0);
#pragma unroll
        for (int l = 0; l < NN; l++) {{
            tx = tx + X[l][0] * dpsi[l];
            ty = ty + X[l][1] * dpsi[l];
            u = u + U[l] * psi[l];
            lm = lm + Lm[l] * psi[l];
            if (HV) {{
                yu = yu + Yu[l] * psi[l];
                yl = yl + Yl[l] * psi[l];
            }}
        }}
        (void)u;
{x_block}
        const T W = wq * sqrt(tx * tx + ty * ty);
{n_block}
        T Q[2];
        T QH = T(0);   // d2 c / du2 (Hessian-vector products)
        if (HV) {{
{phase_h}
        }}
        if (a.deriv) {{
{phase_a_deriv}
        }} else {{
{phase_a}
        }}
#pragma unroll
        for (int i = 0; i < NN; i++) {{
            R[2 * i] = R[2 * i] + W * (lm * psi[i]);
            R[2 * i + 1] = R[2 * i + 1] + W * (Q[0] * psi[i]);
            if (a.flag == 1) {{
#pragma unroll
                for (int j = 0; j < NN; j++) {{
                    if (HV == 1) {{
                        // sum_k dJ/dx_k y_k: only J[l_i][u_j] = int c_u psi_j psi_i depends on u
                        J[2 * i + 1][2 * j] = J[2 * i + 1][2 * j] + W * (((QH * yu) * psi[j]) * psi[i]);
                    }} else if (HV == 2) {{
                        // sum_r y_r d2R_r/dx_a dx_b: the constraint rows' int c_uu psi_a psi_b psi_r
                        J[2 * i][2 * j] = J[2 * i][2 * j] + W * (((QH * yl) * psi[j]) * psi[i]);
                    }} else {{
                        J[2 * i][2 * j + 1] = J[2 * i][2 * j + 1] + W * (psi[j] * psi[i]);
                        J[2 * i + 1][2 * j] = J[2 * i + 1][2 * j] + W * ((Q[1] * psi[j]) * psi[i]);
                    }}
                }}
            }}
        }}
    }}
    // derivative in the parameter: only the constraint rows (l does not depend on it); no
    // time derivative (the mass matrix is zero)
    const bool zero_j = a.flag == 2;
#pragma unroll
    for (int i = 0; i < NS; i++) {{
        const bool zero_row = a.deriv && (i % 2 == 0);
        a.Er[(int64_t)e * NS + i] = zero_row ? T(0) : R[i];
        if (a.Ej) {{
#pragma unroll
            for (int j = 0; j < NS; j++)
                a.Ej[((int64_t)e * NS + i) * NS + j] = (zero_j || zero_row) ? T(0) : J[i][j];
        }}
    }}
}}
}}  // namespace pyhip_gen_{name}
extern "C" __global__ __launch_bounds__(256) void {name}_f64(pyhip::ElemArgs<double> a) {{
    pyhip_gen_{name}::body<double, 0>(a);
}}
extern "C" __global__ __launch_bounds__(256) void {name}_f64_tr(pyhip::ElemArgs<double> a) {{
    pyhip_gen_{name}::body<double, 0>(a);
}}
extern "C" __global__ __launch_bounds__(256) void {name}_f32(pyhip::ElemArgs<float> a) {{
    pyhip_gen_{name}::body<float, 0>(a);
}}
extern "C" __global__ __launch_bounds__(256) void {name}_f32_tr(pyhip::ElemArgs<float> a) {{
    pyhip_gen_{name}::body<float, 0>(a);
}}
extern "C" __global__ __launch_bounds__(256) void {name}_f64_hv(pyhip::ElemArgs<double> a) {{
    pyhip_gen_{name}::body<double, 1>(a);
}}
extern "C" __global__ __launch_bounds__(256) void {name}_f64_hvt(pyhip::ElemArgs<double> a) {{
    pyhip_gen_{name}::body<double, 2>(a);
}}
'''


def emit_constraint_kernels(form, kind):
    """HIP source of the constraint form's kernels on the line faces of element family
    ``kind``. Returns (source, kernel base name, info)."""
    from .scalar import KIND_IDS
    kid = KIND_IDS[kind] if isinstance(kind, str) else int(kind)
    if kid not in FACE_RULES:
        raise ValueError("constraint forms are generated for P1 / P2 triangles and Q2 quadrilaterals")
    nn, rule = FACE_RULES[kid]
    s, F = form.build()
    c, dc = _pieces(s, F)
    if form.global_param is not None:
        cd, dcd = _pieces(s, F, s.p[form.global_param])
    else:
        cd, dcd = E.ZERO, E.ZERO

    def lines(exprs, prefix):
        return E.CPrinter("T", prefix=prefix).assign([("Q[%d]" % k, e) for k, e in enumerate(exprs)])

    normal = _uses(F, s.n)
    x_block = ""
    if _uses(F, s.x):
        xl = []
        for a_ in range(2):
            xl.append("T x%d = T(0);" % a_)
            xl.append("#pragma unroll")
            xl.append("for (int l = 0; l < NN; l++) x%d = x%d + X[l][%d] * psi[l];" % (a_, a_, a_))
        x_block = _indent(xl, 8)
    n_block = ""
    if normal:
        n_block = _indent(["const T tl = sqrt(tx * tx + ty * ty), side = T(cf[%d]);" % form.ncoef,
                           "const T n0 = side * ty / tl, n1 = -(side * tx) / tl;",
                           "(void)n0;", "(void)n1;"], 8)
    src = CONSTRAINT_KERNEL.format(
        name=form.name, nn=nn, rule=rule, dncoef=form.ncoef + (1 if normal else 0),
        x_block=x_block, n_block=n_block,
        phase_h=_indent(E.CPrinter("T", prefix="h").assign([("QH", E.diff(dc, s.u))]), 12),
        phase_a=_indent(lines([c, dc], "a"), 12),
        phase_a_deriv=_indent(lines([cd, dcd], "d"), 12))
    info = {"kind": kid, "face_nodes": nn, "ncoef": form.ncoef, "symmetric": False,
            "has_param_derivative": form.global_param is not None, "normal": normal}
    return src, form.name, info
