"""Taylor-Hood (Q2 velocity / Q1 pressure) weak forms and their HIP element kernels.

The reference's generic residual for a form with a C2 vector field u = (u_0, u_1) and a C1
field p (src/codegen.cpp:2015-2257, one residual per test function, fields in the
element's local dof order), written here as

    R_{a,i} = int  F_a(U) . grad psi_i + S_a(U) psi_i [+ m(U) du_a/dt psi_i]   (velocity rows)
    R_{p,c} = int  S_p(U) phi_c                                                 (pressure rows)

with U = (u, grad u, p, x, parameters), psi the Q2 basis, phi the Q1 pressure sub-basis of
the Q2 element (src/elements.cpp:9288-9305). The Jacobian is derived on the expression tree
(derivatives in u_b, d_e u_b and p, times psi_j, d_e psi_j and phi_c), and so is dR/dp for
a global parameter (``dResidual{i}dParameter_{p}``, src/codegen.cpp:4733-4752). The mass
coefficient m gives the BDF time term and the flag-2 mass matrix m psi_j psi_i on the
diagonal velocity blocks.

``emit_th_kernels`` prints the batched element kernel in the shape of the hand-written
Taylor-Hood kernel (element_kernels.h): 22 slots per element in oomph's node-major order
(u_0, u_1 at every node, p at the corners), nodal data in LDS, per integration point the
geometry and the point quantities, per slot row the residual entry and the Jacobian row.
Four extern "C" kernels <name>_f64, <name>_f64_tr, <name>_f32, <name>_f32_tr, and the
Hessian-vector products <name>_f64_hv / _hvt; the library registers them with the
Taylor-Hood value layout (``pyhip_codegen_set_layout``).
"""
from . import expr as E
from .scalar import _store

NN, NS = 9, 22


class THSymbols:
    """The point symbols a Taylor-Hood form's expressions are written in."""

    def __init__(self, nparams):
        self.u = [E.sym("u0"), E.sym("u1")]
        self.grad_u = [[E.sym("g%d%d" % (a, d)) for d in range(2)] for a in range(2)]   # d_d u_a
        self.p = E.sym("pp")
        self.x = [E.sym("x0"), E.sym("x1")]
        self.par = [E.sym("T(par[%d])" % k) for k in range(nparams)]
        # the Hessian direction at the point (Hessian-vector product kernels only)
        self.yv = [E.sym("y0"), E.sym("y1")]
        self.gy = [[E.sym("gy%d%d" % (a, d)) for d in range(2)] for a in range(2)]
        self.yp = E.sym("yp")


class TaylorHoodWeakForm:
    """flux(s) -> [[F_00, F_01], [F_10, F_11]] (F_a . grad psi_i = F_a0 d_0 psi_i + F_a1 d_1
    psi_i), source(s) -> [S_0, S_1], continuity(s) -> S_p, mass(s) -> m or None (steady
    only), with s a THSymbols. nparams: parameters per sub-problem; global_param: index of
    the parameter with a derivative entry (or None)."""

    def __init__(self, name, flux, source, continuity, *, mass=None, nparams=0, global_param=None,
                 global_param_name="lambda"):
        if not name.isidentifier():
            raise ValueError("form name must be a C identifier")
        self.name = name
        self.flux_fn, self.source_fn, self.cont_fn, self.mass_fn = flux, source, continuity, mass
        self.nparams = int(nparams)
        self.global_param = global_param
        self.global_param_name = global_param_name if global_param is not None else None
        self.aux = False

    def build(self):
        s = THSymbols(self.nparams)
        F = [[E._wrap(v) for v in row] for row in self.flux_fn(s)]
        S = [E._wrap(v) for v in self.source_fn(s)]
        Sp = E._wrap(self.cont_fn(s))
        m = E._wrap(self.mass_fn(s)) if self.mass_fn is not None else E.ZERO
        if len(F) != 2 or any(len(r) != 2 for r in F) or len(S) != 2:
            raise ValueError("a 2D Taylor-Hood form has a 2 x 2 flux and 2 sources")
        return s, F, S, Sp, m

    def symmetric(self, dim=2):
        return False   # saddle-point Jacobians: GMRES


class _B:
    """Phase-B symbols."""
    Pi, Pj = E.sym("P[lt]"), E.sym("P[ls]")
    Ci, Cj = E.sym("C[pi]"), E.sym("C[pj]")
    dxi = [E.sym("dxl[0]"), E.sym("dxl[1]")]
    dxj = [E.sym("dxs[0]"), E.sym("dxs[1]")]


def _dot(a, b):
    acc = E.ZERO
    for x, y in zip(a, b):
        acc = acc + x * y
    return acc


def _blocks(s, F, S, Sp, q):
    """Residual and Jacobian integrands (without W) per (row kind, column kind); q stores a
    point quantity and returns its phase-B symbol."""
    B = _B
    res = {}
    for a in range(2):
        res[a] = _dot([q(F[a][d]) for d in range(2)], B.dxi) + q(S[a]) * B.Pi
    res[2] = q(Sp) * B.Ci
    jac = {}
    for a in range(2):
        for b in range(2):
            # flux: sum_d (sum_e dF_ad/d(d_e u_b) d_e psi_j) d_d psi_i + (dF_ad/du_b psi_j) d_d psi_i
            kj = [_dot([q(E.diff(F[a][d], s.grad_u[b][e])) for e in range(2)], B.dxj) for d in range(2)]
            v = _dot(kj, B.dxi)
            g = _dot([q(E.diff(F[a][d], s.u[b])) for d in range(2)], B.dxi)
            if not E.is_zero(g):
                v = v + g * B.Pj
            h = _dot([q(E.diff(S[a], s.grad_u[b][e])) for e in range(2)], B.dxj)
            if not E.is_zero(h):
                v = v + h * B.Pi
            dd = q(E.diff(S[a], s.u[b]))
            if not E.is_zero(dd):
                v = v + (dd * B.Pj) * B.Pi
            jac[(a, b)] = v
        # pressure column c: sum_d dF_ad/dp phi_c d_d psi_i + dS_a/dp phi_c psi_i
        v = _dot([q(E.diff(F[a][d], s.p)) for d in range(2)], B.dxi) * B.Cj
        dsp = q(E.diff(S[a], s.p))
        if not E.is_zero(dsp):
            v = v + (dsp * B.Cj) * B.Pi
        jac[(a, 2)] = v
    for b in range(2):
        v = _dot([q(E.diff(Sp, s.grad_u[b][e])) for e in range(2)], B.dxj)
        du = q(E.diff(Sp, s.u[b]))
        if not E.is_zero(du):
            v = v + du * B.Pj
        jac[(2, b)] = v * B.Ci
    jac[(2, 2)] = (q(E.diff(Sp, s.p)) * B.Cj) * B.Ci
    return res, jac


KERNEL = r'''
#include "element_common.h"
namespace pyhip_gen_{name} {{
using namespace pyhip;
__device__ __constant__ const int8_t kSlotNode[22] = {{0, 0, 0, 1, 1, 2, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 6, 7, 7, 8, 8, 8}};
__device__ __constant__ const int8_t kSlotKind[22] = {{0, 1, 2, 0, 1, 0, 1, 2, 0, 1, 0, 1, 0, 1, 0, 1, 2, 0, 1, 0, 1, 2}};
__device__ __constant__ const int8_t kSlotP[22] = {{-1, -1, 0, -1, -1, -1, -1, 1, -1, -1, -1, -1, -1, -1, -1, -1, 2, -1, -1, -1, -1, 3}};
__device__ __constant__ const int8_t kNodeSlot[9] = {{0, 3, 5, 8, 10, 12, 14, 17, 19}};
template <typename T, bool TR, int HV>
__device__ __forceinline__ void body(const ElemArgs<T> &a) {{
    constexpr int NN = 9, DIM = 2, NQ = 9, NS = 22;
    constexpr int EPB = 256 / NS;
    constexpr int NQV = {nqv};                 // point quantities
    constexpr int NQH = HV ? {nqh} : 0;        // Hessian-product quantities (HV 1 plain, 2 transposed)
    constexpr int NQS = NQV + (TR ? 3 : 0) + NQH;   // [+ du_0/dt, du_1/dt, m]
    __shared__ T sX[EPB][NN][DIM];
    __shared__ T sV[EPB][NS];
    __shared__ int sEq[EPB][NS];
    __shared__ T sW[EPB][NQ];
    __shared__ T sDx[EPB][NQ][NN][DIM];
    __shared__ T sQ[EPB][NQ][NQS];
    __shared__ T sH[TR ? EPB : 1][NS];
    __shared__ T sY[HV ? EPB : 1][NS];
    const int le = threadIdx.x / NS, slot = threadIdx.x % NS;
    const int e = a.e0 + (int)blockIdx.x * EPB + le;
    const bool active = (le < EPB) && (e < a.e1);
    if (active) {{
        const int ln = kSlotNode[slot], kind = kSlotKind[slot];
        const int node = a.conn[(int64_t)ln * a.nel + e];
        const int eq = a.seqn[(int64_t)slot * a.nel + e];
        if (slot < NN) {{
            const int gn = a.conn[(int64_t)slot * a.nel + e];
            sX[le][slot][0] = T(a.coords[(int64_t)gn * 3 + 0]);
            sX[le][slot][1] = T(a.coords[(int64_t)gn * 3 + 1]);
        }}
        sV[le][slot] = T(eq >= 0 ? a.dofs[eq] : a.pinned[(int64_t)node * a.nvalue + kind]);
        if constexpr (TR) {{
            const double pin = a.pinned[(int64_t)node * a.nvalue + kind];
            const double h1 = eq >= 0 ? a.hist1[eq] : pin, h2 = eq >= 0 ? a.hist2[eq] : pin;
            sH[le][slot] = T(a.tw1 * h1 + a.tw2 * h2);
        }}
        sEq[le][slot] = eq;
        if constexpr (HV != 0) sY[le][slot] = T(eq >= 0 ? a.hvec[eq] : 0.0);   // 0 on pinned values
    }}
    __syncthreads();
    const double *par = nullptr;
    if (active) par = a.params + (int64_t)a.sub[e] * a.pstride;
    if (active && slot < NQ) {{
        const int q = slot;
        const T *D = a.tab->dpsi + q * NN * DIM;
        const T *P = a.tab->psi + q * NN;
        const T *C = a.tab->psi1 + q * 4;
        sW[le][q] = point_geometry<T, NN, DIM>(sX[le], D, a.tab->w[q], sDx[le][q]);
        T u0 = T(0), u1 = T(0), pp = T(0);
#pragma unroll
        for (int l = 0; l < NN; l++) u0 = u0 + sV[le][kNodeSlot[l]] * P[l];
#pragma unroll
        for (int l = 0; l < NN; l++) u1 = u1 + sV[le][1 + kNodeSlot[l]] * P[l];
        T gv[2][2];
#pragma unroll
        for (int b = 0; b < 2; b++) {{
            T s0 = T(0), s1 = T(0);
#pragma unroll
            for (int l = 0; l < NN; l++) s0 = s0 + sV[le][kNodeSlot[l]] * sDx[le][q][l][b];
#pragma unroll
            for (int l = 0; l < NN; l++) s1 = s1 + sV[le][kNodeSlot[l] + 1] * sDx[le][q][l][b];
            gv[0][b] = s0;
            gv[1][b] = s1;
        }}
        pp = pp + sV[le][2] * C[0];
        pp = pp + sV[le][7] * C[1];
        pp = pp + sV[le][16] * C[2];
        pp = pp + sV[le][21] * C[3];
        const T g00 = gv[0][0], g01 = gv[0][1], g10 = gv[1][0], g11 = gv[1][1];
        (void)g00; (void)g01; (void)g10; (void)g11; (void)u0; (void)u1; (void)pp;
{x_block}
        T *Q = sQ[le][q];
        if constexpr (HV != 0) {{
            // the direction y at the point: values, gradients, pressure
            T y0 = T(0), y1 = T(0), yp = T(0), gy[2][2];
#pragma unroll
            for (int l = 0; l < NN; l++) y0 = y0 + sY[le][kNodeSlot[l]] * P[l];
#pragma unroll
            for (int l = 0; l < NN; l++) y1 = y1 + sY[le][1 + kNodeSlot[l]] * P[l];
#pragma unroll
            for (int b = 0; b < 2; b++) {{
                T s0 = T(0), s1 = T(0);
#pragma unroll
                for (int l = 0; l < NN; l++) s0 = s0 + sY[le][kNodeSlot[l]] * sDx[le][q][l][b];
#pragma unroll
                for (int l = 0; l < NN; l++) s1 = s1 + sY[le][kNodeSlot[l] + 1] * sDx[le][q][l][b];
                gy[0][b] = s0;
                gy[1][b] = s1;
            }}
            yp = yp + sY[le][2] * C[0];
            yp = yp + sY[le][7] * C[1];
            yp = yp + sY[le][16] * C[2];
            yp = yp + sY[le][21] * C[3];
            const T gy00 = gy[0][0], gy01 = gy[0][1], gy10 = gy[1][0], gy11 = gy[1][1];
            (void)gy00; (void)gy01; (void)gy10; (void)gy11; (void)y0; (void)y1; (void)yp;
            if constexpr (HV == 1) {{
{phase_hv}
            }} else {{
{phase_hvt}
            }}
        }}
        if (a.deriv) {{
{phase_a_deriv}
        }} else {{
{phase_a}
        }}
        if constexpr (TR) {{
            T h0 = T(0), h1 = T(0);
#pragma unroll
            for (int l = 0; l < NN; l++) h0 = h0 + sH[le][kNodeSlot[l]] * P[l];
#pragma unroll
            for (int l = 0; l < NN; l++) h1 = h1 + sH[le][kNodeSlot[l] + 1] * P[l];
            Q[NQV] = T(a.tw0) * u0 + h0;
            Q[NQV + 1] = T(a.tw0) * u1 + h1;
{mass_line}
        }}
    }}
    __syncthreads();
    const bool myrow = active && sEq[le][slot] >= 0;
    if (!myrow) return;
    const int kt = kSlotKind[slot], lt = kSlotNode[slot];
    const int pi = kt == 2 ? kSlotP[slot] : 0;
    const bool tr = TR && !a.deriv;
    T Racc = T(0);
    T Jacc[NS];
#pragma unroll
    for (int j = 0; j < NS; j++) Jacc[j] = T(0);
    for (int q = 0; q < NQ; q++) {{
        const T *P = a.tab->psi + q * NN;
        const T *C = a.tab->psi1 + q * 4;
        const T W = sW[le][q];
        const T *dxl = sDx[le][q][lt];
        const T *Q = sQ[le][q];
        T r;
        if (a.deriv) {{
            r = kt == 0 ? ({rd0}) : (kt == 1 ? ({rd1}) : ({rdp}));
        }} else {{
            r = kt == 0 ? ({r0}) : (kt == 1 ? ({r1}) : ({rp}));
            if (tr && kt < 2) r = r + (Q[NQV + 2] * Q[NQV + kt]) * P[lt];
        }}
        Racc = Racc + W * r;
        if (a.flag == 2 && !TR) continue;
        if (!a.flag) continue;
#pragma unroll
        for (int ss = 0; ss < NS; ss++) {{
            if (sEq[le][ss] < 0) continue;
            const int ks = kSlotKind[ss], ls = kSlotNode[ss];
            const int pj = ks == 2 ? kSlotP[ss] : 0;
            const T *dxs = sDx[le][q][ls];
            (void)pj; (void)dxs;
            T v;
            if (a.flag == 2) {{
                if (kt < 2 && kt == ks) v = W * ((Q[NQV + 2] * P[ls]) * P[lt]);
                else continue;
            }} else if (HV == 1) {{
{jac_hv}
            }} else if (HV == 2) {{
{jac_hvt}
            }} else if (a.deriv) {{
{jac_deriv}
            }} else {{
{jac}
                if (tr && kt < 2 && kt == ks) v = v + W * (((Q[NQV + 2] * T(a.tw0)) * P[ls]) * P[lt]);
            }}
            Jacc[ss] = Jacc[ss] + v;
        }}
    }}
    a.Er[(int64_t)e * NS + slot] = Racc;
    if (a.flag) {{
        T *dst = a.Ej + ((int64_t)e * NS + slot) * NS;
#pragma unroll
        for (int j = 0; j < NS; j++) dst[j] = Jacc[j];
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
// Hessian-vector products: the Jacobians of the form's derivative in the direction y
// (plain) and of the gradient of y . R (transposed), steady
extern "C" __global__ __launch_bounds__(256) void {name}_f64_hv(pyhip::ElemArgs<double> a) {{
    pyhip_gen_{name}::body<double, false, 1>(a);
}}
extern "C" __global__ __launch_bounds__(256) void {name}_f64_hvt(pyhip::ElemArgs<double> a) {{
    pyhip_gen_{name}::body<double, false, 2>(a);
}}
'''


def _indent(lines, n):
    return "\n".join(" " * n + l for l in lines)


def _jac_switch(jac, text):
    """Phase-B selection of the Jacobian entry by (row kind kt, column kind ks)."""
    lines = []
    first = True
    for kt in range(3):
        for ks in range(3):
            e = jac[(kt, ks)]
            cond = "kt == %d && ks == %d" % (kt, ks)
            if E.is_zero(e):
                body = "continue;"
            else:
                body = "v = W * (%s);" % text(e)
            lines.append(("if (%s) " if first else "else if (%s) ") % cond + "{ " + body + " }")
            first = False
    lines.append("else { continue; }")
    return lines


def emit_th_kernels(form):
    """HIP source of a Taylor-Hood form's element kernels. Returns (source, name, info)."""
    s, F, S, Sp, m = form.build()
    uses_x = any(_uses(e, s.x) for e in [v for r in F for v in r] + S + [Sp, m])
    quantities, stored = [], {}
    q = lambda e: _store(quantities, e, stored)  # noqa: E731
    res, jac = _blocks(s, F, S, Sp, q)
    main_q = list(quantities)
    if form.global_param is not None:
        gp = s.par[form.global_param]
        Fd = [[E.diff(v, gp) for v in r] for r in F]
        Sd = [E.diff(v, gp) for v in S]
        Spd = E.diff(Sp, gp)
        q2, st2 = [], {}
        qd = lambda e: _store(q2, e, st2)  # noqa: E731
        res_d, jac_d = _blocks(s, Fd, Sd, Spd, qd)
    else:
        q2 = []
        res_d = {0: E.ZERO, 1: E.ZERO, 2: E.ZERO}
        jac_d = {(a, b): E.ZERO for a in range(3) for b in range(3)}
    nqv = max(len(main_q), len(q2), 1)
    text = lambda e: E.CPrinter("T", prefix="b").expr_text(e)  # noqa: E731

    # Hessian-vector products: the Jacobian blocks of the directional derivative of (F, S,
    # S_p) in y (plain), and of the gradient of phi = sum_a F_a . grad y_a + S_a y_a + S_p y_p
    # (transposed: its Jacobian is sum_r y_r d2R_r / dx dx), point quantities after Q[nqv)
    def dirv(e):
        acc = E.ZERO
        for b in range(2):
            acc = acc + E.diff(e, s.u[b]) * s.yv[b]
            for d in range(2):
                acc = acc + E.diff(e, s.grad_u[b][d]) * s.gy[b][d]
        return acc + E.diff(e, s.p) * s.yp

    phi = Sp * s.yp
    for a_ in range(2):
        phi = phi + S[a_] * s.yv[a_]
        for d in range(2):
            phi = phi + F[a_][d] * s.gy[a_][d]
    hv_sets = []
    for Fh, Sh, Sph in (([[dirv(v) for v in r] for r in F], [dirv(v) for v in S], dirv(Sp)),
                        ([[E.diff(phi, s.grad_u[a_][d]) for d in range(2)] for a_ in range(2)],
                         [E.diff(phi, s.u[a_]) for a_ in range(2)], E.diff(phi, s.p))):
        qh, sth = [], {}

        def qstore(e, qh=qh, sth=sth):
            if e.op == "const":
                return e
            k = sth.get(e)
            if k is None:
                k = len(qh)
                sth[e] = k
                qh.append(e)
            return E.sym("Q[%d]" % (nqv + k))
        _, jac_h = _blocks(s, Fh, Sh, Sph, qstore)
        hv_sets.append((qh, jac_h))
    nqh = max(len(hv_sets[0][0]), len(hv_sets[1][0]), 1)

    def phase_h_lines(qs):
        return E.CPrinter("T", prefix="h").assign([("Q[%d]" % (nqv + k), e) for k, e in enumerate(qs)]) or ["(void)Q;"]

    def phase_a_lines(qs, prefix):
        return E.CPrinter("T", prefix=prefix).assign([("Q[%d]" % k, e) for k, e in enumerate(qs)])

    x_block = ""
    if uses_x:
        xl = []
        for a_ in range(2):
            xl.append("T x%d = T(0);" % a_)
            xl.append("#pragma unroll")
            xl.append("for (int l = 0; l < NN; l++) x%d = x%d + sX[le][l][%d] * P[l];" % (a_, a_, a_))
        x_block = _indent(xl, 8)
    # the mass coefficient at the point (TR only); in derivative mode its derivative in the
    # global parameter (the flag-2 pass then gives dM/dp)
    md = E.diff(m, s.par[form.global_param]) if form.global_param is not None else E.ZERO
    mlines = (E.CPrinter("T", prefix="m").assign([("Q[NQV + 2]", m)]) if not E.is_zero(m)
              else ["Q[NQV + 2] = T(0);"])
    mdlines = (E.CPrinter("T", prefix="md").assign([("Q[NQV + 2]", md)]) if not E.is_zero(md)
               else ["Q[NQV + 2] = T(0);"])
    mass_line = _indent(["if (a.deriv) {"] + ["    " + l for l in mdlines] + ["} else {"] +
                        ["    " + l for l in mlines] + ["}"], 12)
    src = KERNEL.format(
        name=form.name, nqv=nqv, nqh=nqh, x_block=x_block,
        phase_hv=_indent(phase_h_lines(hv_sets[0][0]), 16), phase_hvt=_indent(phase_h_lines(hv_sets[1][0]), 16),
        jac_hv=_indent(_jac_switch(hv_sets[0][1], text), 16), jac_hvt=_indent(_jac_switch(hv_sets[1][1], text), 16),
        phase_a=_indent(phase_a_lines(main_q, "a") or ["(void)Q;"], 12),
        phase_a_deriv=_indent(phase_a_lines(q2, "d") or ["(void)Q;"], 12),
        mass_line=mass_line,
        r0=text(res[0]), r1=text(res[1]), rp=text(res[2]),
        rd0=text(res_d[0]), rd1=text(res_d[1]), rdp=text(res_d[2]),
        jac=_indent(_jac_switch(jac, text), 16), jac_deriv=_indent(_jac_switch(jac_d, text), 16))
    info = {"kind": 3, "dim": 2, "nslot": NS, "nquantities": nqv, "symmetric": False,
            "has_param_derivative": form.global_param is not None, "layout": 1}
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
