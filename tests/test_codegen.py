"""The HIP code generator on the CPU.

* The expression IR's derivatives equal sympy's (an independent differentiation) at random
  points, for the forms' fluxes and sources and for composite expressions.
* The printed C evaluates the tree as built: printing then evaluating the C text with
  Python's float arithmetic gives the value ``expr.evaluate`` computes, for every output
  and temporary.
* The weak forms' Jacobian coefficients are what the reference's generic Jacobian needs
  (identity flux Jacobian for the Laplacian forms; the dF/du coupling of nonlinear
  diffusion) and the symmetry flags follow.
* Every generated kernel compiles with hiprtc for gfx950 inside the library (no GPU).
"""
import math
import re

import numpy as np
import pytest
import sympy as sp

from pyoomph_amd.codegen import expr as E
from pyoomph_amd.codegen import library, runtime, scalar


def to_sympy(e, env):
    memo = {}

    def go(x):
        if x in memo:
            return memo[x]
        op, a = x.op, x.args
        if op == "sym":
            r = env[x.val]
        elif op == "const":
            r = sp.Float(x.val)
        elif op == "add":
            r = go(a[0]) + go(a[1])
        elif op == "sub":
            r = go(a[0]) - go(a[1])
        elif op == "mul":
            r = go(a[0]) * go(a[1])
        elif op == "div":
            r = go(a[0]) / go(a[1])
        elif op == "neg":
            r = -go(a[0])
        elif op == "exp":
            r = sp.exp(go(a[0]))
        elif op == "pow":
            r = go(a[0]) ** x.val
        memo[x] = r
        return r

    return go(e)


FORMS = [library.poisson(3), library.bratu(), library.reaction_diffusion(), library.nonlinear_diffusion()]


@pytest.mark.parametrize("form", FORMS, ids=lambda f: f.name)
@pytest.mark.parametrize("dim", [2, 3])
def test_ir_derivatives_equal_sympy(form, dim):
    if form.name.startswith("gen_poisson") and dim == 3:
        pytest.skip("the Gaussian source is 2D")
    s, F, S, _ = form.build(dim)
    names = [s.u.val] + [g.val for g in s.grad_u] + [x.val for x in s.x] + [s.f.val] + [p.val for p in s.p]
    env = {n: sp.Symbol("v%d" % i) for i, n in enumerate(names)}
    rng = np.random.default_rng(3)
    vals = {env[n]: float(rng.uniform(0.2, 0.9)) for n in names}
    for e in F + [S]:
        se = to_sympy(e, env)
        for v in [s.u] + list(s.grad_u) + list(s.p):
            mine = float(to_sympy(E.diff(e, v), env).subs(vals))
            ref = float(sp.diff(se, env[v.val]).subs(vals))
            assert math.isclose(mine, ref, rel_tol=1e-12, abs_tol=1e-12), (form.name, v.val)


def test_composite_derivatives():
    u, v = E.sym("u"), E.sym("v")
    e = E.exp(u * v) / (1.0 + u ** 3) - (u - v * 2.0) * E.exp(-u)
    env = {"u": sp.Symbol("u"), "v": sp.Symbol("v")}
    se = to_sympy(e, env)
    pt = {env["u"]: 0.37, env["v"]: -1.3}
    for x in (u, v):
        assert math.isclose(float(to_sympy(E.diff(e, x), env).subs(pt)), float(sp.diff(se, env[x.val]).subs(pt)),
                            rel_tol=1e-13)
    # second derivative through the same rules
    d2 = E.diff(E.diff(e, u), v)
    assert math.isclose(float(to_sympy(d2, env).subs(pt)), float(sp.diff(se, env["u"], env["v"]).subs(pt)),
                        rel_tol=1e-12)


def test_printer_evaluates_the_tree_as_built():
    u, w = E.sym("u"), E.sym("w")
    outs = [("o0", (u * u) * u - w), ("o1", E.exp(-(u * u) / (w * w)) * 3.0 + (u * u)), ("o2", -(u * w))]
    p = E.CPrinter("T")
    lines = p.assign(outs)
    env = {"u": 0.731, "w": -1.37, "texp": math.exp, "T": float}
    for line in lines:
        m = re.match(r"(?:const T )?(\w+) = (.*);$", line)
        env[m.group(1)] = eval(m.group(2), {}, env)
    for name, e in outs:
        assert env[name] == E.evaluate(e, {"u": 0.731, "w": -1.37})
    # the shared u*u went into a temporary once
    assert sum(1 for ln in lines if ln.startswith("const T")) >= 1


def test_identities_keep_results_and_structure():
    u = E.sym("u")
    assert (u + 0.0) is u and (0.0 + u) is u and (u * 1.0) is u and E.is_zero(u * 0.0)
    assert (-(-u)) is u and (u / 1.0) is u
    # no reordering: u*3 and 3*u are different trees
    assert (u * 3.0) is not (3.0 * u)
    # u^3 prints as ((u*u)*u), its derivative as 3*(u*u)
    p = E.CPrinter("T")
    assert p.expr_text(u ** 3) == "((u * u) * u)"
    assert p.expr_text(E.diff(u ** 3, u)) == "(T(3.0) * (u * u))"


def test_jacobian_structure_and_symmetry():
    for f in (library.bratu(), library.reaction_diffusion(), library.poisson(2)):
        s, F, S, _ = f.build(2)
        K = [[E.diff(F[a], s.grad_u[b]) for b in range(2)] for a in range(2)]
        assert all(E.is_one(K[a][a]) and E.is_zero(K[a][1 - a]) for a in range(2))
        assert f.symmetric(2)
    nd = library.nonlinear_diffusion()
    s, F, S, _ = nd.build(2)
    assert not E.is_zero(E.diff(F[0], s.u))
    assert not nd.symmetric(2)


@pytest.mark.parametrize("form,kind", [(library.bratu(), "tri_p1"), (library.bratu(), "tri_p2"),
                                       (library.poisson(8), "tri_p2"), (library.reaction_diffusion(), "hex_q2"),
                                       (library.reaction_diffusion(), "quad_q2"),
                                       (library.nonlinear_diffusion(), "tri_p2")],
                         ids=lambda v: getattr(v, "name", v))
def test_generated_kernels_compile_with_hiprtc(form, kind):
    size = runtime.compile_form(form, kind, load=False)
    assert size > 1000


@pytest.mark.parametrize("form", [library.navier_stokes(), library.stokes_brinkman()], ids=lambda f: f.name)
def test_taylor_hood_kernels_compile_with_hiprtc(form):
    assert runtime.compile_form(form, "quad_q2", load=False) > 1000


def test_taylor_hood_form_jacobian_blocks():
    """The generated Navier-Stokes form's derivatives: the viscous block of the symmetric
    gradient (2 mu on the own component's direction, mu across), the convection block rho
    (u . grad psi_j + psi_j d_b u_a), the pressure column -phi d_a psi_i and the continuity
    row d_b psi_j phi -- checked against sympy at a random point."""
    from pyoomph_amd.codegen import mixed
    form = library.navier_stokes()
    s, F, S, Sp, m = form.build()
    names = ([s.u[0].val, s.u[1].val, s.p.val] + [g.val for r in s.grad_u for g in r] +
             [v.val for v in s.par])
    env = {n: sp.Symbol("v%d" % i) for i, n in enumerate(names)}
    rng = np.random.default_rng(5)
    vals = {env[n]: float(rng.uniform(0.2, 0.9)) for n in names}
    for e in [v for r in F for v in r] + S + [Sp]:
        se = to_sympy(e, env)
        for v in list(s.u) + [s.p] + [g for r in s.grad_u for g in r]:
            mine = float(to_sympy(E.diff(e, v), env).subs(vals))
            ref = float(sp.diff(se, env[v.val]).subs(vals))
            assert math.isclose(mine, ref, rel_tol=1e-12, abs_tol=1e-12)
    src, name, info = mixed.emit_th_kernels(form)
    assert info["layout"] == 1 and info["nslot"] == 22 and not info["symmetric"]
    # the residual entries in the hand-written kernel's operation order (element_kernels.h)
    assert "((T(par[0]) * (g00 + g00)) - pp)" in src
    assert "(T(par[1]) * ((u0 * g00) + (u1 * g01)))" in src


def test_emitted_bratu_matches_the_compiled_operation_order():
    """The generated Bratu residual / Jacobian entries are the compiled kernel's expressions
    (element_kernels.h): W*((g.dxi) + src*P_i) and W*((dxj.dxi) + (dsrc*P_j)*P_i), with
    src = dsrc = -(lambda*exp(u)) stored once."""
    src, _, _ = scalar.emit_kernels(library.bratu(), "tri_p2")
    assert "Q[2] = (-(T(par[0]) * texp(u)));" in src
    assert "W * ((((Q[0] * dxi[0]) + (Q[1] * dxi[1])) + (Q[2] * P[row])))" in src
    assert "W * ((((dxj[0] * dxi[0]) + (dxj[1] * dxi[1])) + ((Q[2] * P[j]) * P[row])))" in src
    assert "W * (((Q[0] * P[j]) * P[row]))" in src   # d/dlambda: -(exp(u)) psi_j psi_i


def test_face_forms_compile_and_differentiate():
    """Generated boundary-face forms: hiprtc compiles the kernels of every line-face family
    (no GPU needed), and the printed point quantities are the flux and its u-derivative."""
    from pyoomph_amd.codegen import expr as E
    from pyoomph_amd.codegen.face import FaceFluxForm, emit_face_kernels
    from pyoomph_amd.codegen.runtime import compile_face_form
    f = FaceFluxForm("t_face_cubic", lambda s: (s.c[0] * s.u ** 3 + s.c[1] * s.x[1]) - s.p[0] * E.exp(s.u),
                     ncoef=2, nparams=1, global_param=0)
    for kind in ("tri_p1", "tri_p2", "quad_q2"):
        assert compile_face_form(f, kind, load=False) > 0
    src, name, info = emit_face_kernels(f, "tri_p2")
    assert info["has_param_derivative"] and info["face_nodes"] == 3
    s, F = f.build()
    env = {"u": 0.7, "x0": 0.1, "x1": 0.4, "T(cf[0])": 2.0, "T(cf[1])": 0.3, "T(par[0])": 0.8}
    import math
    assert abs(E.evaluate(E.diff(F, s.u), env) - (2.0 * 3 * 0.7 ** 2 - 0.8 * math.exp(0.7))) < 1e-14
    assert abs(E.evaluate(E.diff(F, s.p[0]), env) + math.exp(0.7)) < 1e-14
    with pytest.raises(ValueError):
        emit_face_kernels(f, "hex_q2")


def test_constraint_and_normal_forms_compile_for_the_line_faces():
    """EnforcedBC constraint forms and flux forms reading the normal compile with hiprtc for
    the line faces of every family (no GPU); the normal adds the face's side as one more
    coefficient; bricks are refused."""
    from pyoomph_amd.codegen.face import ConstraintForm, FaceFluxForm
    from pyoomph_amd.codegen.runtime import compile_constraint_form, compile_face_form
    con = ConstraintForm("cg_con_test", lambda s: s.u ** 3 + s.c[0] * s.n[1] - s.p[0] * s.u, ncoef=1, nparams=1,
                         global_param=0)
    flux = FaceFluxForm("cg_nflux_test", lambda s: s.c[0] * s.n[0] * s.u, ncoef=1)
    assert con.uses_normal() and flux.uses_normal() and flux.device_ncoef() == 2
    for kind in ("tri_p1", "tri_p2", "quad_q2"):
        assert compile_constraint_form(con, kind, load=False) > 0
        assert compile_face_form(flux, kind, load=False) > 0
    with pytest.raises(ValueError):
        compile_constraint_form(con, "hex_q2", load=False)
