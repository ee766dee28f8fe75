"""Oracle element loop, residual and CSR assembly (oomph vectors_of_pairs semantics).

Test infrastructure only (see ``oracle/__init__``).

* ``OracleProblem`` binds a mesh, the values per node, pinning and the weak form, and
  builds the per-element local equation tables the way oomph's
  ``FiniteElement::assign_nodal_local_eqn_numbers`` does (nodes in element order, values
  in value order, pinned values skipped) -- the element's local dof order that indexes the
  dense element Jacobian.
* ``get_residuals`` / ``get_jacobian`` restate oomph ``Problem::get_residuals``
  (problem.cc:3810-3900) and ``Problem::get_jacobian`` with the default
  ``vectors_of_pairs`` method (problem.cc:4038-4170, 5332-5685), done in ``oracle.c``.
"""
import ctypes
import os

import numpy as np

from . import tables as _tables

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

FORM_POISSON, FORM_BRATU, FORM_RD, FORM_NS = 1, 2, 3, 4

# Taylor-Hood slot table (node, value) in oomph local order; must match oracle.c.
TH_SLOTS = [(0, 0), (0, 1), (0, 2), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2), (3, 0), (3, 1),
            (4, 0), (4, 1), (5, 0), (5, 1), (6, 0), (6, 1), (6, 2), (7, 0), (7, 1), (8, 0),
            (8, 1), (8, 2)]


def lib():
    """The oracle's C library: the exact build (-ffp-contract=off, the checker), or with
    ORACLE_BUILD=native the build with the reference's compiler defaults (-O3
    -march=native, contraction on; ccompiler.py:213-221), used only to time the CPU
    baseline."""
    global _LIB
    if _LIB is None:
        native = os.environ.get("ORACLE_BUILD") == "native"
        path = os.path.join(_HERE, "_build", "liboracle_native.so" if native else "liboracle.so")
        if not os.path.exists(path):
            import subprocess
            subprocess.check_call(["make", "-s", "-C", _HERE] + (["native"] if native else []))
        L = ctypes.CDLL(path)
        dp = ctypes.POINTER(ctypes.c_double)
        lp = ctypes.POINTER(ctypes.c_long)
        L.oracle_elements.restype = ctypes.c_int
        L.oracle_elements.argtypes = [ctypes.c_int] * 6 + [dp] * 4 + [ctypes.c_int, dp, dp, dp,
                                                                       lp, lp, dp, ctypes.c_int,
                                                                       ctypes.c_int, ctypes.c_int,
                                                                       ctypes.c_int, dp,
                                                                       ctypes.c_double, dp, dp, dp]
        L.oracle_vectors_of_pairs.restype = ctypes.c_long
        L.oracle_vectors_of_pairs.argtypes = [ctypes.c_long, ctypes.c_int, ctypes.c_int, lp, lp,
                                              dp, dp, dp, lp, ctypes.POINTER(lp),
                                              ctypes.POINTER(dp)]
        L.oracle_free.argtypes = [ctypes.c_void_p]
        L.oracle_assemble_residual.restype = None
        L.oracle_assemble_residual.argtypes = [ctypes.c_int, ctypes.c_int, lp, lp, dp, dp]
        _LIB = L
    return _LIB


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double)) if a is not None else None


def _lp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_long))


class OracleProblem:
    """A steady problem on one mesh.

    mesh      -- ``oracle.meshes.OracleMesh``
    form      -- FORM_* id
    nvalue    -- values per node (1 scalar, 3 for Taylor-Hood [ux, uy, p])
    pinned    -- bool [nnode, nvalue]
    bc        -- float [nnode, nvalue] values of pinned entries
    params    -- form parameters (Poisson: 4 per Gaussian cx,cy,w,a; Bratu: lambda;
                 Navier-Stokes: mu, rho)
    aux       -- nodal auxiliary field (reaction-diffusion source f) or None
    faces     -- [nface, 2|3] boundary-face nodes (``faces.boundary_faces``) carrying the
                 flux face_alpha*u + face_beta (``oracle/faces.py``; scalars, or one value
                 per face), or None; the face elements follow the bulk elements in
                 element order
    face_expr -- (flux, dflux/du, dflux/dparam, d2flux/du dparam; the last two may be None)
                 callables (u, x0, x1, C) replacing
                 the flux alpha*u + beta of the faces by an expression, C = face_coef
                 [nface, ncoef] (``faces.expression_face_outputs``); with face_normal the
                 expressions also take the outward unit normal (n0, n1)
    enf_expr  -- (c, dc/du, dc/dp, d2c/du dp) of a constraint expression replacing
                 enf_alpha*u + enf_beta (functions of u, x0, x1, C = enf_coef [nenf, ncoef];
                 the last two may be None), or None
    enforced  -- [nenf, 2|3] boundary-face nodes of EnforcedBC(u=enf_alpha*u + enf_beta)
                 (``faces.enforced_face_outputs``), or None; the scalar forms then have
                 nvalue 2 (u, the multiplier); these faces follow the flux faces
    """

    def __init__(self, mesh, form, nvalue, pinned, bc, params, aux=None, faces=None, face_alpha=0.0,
                 face_beta=0.0, enforced=None, enf_alpha=1.0, enf_beta=0.0, face_expr=None, face_coef=None,
                 face_normal=False, enf_expr=None, enf_coef=None):
        from .meshes import number_equations
        self.mesh, self.form, self.nvalue = mesh, form, nvalue
        self.pinned = np.asarray(pinned, dtype=bool)
        self.bc = np.asarray(bc, dtype=np.float64)
        self.params = np.ascontiguousarray(params, dtype=np.float64)
        self.aux = None if aux is None else np.ascontiguousarray(aux, dtype=np.float64)
        self.eqn, self.ndof = number_equations(mesh.nnode, nvalue, self.pinned)
        self.tab = _tables.tables(mesh.kind)
        E = mesh.elements
        if form == FORM_NS:
            self.slots = TH_SLOTS
        else:
            self.slots = [(l, 0) for l in range(E.shape[1])]
        ns = len(self.slots)
        sn = np.array([s[0] for s in self.slots])
        sv = np.array([s[1] for s in self.slots])
        self.slot_node = E[:, sn]                     # [nel, nslot] global node
        self.slot_val = np.broadcast_to(sv, self.slot_node.shape)
        g = self.eqn[self.slot_node, self.slot_val]   # [nel, nslot] global eqn or -1
        self.slot_geqn = g
        self.leqn = np.full(g.shape, -1, dtype=np.int64)
        free = g >= 0
        self.leqn[free] = (np.cumsum(free, axis=1) - 1)[free]
        self.nvar = free.sum(axis=1).astype(np.int64)
        self.nslot = ns
        nenf_slots = 0 if enforced is None or not len(enforced) else 2 * np.asarray(enforced).shape[1]
        self.maxdof = max(ns, nenf_slots)
        self.geqn = np.full((E.shape[0], self.maxdof), -1, dtype=np.int64)
        rows = np.nonzero(free)
        self.geqn[rows[0], self.leqn[rows]] = g[rows]
        self.faces = None
        self.face_expr = face_expr
        self.face_normal = bool(face_normal)
        self.enf_expr = enf_expr
        self.enf_coef = None if enf_coef is None else np.asarray(enf_coef, dtype=np.float64)
        self.face_coef = None if face_coef is None else np.asarray(face_coef, dtype=np.float64)
        if faces is not None and len(faces):
            nf = len(faces)
            nc = 2 if form == FORM_NS else 1      # velocity components of a Taylor-Hood traction
            self.face_alpha = np.broadcast_to(np.asarray(face_alpha, dtype=np.float64), (nf,)).copy()
            fb = np.asarray(face_beta, dtype=np.float64)
            self.face_beta = np.broadcast_to(fb.reshape(-1, nc) if fb.ndim else fb, (nf, nc)).copy()
            if not (nvalue in (1, 2) and form != FORM_NS) and not (nvalue == 3 and form == FORM_NS):
                raise ValueError("boundary-face fluxes are built for the scalar forms and Taylor-Hood")
            self.faces = np.asarray(faces, dtype=np.int64)
            # face slots: (node l, component c) = l*nc + c
            fg = self.eqn[np.repeat(self.faces, nc, axis=1), np.tile(np.arange(nc), self.faces.shape[1])]
            ffree = fg >= 0
            self.face_leqn = np.full(fg.shape, -1, dtype=np.int64)
            self.face_leqn[ffree] = (np.cumsum(ffree, axis=1) - 1)[ffree]
            fnv = ffree.sum(axis=1).astype(np.int64)
            fgeqn = np.full((fg.shape[0], self.maxdof), -1, dtype=np.int64)
            r = np.nonzero(ffree)
            fgeqn[r[0], self.face_leqn[r]] = fg[r]
            self.nvar = np.concatenate([self.nvar, fnv])
            self.geqn = np.concatenate([self.geqn, fgeqn])
        self.enforced = None
        if enforced is not None and len(enforced):
            if nvalue != 2 or form == FORM_NS:
                raise ValueError("EnforcedBC multipliers: scalar forms with nvalue 2 (u, multiplier)")
            self.enforced = np.asarray(enforced, dtype=np.int64)
            ne, enn = self.enforced.shape
            self.enf_alpha = np.broadcast_to(np.asarray(enf_alpha, dtype=np.float64), (ne,)).copy()
            self.enf_beta = np.broadcast_to(np.asarray(enf_beta, dtype=np.float64), (ne,)).copy()
            # slots (node l, value c) = 2l + c: u and the multiplier
            eg = self.eqn[np.repeat(self.enforced, 2, axis=1), np.tile(np.arange(2), enn)]
            efree = eg >= 0
            self.enf_leqn = np.full(eg.shape, -1, dtype=np.int64)
            self.enf_leqn[efree] = (np.cumsum(efree, axis=1) - 1)[efree]
            egeqn = np.full((ne, self.maxdof), -1, dtype=np.int64)
            r = np.nonzero(efree)
            egeqn[r[0], self.enf_leqn[r]] = eg[r]
            self.nvar = np.concatenate([self.nvar, efree.sum(axis=1).astype(np.int64)])
            self.geqn = np.concatenate([self.geqn, egeqn])

    # -- time stepping (scalar forms) --------------------------------------------------
    # du/dt = tw[0] u + (tw[1] u(t-1) + tw[2] u(t-2)), nodal history combined per node;
    # pinned nodes keep their boundary value at every time level.
    time_weights = None
    hist = None

    def set_time(self, tw, h1, h2):
        """BDF weights (current, previous, second previous) and the previous and
        second-previous dof vectors; tw None returns to the steady problem."""
        if tw is None:
            self.time_weights, self.hist = None, None
        else:
            self.time_weights = np.asarray(tw, dtype=np.float64)
            self.hist = (np.asarray(h1, dtype=np.float64), np.asarray(h2, dtype=np.float64))

    def _history_slots(self):
        if self.time_weights is None:
            return None, 0.0
        tw = self.time_weights
        U1 = self.nodal_values(self.hist[0])
        U2 = self.nodal_values(self.hist[1])
        H = tw[1] * U1[self.slot_node, self.slot_val] + tw[2] * U2[self.slot_node, self.slot_val]
        return np.ascontiguousarray(H), float(tw[0])

    def nodal_values(self, x):
        U = self.bc.copy()
        m = self.eqn >= 0
        U[m] = x[self.eqn[m]]
        return U

    def element_outputs(self, x, flag, deriv=0, y=None):
        """Per-element R [nel, maxdof] and J [nel, maxdof, maxdof] in local numbering;
        with y (a dof vector) J holds the Hessian-vector product instead."""
        mesh, tab = self.mesh, self.tab
        U = self.nodal_values(x)
        nel, nn = mesh.elements.shape
        dim = mesh.coords.shape[1]
        X = np.ascontiguousarray(mesh.coords[mesh.elements])            # [nel, nn, dim]
        V = np.ascontiguousarray(U[self.slot_node, self.slot_val])       # [nel, nslot]
        aux = None
        if self.aux is not None:
            aux = np.ascontiguousarray(self.aux[mesh.elements])
        R = np.zeros((nel, self.maxdof))
        J = np.zeros((nel, self.maxdof, self.maxdof)) if flag else None
        H, tw0 = self._history_slots()
        Y = None
        if y is not None:
            Yn = np.zeros_like(self.bc)            # pinned values do not vary: Y = 0 there
            m = self.eqn >= 0
            Yn[m] = np.asarray(y, dtype=np.float64)[self.eqn[m]]
            Y = np.ascontiguousarray(Yn[self.slot_node, self.slot_val])
        psi1 = tab.get("psi1")
        psi1 = None if psi1 is None else np.ascontiguousarray(psi1)
        rc = lib().oracle_elements(
            self.form, nel, dim, nn, len(tab["w"]), 4 if psi1 is not None else 0,
            _dp(np.ascontiguousarray(tab["w"])), _dp(np.ascontiguousarray(tab["psi"])),
            _dp(np.ascontiguousarray(tab["dpsi"])), _dp(psi1), self.nslot, _dp(X), _dp(V),
            _dp(aux), _lp(np.ascontiguousarray(self.leqn)), _lp(self.nvar), _dp(self.params),
            len(self.params), int(flag), int(deriv), self.maxdof, _dp(H), tw0, _dp(Y), _dp(R), _dp(J))
        if rc != 0:
            raise RuntimeError("oracle_elements failed: %d" % rc)
        if self.faces is not None:
            R, J = self._append_faces(U, R, J, flag, deriv, y)
        if self.enforced is not None:
            R, J = self._append_enforced(U, R, J, flag, deriv, y)
        return R, J

    def _append_enforced(self, U, R, J, flag, deriv, y):
        """EnforcedBC face blocks after the flux faces, in the compact local numbering."""
        from .faces import enforced_face_outputs
        F = self.enforced
        nf, nn = F.shape
        quad = self.mesh.kind == "quad_q2"
        Xf = np.ascontiguousarray(self.mesh.coords[F][:, :, :2])
        if self.enf_expr is not None:
            # a constraint expression: its parameter derivatives stay in the constraint rows
            Rs, Js = enforced_face_outputs(Xf, np.ascontiguousarray(U[F, 0]), np.ascontiguousarray(U[F, 1]),
                                           None, None, 1 if flag == 1 else 0, quad=quad, cexpr=self.enf_expr,
                                           C=self.enf_coef, deriv=bool(deriv))
            if flag == 2 or y is not None:
                Js[:] = 0.0
        else:
            Rs, Js = enforced_face_outputs(Xf, np.ascontiguousarray(U[F, 0]), np.ascontiguousarray(U[F, 1]),
                                           self.enf_alpha, self.enf_beta, 1 if flag == 1 else 0, quad=quad)
            if deriv:
                Rs[:] = 0.0
            if deriv or flag == 2 or y is not None:
                Js[:] = 0.0
        md = self.maxdof
        Rf = np.zeros((nf, md))
        Jf = np.zeros((nf, md * md))
        nel = R.shape[0]
        for f in range(nf):
            loc = self.enf_leqn[f]
            fr = np.nonzero(loc >= 0)[0]
            nv = len(fr)
            Rf[f, loc[fr]] = Rs[f, fr]
            blk = np.zeros((nv, nv))
            blk[np.ix_(loc[fr], loc[fr])] = Js[f][np.ix_(fr, fr)]
            Jf[f, :nv * nv] = blk.ravel()
            assert nv == self.nvar[nel + f]
        R = np.concatenate([R, Rf])
        if J is not None:
            J = np.concatenate([J, Jf.reshape(nf, md, md)])
        return np.ascontiguousarray(R), None if J is None else np.ascontiguousarray(J)

    def _append_faces(self, U, R, J, flag, deriv, y):
        """Face element blocks after the bulk ones, in the same compact local numbering
        (residual stride maxdof, Jacobian block e at e*maxdof^2 with row stride nvar)."""
        from .faces import face_outputs, face_outputs_quad9
        F = self.faces
        nf, nn = F.shape
        nc = self.face_beta.shape[1]
        fl = 1 if flag == 1 else 0
        if self.face_expr is not None:
            from .faces import expression_face_outputs
            flux, dflux, dflux_dp, d2flux_dudp = self.face_expr
            Xf = np.ascontiguousarray(self.mesh.coords[F][:, :, :2])
            quad = self.mesh.kind == "quad_q2"
            Uf = np.ascontiguousarray(U[F, 0])
            sides = None
            if self.face_normal:
                from .faces import outward_sides
                sides = outward_sides(self.mesh.coords, self.mesh.elements, F)
            if deriv:
                # dR/dp and dJ/dp: the flux's parameter derivative and its u-derivative
                zero = lambda u, *args: np.zeros_like(u)  # noqa: E731
                f = dflux_dp if dflux_dp is not None else zero
                d = d2flux_dudp if d2flux_dudp is not None else zero
                Rs, Js = expression_face_outputs(Xf, Uf, self.face_coef, f, d, fl, quad=quad, sides=sides)
            else:
                Rs, Js = expression_face_outputs(Xf, Uf, self.face_coef, flux, dflux, fl, quad=quad, sides=sides)
            if flag == 2 or y is not None:
                Js[:] = 0.0
            deriv = False   # handled: the derivative entries are the expression's
        elif nn == 9:
            Xf = np.ascontiguousarray(self.mesh.coords[F][:, :, :3])
            Rs, Js = face_outputs_quad9(Xf, np.ascontiguousarray(U[F, 0]), self.face_alpha, self.face_beta[:, 0], fl)
        else:
            Xf = np.ascontiguousarray(self.mesh.coords[F][:, :, :2])
            quad = self.mesh.kind == "quad_q2"
            Rs = np.zeros((nf, nn * nc))
            Js = np.zeros((nf, nn * nc, nn * nc))
            for c in range(nc):
                # each velocity component's rows and columns: slots l*nc + c
                Rc, Jc = face_outputs(Xf, np.ascontiguousarray(U[F, c]), self.face_alpha, self.face_beta[:, c], fl,
                                      quad=quad)
                Rs[:, c::nc] = Rc
                Js[:, c::nc, c::nc] = Jc
        if deriv:
            Rs[:] = 0.0
        if deriv or flag == 2 or y is not None:
            Js[:] = 0.0
        md = self.maxdof
        Rf = np.zeros((nf, md))
        Jf = np.zeros((nf, md * md))
        nel = self.mesh.elements.shape[0]
        for f in range(nf):
            loc = self.face_leqn[f]
            fr = np.nonzero(loc >= 0)[0]
            nv = len(fr)
            Rf[f, loc[fr]] = Rs[f, fr]
            blk = np.zeros((nv, nv))
            blk[np.ix_(loc[fr], loc[fr])] = Js[f][np.ix_(fr, fr)]
            Jf[f, :nv * nv] = blk.ravel()
            assert nv == self.nvar[nel + f]
        R = np.concatenate([R, Rf])
        if J is not None:
            J = np.concatenate([J, Jf.reshape(nf, md, md)])
        return np.ascontiguousarray(R), None if J is None else np.ascontiguousarray(J)

    def get_residuals(self, x, deriv=0):
        """oomph Problem::get_residuals: residuals[eqn] += el_res[i] in element order."""
        R, _ = self.element_outputs(x, 0, deriv)
        res = np.zeros(self.ndof)
        lib().oracle_assemble_residual(R.shape[0], self.maxdof, _lp(self.nvar),
                                       _lp(np.ascontiguousarray(self.geqn)), _dp(R), _dp(res))
        return res

    def get_mass_matrix(self, x):
        """The mass matrix dR/d(du/dt) (flag 2 of the element functions) assembled like the
        Jacobian: residual, row_start, column_index (first-touch order), values."""
        return self.get_jacobian(x, flag=2)

    def get_hessian_vector_product(self, x, y, transposed=False):
        """sum_k d J / d u_k y_k assembled like the Jacobian (the reference's
        HessianVectorProduct with one vector, flag 1): residual, row_start, column_index,
        values. transposed: entry (i, k) = sum_j y_j d^2 R_j/(du_i du_k) (flag 4,
        SET_DIRECTIONAL_SYMMETRIC_HESSIAN_FROM_TRANSPOSED, src/jitbridge.h:678-691)."""
        return self.get_jacobian(x, flag=4 if transposed else 1, y=y)

    def get_jacobian(self, x, deriv=0, flag=1, y=None):
        """oomph Problem::get_jacobian (vectors_of_pairs). Returns residual, row_start,
        column_index (first-touch order), values."""
        # element Jacobians: block e starts at e*maxdof^2, rows have stride nvar[e]
        R, J = self.element_outputs(x, flag, deriv, y)
        nel = R.shape[0]
        res = np.zeros(self.ndof)
        row_start = np.zeros(self.ndof + 1, dtype=np.int64)
        colp = ctypes.POINTER(ctypes.c_long)()
        valp = ctypes.POINTER(ctypes.c_double)()
        L = lib()
        nnz = L.oracle_vectors_of_pairs(self.ndof, nel, self.maxdof, _lp(self.nvar),
                                        _lp(np.ascontiguousarray(self.geqn)), _dp(R),
                                        _dp(J), _dp(res), _lp(row_start),
                                        ctypes.byref(colp), ctypes.byref(valp))
        if nnz < 0:
            raise MemoryError("oracle_vectors_of_pairs")
        col = np.ctypeslib.as_array(colp, shape=(max(nnz, 1),))[:nnz].copy()
        val = np.ctypeslib.as_array(valp, shape=(max(nnz, 1),))[:nnz].copy()
        L.oracle_free(ctypes.cast(colp, ctypes.c_void_p))
        L.oracle_free(ctypes.cast(valp, ctypes.c_void_p))
        return res, row_start, col, val


def map_order_csr(row_start, col, val):
    """The rows of a vectors-of-pairs CSR with their columns in ascending order: the form in
    which Problem::sparse_assemble_row_or_column_compressed_base_problem returns the
    matrices of a multi-assembly (src/problem.cpp:2232-2262, one std::map<unsigned, double>
    per row; each value is the same element-order sum of its non-zero element
    contributions, only the order of a row's entries differs)."""
    row_start = np.asarray(row_start)
    col = np.asarray(col)
    val = np.asarray(val)
    oc = np.empty_like(col)
    ov = np.empty_like(val)
    for r in range(len(row_start) - 1):
        a, b = row_start[r], row_start[r + 1]
        order = np.argsort(col[a:b], kind="stable")
        oc[a:b] = col[a:b][order]
        ov[a:b] = val[a:b][order]
    return row_start.copy(), oc, ov
