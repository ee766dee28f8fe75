"""Host-side mirror of pyoomph's ``Problem`` for the assembly-and-solve path.

A ``Problem`` holds one or more sub-problems (``MeshProblem``: one mesh, one set of
equations, Dirichlet conditions). pyoomph's own ``Problem`` with one bulk mesh is a
single sub-problem; several sub-problems are solved as one batch with independent
Newton iterations (the served / swept case).

Reference interfaces mirrored here:
* Newton knobs and defaults: ``newton_solver_tolerance`` 1e-8,
  ``max_newton_iterations`` 10, ``max_residuals`` 1e10,
  ``always_take_one_newton_step`` True (pyoomph/generic/problem.py:409-411, 536).
* ``Problem.solve()`` (problem.py:4721) -> steady Newton; raises ``NewtonSolverError``
  (a ``RuntimeError``) like oomph's ``NewtonSolverError`` (problem.cc:9283-9298).
* ``Problem.assemble_jacobian()`` (problem.py:711-717): residual vector and the CSR
  Jacobian in oomph's form (first-touch column order, exact zeros dropped, int32 /
  float64), as ``_assemble_residual_jacobian`` returns it (src/pybind/problem.cpp:572-602).
* ``Problem._assemble_multiassembly`` (src/pybind/problem.cpp:631-639) and
  ``MultiAssembleRequest`` (pyoomph/generic/bifurcation_tools.py:442-520): several
  residual-like vectors and matrices at the same dofs in one call.
* ``get_current_dofs`` / ``set_current_dofs``.
* Equation numbering: nodes in mesh order, values in value order, pinned values skipped
  (oomph mesh.cc:686-708, nodes.cc:896-927).
All numerical work runs in libpyoomph_hip.so on the GPU; this module only builds the
numbering and moves arrays.
"""
import ctypes

import numpy as np

from . import _lib
from .equations import Equations, NavierStokesEquations

__all__ = ["MeshProblem", "Problem", "NewtonSolverError", "MultiAssembleRequest"]


class NewtonSolverError(RuntimeError):
    """Raised when a Newton solve exceeds max_residuals or max_newton_iterations."""


class MeshProblem:
    """One mesh with its equations and Dirichlet conditions.

    dirichlet: list of (boundary names, {value_index: value}) applied in order (a node on
    several boundaries keeps the value of the last condition that names it), as
    successive ``DirichletBC(...) @ boundary`` do in pyoomph.
    flux: list of (boundary names, alpha, beta) -- ``NeumannBC(u=alpha*u + beta) @ boundary``
    for each name (pyoomph/meshes/bcs.py:71-89: the residual gains the face integral of
    (alpha*u + beta) * test), in order, on P1 / P2 triangle, Q2 quadrilateral and Q2 brick
    meshes of the scalar forms; for Taylor-Hood, beta = (beta_x, beta_y) and the condition is
    ``NeumannBC(velocity_x=alpha*u_x + beta_x, velocity_y=alpha*u_y + beta_y)``; one
    (names, alpha, beta) triple is accepted for a single condition; None: none.
    enforced: list of (boundary names, alpha, beta) -- ``EnforcedBC(u=alpha*u + beta) @
    boundary`` (pyoomph/meshes/bcs.py:129-253) for scalar forms on P1 / P2 triangle and Q2
    quadrilateral meshes: a Lagrange-multiplier field l on those boundaries, in u's space,
    held as value 1 of the nodes (``nvalue`` 2: u, l; numbered node by node after u, as
    oomph numbers a node's values), pinned to 0 where u is pinned and off the boundaries
    (EnforcedBC pins redundant multipliers, bcs.py:199-243); the residual gains
    weak(alpha*u + beta, l_test) + weak(l, u_test) on every face. The Jacobian becomes a
    saddle point: solve with GMRES and ILU(0).
    enforced_form: a generated constraint form (``codegen.face.ConstraintForm``) whose
    constraint expression replaces alpha*u + beta; enforced then lists (boundary names,
    coefficients) with the form's ``ncoef`` coefficients per condition.
    flux_form: a generated boundary-face form (``codegen.face.FaceFluxForm``) whose flux
    expression replaces alpha*u + beta on the flux faces; flux then lists (boundary names,
    coefficients) with the form's ``ncoef`` coefficients per condition (a form reading the
    normal gets each face's side appended).
    """

    def __init__(self, mesh, equations, dirichlet=(), pin_extra=(), flux=None, enforced=None, flux_form=None,
                 enforced_form=None):
        if not isinstance(equations, Equations):
            raise TypeError("equations must be an Equations instance")
        self.mesh = mesh
        self.equations = equations
        equations.check_mesh(mesh)
        nv = equations.nvalue
        if enforced_form is None and enforced is not None and len(enforced) == 3 and np.isscalar(enforced[1]):
            enforced = [enforced]
        if enforced and nv != 1:
            raise ValueError("EnforcedBC multipliers are built for the scalar forms")
        self.nvalue = nv + (1 if enforced else 0)
        self.pinned = np.zeros((mesh.nnode, self.nvalue), dtype=bool)
        self.values = np.zeros((mesh.nnode, self.nvalue), dtype=np.float64)
        for names, vals in dirichlet:
            m = mesh.on_boundary(list(names))
            for vi, v in vals.items():
                self.pinned[m, vi] = True
                self.values[m, vi] = v
        # values a node does not carry (e.g. pressure on Q2-only nodes) are pinned dummies
        # (BulkElementBase::pin_dummy_values, src/elements.cpp:2271-2330)
        dummy = equations.dummy_values(mesh)
        if dummy is not None:
            self.pinned |= dummy
        for node, vi, v in pin_extra:
            self.pinned[node, vi] = True
            self.values[node, vi] = v
        self.enforced, self.enforced_coef = None, None
        self.enforced_form = enforced_form
        if enforced:
            from .meshes import boundary_faces, face_sides
            faces, coef = [], []
            normal = enforced_form is not None and enforced_form.uses_normal()
            for cond in enforced:
                if enforced_form is None:
                    names, alpha, beta = cond
                    cf = [float(alpha), float(beta)]
                else:
                    names, cf = cond
                    cf = np.asarray(cf, dtype=np.float64).reshape(-1)
                    if cf.size != enforced_form.ncoef:
                        raise ValueError("the constraint form takes %d coefficients per condition"
                                         % enforced_form.ncoef)
                names = [names] if isinstance(names, str) else list(names)
                f, owner = boundary_faces(mesh, names, return_elements=True)
                faces.append(f)
                c = np.tile(cf, (len(f), 1))
                if normal:
                    c = np.concatenate([c, face_sides(mesh, f, owner)[:, None]], axis=1)
                coef.append(c)
            self.enforced = np.ascontiguousarray(np.concatenate(faces), dtype=np.int32)
            self.enforced_coef = np.ascontiguousarray(np.concatenate(coef).reshape(len(self.enforced), -1),
                                                      dtype=np.float64)
            on = np.zeros(mesh.nnode, dtype=bool)
            on[self.enforced.reshape(-1)] = True
            # the multiplier exists on the constrained boundaries only, and is redundant
            # (pinned to 0) where u itself is pinned
            self.pinned[:, 1] = ~on | self.pinned[:, 0]
            self.values[:, 1] = 0.0
        free = ~self.pinned.reshape(-1)
        eqn = np.full(free.shape, -1, dtype=np.int64)
        eqn[free] = np.arange(int(free.sum()))
        self.eqn = eqn.reshape(mesh.nnode, self.nvalue).astype(np.int32)
        self.ndof = int(free.sum())
        self.faces, self.face_coef = None, None
        self.flux_form = flux_form
        if flux_form is not None:
            from .meshes import boundary_faces, face_sides
            if nv != 1:
                raise ValueError("generated boundary-face forms are built for the scalar forms")
            normal = flux_form.uses_normal()
            faces, coef = [], []
            for names, c in flux:
                names = [names] if isinstance(names, str) else list(names)
                c = np.asarray(c, dtype=np.float64).reshape(-1)
                if c.size != flux_form.ncoef:
                    raise ValueError("the face form takes %d coefficients per condition" % flux_form.ncoef)
                f, owner = boundary_faces(mesh, names, return_elements=True)
                faces.append(f)
                cf = np.tile(c, (len(f), 1))
                if normal:
                    # the face's side for the outward normal, after the form's coefficients
                    cf = np.concatenate([cf, face_sides(mesh, f, owner)[:, None]], axis=1)
                coef.append(cf)
            self.faces = np.ascontiguousarray(np.concatenate(faces), dtype=np.int32)
            self.face_coef = np.ascontiguousarray(np.concatenate(coef).reshape(len(self.faces), -1),
                                                  dtype=np.float64)
        elif flux is not None:
            from .meshes import boundary_faces
            if nv not in (1, 3):
                raise ValueError("boundary-face fluxes are built for the scalar forms and Taylor-Hood")
            if len(flux) == 3 and np.isscalar(flux[1]):
                flux = [flux]
            faces, coef = [], []
            for names, alpha, beta in flux:
                names = [names] if isinstance(names, str) else list(names)
                f = boundary_faces(mesh, names)
                faces.append(f)
                # Taylor-Hood: beta = (beta_x, beta_y), the traction on the velocity components
                b = [float(beta)] if nv == 1 else [float(beta[0]), float(beta[1])]
                coef.append(np.tile([float(alpha)] + b, (len(f), 1)))
            self.faces = np.ascontiguousarray(np.concatenate(faces), dtype=np.int32)
            self.face_coef = np.ascontiguousarray(np.concatenate(coef), dtype=np.float64)

    def nodal_values(self, dofs):
        U = self.values.copy()
        m = self.eqn >= 0
        U[m] = dofs[self.eqn[m]]
        return U


class Problem:
    """Device-resident steady problem (one or more independent sub-problems)."""

    def __init__(self, parts, device="cuda:0"):
        import torch
        if isinstance(parts, MeshProblem):
            parts = [parts]
        self.parts = list(parts)
        if not self.parts:
            raise ValueError("no sub-problems")
        self.newton_solver_tolerance = 1e-8
        self.max_newton_iterations = 10
        self.max_residuals = 1e10
        self.always_take_one_newton_step = True
        self.newton_relaxation_factor = 1.0
        # > 0: Newton also stops only once max|dx| <= this * max|x| (FP32 Jacobians)
        self.newton_correction_tolerance = 0.0
        self.jacobian_precision = "f64"
        self.linear_precision = "f64"
        self.preconditioner = "ilu0"
        # "auto": conjugate gradients where every sub-problem's Jacobian is symmetric
        # positive definite (Poisson, Bratu, reaction-diffusion), GMRES otherwise
        self.krylov = "auto"
        self.gmres_restart = 50
        # ILU(0) row order: "auto" (multicolour when the natural order has more than 64
        # triangular-solve levels), "natural" or "multicolor"
        self.ilu_order = "auto"
        # multigrid: sub-problems on P2 / Q2 elements coarsen first to the linear functions
        # on their vertices (False: aggregation from the first level)
        self.amg_element_level = True
        self.gmres_max_iters = 5000
        self.gmres_rtol = 1e-10
        self.rtol_f32 = 1e-5
        self.max_refinements = 10
        self.refine_rtol = 1e-10
        self.torch = torch
        self.device = torch.device(device)
        self._handle = None
        self._keep = []
        self._build()

    # -- set-up ---------------------------------------------------------------------
    def _build(self):
        L = _lib.load()
        torch = self.torch
        descs = (_lib.SubproblemDesc * len(self.parts))()
        keep = []
        for i, mp in enumerate(self.parts):
            eq = mp.equations
            coords = np.ascontiguousarray(mp.mesh.coords, dtype=np.float64)
            conn = np.ascontiguousarray(mp.mesh.elements, dtype=np.int32)
            eqn = np.ascontiguousarray(mp.eqn, dtype=np.int32)
            pv = np.ascontiguousarray(mp.values, dtype=np.float64)
            par = np.ascontiguousarray(eq.params(), dtype=np.float64)
            aux = eq.aux(mp.mesh)
            aux = None if aux is None else np.ascontiguousarray(aux, dtype=np.float64)
            keep += [coords, conn, eqn, pv, par, aux]
            d = descs[i]
            d.form = eq.form
            d.elem_kind = _lib.ELEM_KIND[mp.mesh.kind]
            d.nnode = mp.mesh.nnode
            d.nelem = mp.mesh.nelement
            d.coords = coords.ctypes.data_as(_lib.c_dp)
            d.conn = conn.ctypes.data_as(_lib.c_i32p)
            d.nvalue = mp.nvalue
            d.eqn = eqn.ctypes.data_as(_lib.c_i32p)
            d.pinned_values = pv.ctypes.data_as(_lib.c_dp)
            d.ndof = mp.ndof
            d.params = par.ctypes.data_as(_lib.c_dp) if par.size else None
            d.nparams = int(par.size)
            d.aux = aux.ctypes.data_as(_lib.c_dp) if aux is not None else None
            if mp.faces is not None and len(mp.faces):
                fc = np.ascontiguousarray(mp.faces, dtype=np.int32)
                cf = np.ascontiguousarray(mp.face_coef, dtype=np.float64)
                keep += [fc, cf]
                d.nface = fc.shape[0]
                d.face_conn = fc.ctypes.data_as(_lib.c_i32p)
                d.face_coef = cf.ctypes.data_as(_lib.c_dp) if cf.size else None
                if mp.flux_form is not None:
                    from .codegen.runtime import compile_face_form
                    d.face_form = compile_face_form(mp.flux_form, mp.mesh.kind)
            if mp.enforced is not None and len(mp.enforced):
                keep += [mp.enforced, mp.enforced_coef]
                d.nenforced = mp.enforced.shape[0]
                d.enforced_conn = mp.enforced.ctypes.data_as(_lib.c_i32p)
                d.enforced_coef = mp.enforced_coef.ctypes.data_as(_lib.c_dp) if mp.enforced_coef.size else None
                if mp.enforced_form is not None:
                    from .codegen.runtime import compile_constraint_form
                    d.enforced_form = compile_constraint_form(mp.enforced_form, mp.mesh.kind)
        torch.cuda.set_device(self.device)
        # the problem runs on the stream current at creation; every host wait below is on
        # this stream only, so problems on other streams keep running
        self._stream = torch.cuda.current_stream(self.device)
        stream = self._stream.cuda_stream
        h = ctypes.c_void_p()
        _lib.check(L.pyhip_problem_create(descs, len(self.parts), ctypes.c_void_p(stream),
                                          ctypes.byref(h)))
        self._handle = h
        ndof, nnz, nc, nb = (ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64())
        _lib.check(L.pyhip_problem_info(h, ctypes.byref(ndof), ctypes.byref(nnz), ctypes.byref(nc),
                                        ctypes.byref(nb)))
        self.ndof, self.nnz, self.ncontrib = ndof.value, nnz.value, nc.value
        self.device_bytes = nb.value
        off = np.zeros(len(self.parts) + 1, dtype=np.int64)
        _lib.check(L.pyhip_problem_dof_offsets(h, off.ctypes.data_as(_lib.c_i64p)))
        self.dof_offsets = off
        with torch.cuda.stream(self._stream):
            self.dofs = torch.zeros(max(self.ndof, 1), dtype=torch.float64, device=self.device)
        self._stream.synchronize()

    def _host(self, t):
        """Copy a device tensor of this problem to a numpy array (waits on its stream)."""
        with self.torch.cuda.stream(self._stream):
            return t.cpu().numpy()

    def __del__(self):
        try:
            if self._handle is not None:
                _lib.load().pyhip_problem_destroy(self._handle)
                self._handle = None
        except Exception:
            pass

    # -- dofs -------------------------------------------------------------------------
    def get_current_dofs(self):
        return self._host(self.dofs[:self.ndof]).copy()

    def set_current_dofs(self, x):
        x = np.asarray(x, dtype=np.float64)
        if x.shape != (self.ndof,):
            raise ValueError("expected %d dofs" % self.ndof)
        with self.torch.cuda.stream(self._stream):
            self.dofs[:self.ndof].copy_(self.torch.from_numpy(x))
        self._stream.synchronize()

    def sub_dofs(self, i, x=None):
        x = self.get_current_dofs() if x is None else x
        return x[self.dof_offsets[i]:self.dof_offsets[i + 1]]

    # -- assembly ---------------------------------------------------------------------
    def get_residuals(self, dparameter=None):
        """Residual vector (oomph Problem::get_residuals); ``dparameter="lambda"`` gives
        dR/dlambda (get_derivative_wrt_global_parameter, src/problem.cpp:922)."""
        torch = self.torch
        with torch.cuda.stream(self._stream):
            res = torch.empty(max(self.ndof, 1), dtype=torch.float64, device=self.device)
        deriv = -1 if dparameter is None else self._param_index(dparameter)
        _lib.check(_lib.load().pyhip_assemble(self._handle, _lib.ptr(self.dofs), 0, deriv, _lib.F64,
                                               _lib.ptr(res), None))
        return self._host(res[:self.ndof])

    def _param_index(self, name):
        """The device index of a global parameter: every form has at most one (index 0)."""
        names = {n for mp in self.parts for n in getattr(mp.equations, "global_parameters", ())}
        if name not in names:
            raise ValueError("unknown global parameter %r" % name)
        return 0

    def assemble_device(self, precision="f64", dparameter=None):
        """Residual and column-sorted CSR values on the device (torch tensors)."""
        torch = self.torch
        prec = _lib.F64 if precision == "f64" else _lib.F32
        with torch.cuda.stream(self._stream):
            res = torch.empty(max(self.ndof, 1), dtype=torch.float64, device=self.device)
            vals = torch.empty(max(self.nnz, 1), dtype=torch.float64 if prec == _lib.F64 else torch.float32,
                               device=self.device)
        deriv = -1 if dparameter is None else self._param_index(dparameter)
        _lib.check(_lib.load().pyhip_assemble(self._handle, _lib.ptr(self.dofs), 1, deriv, prec,
                                               _lib.ptr(res), _lib.ptr(vals)))
        return res, vals

    def structural_pattern(self):
        rs = np.zeros(self.ndof + 1, dtype=np.int32)
        col = np.zeros(max(self.nnz, 1), dtype=np.int32)
        _lib.check(_lib.load().pyhip_pattern_copy(self._handle, rs.ctypes.data_as(_lib.c_i32p),
                                                   col.ctypes.data_as(_lib.c_i32p)))
        return rs, col[:self.nnz]

    def _set_solved_residual(self, name, raise_error=True):
        """The residual the following assemblies and solves use (the reference's
        Problem._set_solved_residual, src/problem.cpp:621-633): "" the default one, or a
        named residual table of the generated forms; returns whether any equations define it."""
        n = ctypes.c_int()
        _lib.check(_lib.load().pyhip_problem_set_solved_residual(self._handle, str(name).encode(), ctypes.byref(n)))
        if name and n.value == 0:
            _lib.load().pyhip_problem_set_solved_residual(self._handle, self._solved_residual.encode(), None)
            if raise_error:
                raise RuntimeError("Cannot activate the residual-Jacobian pair named '%s', since it is defined "
                                   "in no equations at all" % name)
            return False
        self._solved_residual = str(name)
        return True

    _solved_residual = ""

    def _get_solved_residual(self):
        return self._solved_residual

    def assemble_jacobian(self, dparameter=None, which_one=""):
        """(residual, scipy.sparse.csr_matrix) in the reference's CSR form; with
        ``dparameter`` (dR/dp, dJ/dp), dJ/dp's exact-zero element contributions dropped;
        ``which_one``: a named residual table (pyoomph's assemble_jacobian(which_one=...),
        problem.py:711-717), the solved residual being restored afterwards."""
        import scipy.sparse
        old = self._solved_residual
        if which_one != old:
            self._set_solved_residual(which_one)
        try:
            res, vals = self.assemble_device("f64", dparameter)
            rs, col, val = self._export(vals)
        finally:
            if which_one != old:
                self._set_solved_residual(old)
        R = self._host(res[:self.ndof])
        J = scipy.sparse.csr_matrix((val, col, rs), shape=(self.ndof, self.ndof))
        return R, J

    def _export(self, vals):
        torch = self.torch
        with torch.cuda.stream(self._stream):
            rs = torch.empty(self.ndof + 1, dtype=torch.int32, device=self.device)
            col = torch.empty(max(self.nnz, 1), dtype=torch.int32, device=self.device)
            val = torch.empty(max(self.nnz, 1), dtype=torch.float64, device=self.device)
        nnz = ctypes.c_int64()
        _lib.check(_lib.load().pyhip_export_reference_csr(self._handle, _lib.ptr(vals), _lib.ptr(rs),
                                                          _lib.ptr(col), _lib.ptr(val),
                                                          ctypes.byref(nnz)))
        n = nnz.value
        return self._host(rs), self._host(col[:n]), self._host(val[:n])

    # -- Newton -----------------------------------------------------------------------
    def newton_options(self):
        o = _lib.NewtonOpts()
        _lib.load().pyhip_newton_opts_default(ctypes.byref(o))
        o.tol = self.newton_solver_tolerance
        o.max_iter = self.max_newton_iterations
        o.max_residuals = self.max_residuals
        o.always_take_one_newton_step = int(self.always_take_one_newton_step)
        o.relaxation = self.newton_relaxation_factor
        o.jacobian_precision = _lib.F64 if self.jacobian_precision == "f64" else _lib.F32
        o.linear.precision = _lib.F64 if self.linear_precision == "f64" else _lib.F32
        o.linear.precond = {"ilu0": _lib.PRECOND_ILU0, "jacobi": _lib.PRECOND_JACOBI,
                            "amg": _lib.PRECOND_AMG, "block": _lib.PRECOND_BLOCK}[self.preconditioner]
        kry = self.krylov
        saddle = any(mp.enforced is not None for mp in self.parts)
        if kry == "auto":
            kry = "cg" if all(getattr(mp.equations, "symmetric", False) for mp in self.parts) and not saddle \
                else "gmres"
        if saddle and (kry == "cg" or self.preconditioner != "ilu0"):
            raise ValueError("EnforcedBC multipliers make the Jacobian a saddle point: "
                             "use krylov='gmres' (or 'auto') with preconditioner='ilu0'")
        o.linear.krylov = _lib.KRYLOV_CG if kry == "cg" else _lib.KRYLOV_GMRES
        o.linear.rtol_f32 = self.rtol_f32
        o.linear.restart = self.gmres_restart
        o.linear.max_iters = self.gmres_max_iters
        o.linear.rtol = self.gmres_rtol
        o.linear.max_refinements = self.max_refinements
        o.linear.refine_rtol = self.refine_rtol
        o.linear.ilu_order = {"auto": _lib.ILU_ORDER_AUTO, "natural": _lib.ILU_ORDER_NATURAL,
                              "multicolor": _lib.ILU_ORDER_MULTICOLOR}[self.ilu_order]
        o.linear.amg_element_level = 1 if self.amg_element_level else 0
        o.correction_tol = self.newton_correction_tolerance
        return o

    # -- time stepping ------------------------------------------------------------------
    # du/dt (scalar forms; Taylor-Hood velocities) with the reference's BDF weights; the
    # history vectors live on the device next to the dofs.
    time = 0.0
    time_steps_done = 0
    _hist = None
    _dt_prev = None

    @staticmethod
    def bdf_weights(dt, dt_prev=None):
        """Weights of (u, u(t-1), u(t-2)) in du/dt: BDF2 (MultiTimeStepper::set_weights,
        src/timestepper.cpp:32-59), or BDF1 when no step has been completed (the degraded
        BDF2 of prepare_shape_buffer_for_integration, src/elements.cpp:4609-4612)."""
        if dt_prev is None:
            return np.array([1.0 / dt, -1.0 / dt, 0.0])
        return np.array([1.0 / dt + 1.0 / (dt + dt_prev), -(dt + dt_prev) / (dt * dt_prev),
                         dt / ((dt + dt_prev) * dt_prev)])

    def set_initial_condition(self, x=None, t=0.0):
        """Current dofs (x, host array, or the current ones) become the initial condition:
        every history level holds them (impulsive start), time = t, no step done."""
        if x is not None:
            self.set_current_dofs(x)
        torch = self.torch
        with torch.cuda.stream(self._stream):
            self._hist = [self.dofs.clone(), self.dofs.clone()]
        self._stream.synchronize()
        self.time, self.time_steps_done, self._dt_prev = float(t), 0, None

    def set_time_state(self, weights, hist1, hist2):
        """Time derivative for the following assemblies (weights of u, u(t-1), u(t-2);
        hist1 / hist2 device tensors [ndof]); weights None returns to the steady forms."""
        L = _lib.load()
        if weights is None:
            _lib.check(L.pyhip_set_time_stepping(self._handle, 0, None, None, None))
            self._time_keep = None
            return
        w = np.ascontiguousarray(weights, dtype=np.float64)
        self._time_keep = (w, hist1, hist2)
        _lib.check(L.pyhip_set_time_stepping(self._handle, 1, w.ctypes.data_as(_lib.c_dp), _lib.ptr(hist1),
                                              _lib.ptr(hist2)))

    def unsteady_newton_solve(self, dt, raise_on_failure=True):
        """One time step of length dt for every sub-problem (oomph's
        unsteady_newton_solve): the history shifts by one level, time advances, and
        Newton starts from the current values with the BDF weights of this step."""
        if self._hist is None:
            self.set_initial_condition()
        torch = self.torch
        with torch.cuda.stream(self._stream):
            self._hist[1].copy_(self._hist[0])
            self._hist[0].copy_(self.dofs)
        w = self.bdf_weights(dt, None if self.time_steps_done == 0 else self._dt_prev)
        self.set_time_state(w, self._hist[0], self._hist[1])
        try:
            st = self.solve(raise_on_failure=raise_on_failure)
        finally:
            self.set_time_state(None, None, None)
        self.time_steps_done += 1
        self._dt_prev = dt
        self.time += dt
        return st

    def assemble_mass_matrix(self):
        """The mass matrix dR/d(du/dt) (the reference's flag-2 element output) in the
        reference's CSR form (first-touch columns, exact zeros dropped)."""
        import scipy.sparse
        torch = self.torch
        with torch.cuda.stream(self._stream):
            res = torch.empty(max(self.ndof, 1), dtype=torch.float64, device=self.device)
            vals = torch.empty(max(self.nnz, 1), dtype=torch.float64, device=self.device)
            h = self.dofs.clone()
        keep = getattr(self, "_time_keep", None)
        self.set_time_state(self.bdf_weights(1.0), h, h)
        try:
            _lib.check(_lib.load().pyhip_assemble(self._handle, _lib.ptr(self.dofs), 2, -1, _lib.F64,
                                                   _lib.ptr(res), _lib.ptr(vals)))
            rs, col, val = self._export(vals)
        finally:
            if keep is None:
                self.set_time_state(None, None, None)
            else:
                self.set_time_state(*keep)
        return scipy.sparse.csr_matrix((val, col, rs), shape=(self.ndof, self.ndof))

    def assemble_hessian_vector_product(self, y, transposed=False):
        """sum_k dJ/du_k y_k at the current dofs (the reference's HessianVectorProduct) in
        the reference's CSR form; y of shape [nvec, ndof] gives one matrix per vector.
        transposed: entry (i, k) = sum_j y_j d^2 R_j/(du_i du_k) (flag 4), in the
        multi-assembly's CSR form (columns ascending)."""
        y = np.asarray(y, dtype=np.float64)
        if y.ndim == 2:
            return [self.assemble_hessian_vector_product(v, transposed) for v in y]
        if transposed:
            return self.start_multiassembly().dJdU(y, transposed=True).assemble()[0]
        import scipy.sparse
        torch = self.torch
        with torch.cuda.stream(self._stream):
            res = torch.empty(max(self.ndof, 1), dtype=torch.float64, device=self.device)
            vals = torch.empty(max(self.nnz, 1), dtype=torch.float64, device=self.device)
            yd = torch.tensor(np.asarray(y, dtype=np.float64), device=self.device)
        _lib.check(_lib.load().pyhip_assemble_hessian_vector(self._handle, _lib.ptr(self.dofs), _lib.ptr(yd),
                                                              _lib.ptr(res), _lib.ptr(vals)))
        rs, col, val = self._export(vals)
        return scipy.sparse.csr_matrix((val, col, rs), shape=(self.ndof, self.ndof))

    # -- multi-assembly -------------------------------------------------------------------
    def start_multiassembly(self):
        """A request list run by ``assemble()`` (the reference's
        AugmentedAssemblyHandler.start_multiassembly, bifurcation_tools.py:579-580)."""
        return MultiAssembleRequest(self)

    def _assemble_multiassembly(self, what, contributions, params, hessian_vectors, hessian_vector_indices):
        """The reference's ``_assemble_multiassembly`` (src/pybind/problem.cpp:631-639):
        returns (ndof, data, csrdata, return_indices) with data = the vectors, then the
        values of each matrix; csrdata = row_start, column_index of each matrix;
        return_indices[i] >= 0 a vector, -1 - m the m-th matrix. Parameters and Hessian
        vector indices are taken in the order of the requests that need them
        (src/bifurcation.cpp:5354-5392)."""
        torch = self.torch
        what = list(what)
        contributions = list(contributions)
        if len(what) != len(contributions):
            raise RuntimeError("Number of what and contributions must match")
        params = list(params)
        hidx = list(hessian_vector_indices)
        reqs = (_lib.MultiRequest * max(len(what), 1))()
        cnames = [str(c).encode() for c in contributions]
        pi = hi = 0
        for i, w in enumerate(what):
            if w not in _lib.MULTI_WHAT:
                raise RuntimeError("Unknown what type '%s'" % w)
            kind, tr = _lib.MULTI_WHAT[w]
            reqs[i].what, reqs[i].transposed, reqs[i].param, reqs[i].hvec = kind, tr, -1, -1
            reqs[i].contribution = cnames[i]
            if kind in (3, 4, 5):
                if pi >= len(params):
                    raise RuntimeError("You have not provided enough parameters for the what type '%s'" % w)
                reqs[i].param = self._param_index(params[pi])
                pi += 1
            if kind in (6, 7):
                if hi >= len(hidx):
                    raise RuntimeError("You have not provided enough hessian vector indices for the what type '%s'" % w)
                reqs[i].hvec = int(hidx[hi])
                hi += 1
        if pi != len(params):
            raise RuntimeError("You have provided too many parameters")
        if hi != len(hidx):
            raise RuntimeError("You have provided too many hessian vector indices")
        L = _lib.load()
        n = len(what)
        ret = (ctypes.c_int * max(n, 1))()
        nv, nm = ctypes.c_int(), ctypes.c_int()
        nh = len(hessian_vectors)
        _lib.check(L.pyhip_multi_assembly_plan(n, reqs, nh, ret, ctypes.byref(nv), ctypes.byref(nm)))
        nv, nm = nv.value, nm.value
        nd = self.ndof
        with torch.cuda.stream(self._stream):
            H = None
            if nh:
                H = torch.tensor(np.stack([np.asarray(v, dtype=np.float64).reshape(nd) for v in hessian_vectors]),
                                 device=self.device)
            vecs = [torch.empty(max(nd, 1), dtype=torch.float64, device=self.device) for _ in range(nv)]
            rss = [torch.empty(nd + 1, dtype=torch.int32, device=self.device) for _ in range(nm)]
            cols = [torch.empty(max(self.nnz, 1), dtype=torch.int32, device=self.device) for _ in range(nm)]
            vals = [torch.empty(max(self.nnz, 1), dtype=torch.float64, device=self.device) for _ in range(nm)]

        def parr(ts):
            a = (ctypes.c_void_p * max(len(ts), 1))()
            for k, t in enumerate(ts):
                a[k] = t.data_ptr()
            return a

        nnz = (ctypes.c_int64 * max(nm, 1))()
        _lib.check(L.pyhip_assemble_multi(self._handle, _lib.ptr(self.dofs), n, reqs,
                                          _lib.ptr(H), nh, parr(vecs), parr(rss), parr(cols), parr(vals), nnz))
        data = [self._host(v[:nd]) for v in vecs]
        csrdata = []
        for m in range(nm):
            k = nnz[m]
            data.append(self._host(vals[m][:k]))
            csrdata.append(self._host(rss[m]))
            csrdata.append(self._host(cols[m][:k]))
        return nd, data, csrdata, [ret[i] for i in range(n)]

    # -- arclength continuation -----------------------------------------------------------
    _arc = None

    def reset_arc_length_parameters(self):
        """Theta_squared = Parameter_derivative = 1, no step taken, zero dof derivative."""
        self._arc = (_lib.ArclengthState * len(self.parts))()
        for a in self._arc:
            a.theta_squared, a.parameter_derivative, a.step_taken = 1.0, 1.0, 0

    def arclength_continuation(self, parameter, ds, max_ds=None, active=None):
        """One pseudo-arclength step in the global parameter (the Bratu "lambda") for
        every sub-problem (or those with active[i]), each with its own step ds (scalar or
        per sub-problem) and continuation state: pyoomph's Problem.arclength_continuation
        (pyoomph/generic/problem.py:3530-3600; max_ds clamps the step as there) over
        oomph's arc_length_step_solve. Returns the suggested next steps (array). Use
        GMRES: the Jacobian is indefinite past a fold."""
        self._param_index(parameter)
        ns = len(self.parts)
        if self._arc is None:
            self.reset_arc_length_parameters()
        ds = np.broadcast_to(np.asarray(ds, dtype=np.float64), (ns,)).copy()
        if max_ds is not None:
            ds = np.minimum(np.abs(ds), abs(float(max_ds))) * np.where(ds > 0, 1.0, -1.0)
        act = np.ones(ns, dtype=np.int32) if active is None else np.asarray(active, dtype=np.int32).copy()
        nxt = np.zeros(ns)
        o = self.newton_options()
        rc = _lib.load().pyhip_arclength_step(self._handle, _lib.ptr(self.dofs), 0, ctypes.byref(o),
                                               act.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                               ds.ctypes.data_as(_lib.c_dp), self._arc,
                                               nxt.ctypes.data_as(_lib.c_dp))
        self._stream.synchronize()
        if rc == _lib.PYHIP_ERR_NEWTON:
            raise NewtonSolverError(_lib.load().pyhip_last_error().decode())
        _lib.check(rc)
        return nxt

    def set_global_parameter(self, sub, value):
        """Sets the continued parameter (the Bratu lambda) of one sub-problem."""
        par = np.array([value], dtype=np.float64)
        _lib.check(_lib.load().pyhip_problem_set_params(self._handle, int(sub), par.ctypes.data_as(_lib.c_dp), 1))

    def set_sub_data(self, sub, params=None, aux=None):
        """Replaces one sub-problem's parameters and / or auxiliary nodal field (a new
        request in a serving slot on the same mesh)."""
        L = _lib.load()
        if params is not None:
            par = np.ascontiguousarray(params, dtype=np.float64).reshape(-1)
            _lib.check(L.pyhip_problem_set_params(self._handle, int(sub), par.ctypes.data_as(_lib.c_dp),
                                                  int(par.size)))
        if aux is not None:
            a = np.ascontiguousarray(aux, dtype=np.float64).reshape(-1)
            _lib.check(L.pyhip_problem_set_aux(self._handle, int(sub), a.ctypes.data_as(_lib.c_dp), int(a.size)))

    def set_stream(self, stream):
        """Moves the problem (its library calls and its tensors' copies) to a torch stream
        of the same device; the work queued on the old stream completes first."""
        if stream is None or stream == self._stream:
            return
        _lib.check(_lib.load().pyhip_problem_set_stream(self._handle, ctypes.c_void_p(stream.cuda_stream)))
        self._stream = stream

    def current_device_bytes(self):
        """Device bytes held now (mesh, pattern, element buffers, solver workspace)."""
        nb = ctypes.c_int64()
        _lib.check(_lib.load().pyhip_problem_info(self._handle, None, None, None, ctypes.byref(nb)))
        return nb.value

    def newton_step(self, mask, state, opts=None):
        """One Newton iteration for the sub-problems with mask != 0 that have neither
        converged nor failed (``pyhip_newton_step``): ``state`` is a ctypes array of
        ``_lib.NewtonStepState`` [nsub] carried between calls, newton_iters = 0 starting a
        sub-problem's solve from its current dofs."""
        mask = np.ascontiguousarray(mask, dtype=np.int32)
        if mask.shape != (len(self.parts),) or len(state) != len(self.parts):
            raise ValueError("mask and state need one entry per sub-problem")
        o = self.newton_options() if opts is None else opts
        _lib.check(_lib.load().pyhip_newton_step(self._handle, _lib.ptr(self.dofs), ctypes.byref(o),
                                                 mask.ctypes.data_as(_lib.c_i32p), state))

    def global_parameters(self):
        """Current value of the continued parameter per sub-problem (after a step)."""
        if self._arc is None:
            return None
        return np.array([a.parameter for a in self._arc])

    def arclength_stats(self):
        """Per sub-problem, the last step's Newton iterations, rejected tries, continuation
        state, whether its linear solve failed and whether it crossed a fold (turning_point:
        Parameter_derivative changed sign, det J with it)."""
        return [{"newton_iters": a.newton_iters, "rejected": a.rejected, "theta_squared": a.theta_squared,
                 "parameter_derivative": a.parameter_derivative, "linear_failed": bool(a.linear_failed),
                 "turning_point": bool(a.turning_point)}
                for a in self._arc]

    def solve(self, raise_on_failure=True):
        """Steady Newton solve of every sub-problem; returns per-sub-problem stats."""
        o = self.newton_options()
        stats = (_lib.NewtonStats * len(self.parts))()
        rc = _lib.load().pyhip_newton_solve(self._handle, _lib.ptr(self.dofs), ctypes.byref(o), stats)
        self._stream.synchronize()
        out = [{"newton_iters": s.newton_iters, "linear_iters": s.linear_iters,
                "linear_failures": s.linear_failures,
                "converged": bool(s.converged), "final_maxres": s.final_maxres,
                "history": list(s.history[:s.nhistory]), "done_ms": s.done_ms} for s in stats]
        self.last_stats = out
        if rc == _lib.PYHIP_ERR_NEWTON:
            if raise_on_failure:
                raise NewtonSolverError(_lib.load().pyhip_last_error().decode())
            return out
        _lib.check(rc)
        return out


class MultiAssembleRequest:
    """The reference's MultiAssembleRequest (pyoomph/generic/bifurcation_tools.py:442-520) on
    a device ``Problem``: R, J, M, dRdp, dJdp, dMdp, dJdU, dMdU append requests (returning
    the request, so calls chain), ``assemble()`` runs them in one call and returns numpy
    vectors and scipy CSR matrices in request order. A Hessian vector passed twice (the same
    object) is sent once."""

    def __init__(self, problem):
        self._what = []
        self._contributions = []
        self._hessian_vectors = []
        self._hessian_vector_indices = []
        self._parameters = []
        self.problem = problem

    def _resolve_hessian_vector_index(self, V):
        for i, w in enumerate(self._hessian_vectors):
            if V is w:
                return i
        self._hessian_vectors.append(V)
        return len(self._hessian_vectors) - 1

    def _add(self, what, contribution):
        self._what.append(what)
        self._contributions.append(contribution)
        return self

    def R(self, contribution=""):
        return self._add("residuals", contribution)

    def J(self, contribution=""):
        return self._add("jacobian", contribution)

    def M(self, contribution=""):
        return self._add("mass_matrix", contribution)

    def dRdp(self, parameter, contribution=""):
        self._parameters.append(parameter)
        return self._add("dresiduals_dparameter", contribution)

    def dJdp(self, parameter, contribution=""):
        self._parameters.append(parameter)
        return self._add("djacobian_dparameter", contribution)

    def dMdp(self, parameter, contribution=""):
        self._parameters.append(parameter)
        return self._add("dmass_matrix_dparameter", contribution)

    def dJdU(self, vector, contribution="", transposed=False):
        self._hessian_vector_indices.append(self._resolve_hessian_vector_index(vector))
        return self._add("hessian_vector_product_transposed" if transposed else "hessian_vector_product",
                         contribution)

    def dMdU(self, vector, contribution="", transposed=False):
        self._hessian_vector_indices.append(self._resolve_hessian_vector_index(vector))
        return self._add("mass_matrix_hessian_vector_product_transposed" if transposed
                         else "mass_matrix_hessian_vector_product", contribution)

    def assemble(self):
        import scipy.sparse
        n, vectors, csrdatas, return_indices = self.problem._assemble_multiassembly(
            self._what, self._contributions, self._parameters, self._hessian_vectors,
            self._hessian_vector_indices)
        nmatrix = len(csrdatas) // 2
        nvectors = len(vectors) - nmatrix
        matrices = [scipy.sparse.csr_matrix((vectors[nvectors + i], csrdatas[2 * i + 1], csrdatas[2 * i]),
                                            shape=(n, n)) for i in range(nmatrix)]
        return [vectors[r] if r >= 0 else matrices[-(r + 1)] for r in return_indices]
