/*
 * pyoomph HIP assembly-and-solve library: C ABI.
 *
 * Plain pointers and sizes only. "dev" pointers are device (HBM) addresses, everything
 * else is host memory. Every entry point returns PYHIP_OK (0) or a negative status;
 * pyhip_last_error() gives the message of the last failure on the calling thread (the
 * reference throws std::runtime_error via throw_runtime_error, src/exception.cpp, which
 * pybind turns into RuntimeError; the Python mirror raises RuntimeError on a negative
 * status).
 *
 * Which reference interface each entry point replaces is cited beside it.
 */
#ifndef PYOOMPH_HIP_H
#define PYOOMPH_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---------------------------------------------------------------- */
#define PYHIP_OK 0
#define PYHIP_ERR_INVALID (-1)       /* bad argument / inconsistent sizes              */
#define PYHIP_ERR_HIP (-2)           /* HIP runtime error                              */
#define PYHIP_ERR_NOMEM (-3)         /* allocation failed or over the HBM budget        */
#define PYHIP_ERR_NOT_CONVERGED (-4) /* Krylov solver did not reach its tolerance       */
#define PYHIP_ERR_BREAKDOWN (-5)     /* zero pivot / non-finite value in a factorization */
#define PYHIP_ERR_NEWTON (-6)        /* NewtonSolverError (oomph problem.cc:9283-9298)  */

/* ---- element families and weak forms ---------------------------------------------- */
#define PYHIP_ELEM_TRI_P1 1  /* BulkElementTri2dC1   (src/elements.hpp:899-942)   */
#define PYHIP_ELEM_TRI_P2 2  /* BulkElementTri2dC2   (src/elements.hpp:990-1044)  */
#define PYHIP_ELEM_QUAD_Q2 3 /* BulkElementQuad2dC2  (src/elements.hpp:831-897)   */
#define PYHIP_ELEM_HEX_Q2 4  /* BulkElementBrick3dC2 (src/elements.hpp:1143-1213) */

#define PYHIP_FORM_POISSON 1            /* pyoomph/equations/poisson.py:70-74           */
#define PYHIP_FORM_BRATU 2              /* -lap u = lambda e^u, global parameter lambda */
#define PYHIP_FORM_REACTION_DIFFUSION 3 /* -lap u + u^3 = f, f a nodal field            */
#define PYHIP_FORM_NAVIER_STOKES_TH 4   /* pyoomph/equations/navier_stokes.py:255-347,
                                           481-503, Taylor-Hood Q2/Q1, steady           */

#define PYHIP_F64 0
#define PYHIP_F32 1

const char *pyhip_last_error(void);
int pyhip_version(void);
int pyhip_device_count(int *count);

/* ---- element table ---------------------------------------------------------------------
 * Replaces JIT_ELEMENT_init(JITFuncSpec_Table_FiniteElement_t*) (src/jitbridge.h:329-499,
 * emitted by src/codegen.cpp:6398-7090): fills the description of one weak form on one
 * element family. Residual 0 is the steady residual ("ResidualAndJacobianSteady",
 * codegen.cpp:4718-4730); has_param_derivative[p] marks dResidual/dParameter_p
 * (codegen.cpp:4733-4752). The functions themselves are the batched device entry points
 * below, not per-element function pointers. */
typedef struct {
    int form, elem_kind;
    int num_res_jacs;                  /* num_res_jacs                                */
    const char *res_jac_names[2];      /* res_jac_names                               */
    int nodal_values;                  /* values per node (numfields_C2+numfields_C1)  */
    int nslot;                         /* element dofs incl. pinned ones              */
    int nnode;                         /* nodes per element                           */
    int dim;                           /* nodal dimension                             */
    int integration_order;             /* integration_order (0 = element default)     */
    int nqp;                           /* points of the scheme that is used            */
    int numglobal_params;              /* numglobal_params                            */
    const char *global_param_names[2];
    int has_param_derivative[2];       /* ParameterDerivative[0][p] != NULL           */
    int moving_nodes;                  /* always 0 here                               */
} pyhip_element_table;
int pyhip_element_init(int form, int elem_kind, pyhip_element_table *table);

/* ---- problems ----------------------------------------------------------------------------
 * A problem is one or more independent sub-problems (each: one mesh, one weak form, its
 * own parameters) held device-resident with a block-diagonal sparsity pattern. One
 * sub-problem is pyoomph's Problem with one bulk mesh; several are the batched / served
 * case. Numbering, pinning and boundary values come from the host (the reference numbers
 * equations in oomph Mesh::assign_global_eqn_numbers, mesh.cc:686-708). */
typedef struct {
    int form, elem_kind;
    int64_t nnode, nelem;
    const double *coords;        /* [nnode][dim]                                        */
    const int32_t *conn;         /* [nelem][nnode_per_element], oomph local node order  */
    int nvalue;                  /* values per node                                     */
    const int32_t *eqn;          /* [nnode][nvalue] sub-problem equation, -1 pinned     */
    const double *pinned_values; /* [nnode][nvalue] value used where eqn == -1          */
    int64_t ndof;
    const double *params;        /* Poisson: n, then (cx, cy, w, a) per Gaussian source;
                                    Bratu: lambda; Navier-Stokes: mu, rho               */
    int nparams;
    const double *aux;           /* [nnode] auxiliary nodal field, or NULL              */
    /* Fluxes on boundary faces: pyoomph's NeumannBC(u=flux) @ boundary (an
     * InterfaceEquations residual weak(flux, testfunction(u)), pyoomph/meshes/bcs.py:71-89,
     * assembled by interface elements built on the bulk elements' faces,
     * src/mesh.cpp:615-660) with flux = alpha*u + beta per face (Neumann: alpha = 0;
     * Robin: alpha != 0), so several conditions on several boundaries of one mesh are
     * one face list. Scalar forms on P1 / P2 triangles (faces: TElement<1,2> /
     * TElement<1,3> with TGauss<1,2> / TGauss<1,3>, src/elements.hpp:676-730), Q2
     * quadrilaterals (QElement<1,3> with Gauss<1,3>) and Q2 bricks (QElement<2,3> with
     * Gauss<2,3>); for Taylor-Hood the traction NeumannBC(velocity_x=alpha*u_x + beta_x,
     * velocity_y=alpha*u_y + beta_y) on Q2 quad faces. The faces follow the bulk elements
     * in element order. nface = 0: none. */
    int64_t nface;
    const int32_t *face_conn;    /* [nface][2] (P1: end, end), [nface][3] (P2, Q2 quads:
                                    end, midside, end) or [nface][9] (Q2 bricks: the
                                    face's 3x3 node lattice) mesh nodes of each face     */
    const double *face_coef;     /* [nface][2] alpha, beta of each face's flux (Taylor-
                                    Hood: [nface][3] alpha, beta_x, beta_y)               */
    /* Constraints enforced by Lagrange multipliers: pyoomph's EnforcedBC(u=constraint) @
     * boundary (pyoomph/meshes/bcs.py:129-253) with constraint = alpha*u + beta per face,
     * for the scalar forms on P1 / P2 triangles and Q2 quadrilaterals. The multiplier field
     * l lives on the boundary in u's space as value 1 of the nodes (nvalue = 2: u, l; eqn
     * and pinned_values hold both, l pinned to 0 where u is pinned and on nodes off the
     * boundary, as EnforcedBC pins redundant multipliers, bcs.py:199-243). Each face adds
     * the residuals weak(constraint, l_test) and weak(l, u_test) (bcs.py:185-189):
     * R[u_i] += int l psi_i ds, R[l_i] += int (alpha u + beta) psi_i ds, on the face
     * elements of the flux faces above, slots (u, l) per face node. The Jacobian is a
     * saddle point (zero l-l block): GMRES with ILU(0), the multipliers ordered last.
     * The faces follow the flux faces in element order. nenforced = 0: none. */
    int64_t nenforced;
    const int32_t *enforced_conn; /* [nenforced][2|3] mesh nodes of each face          */
    const double *enforced_coef;  /* [nenforced][2] alpha, beta of each face's constraint */
    /* 0: the faces above carry the built-in flux alpha*u + beta; else the id of a generated
     * boundary-face form (pyhip_codegen_set_face_form, scalar forms on P1 / P2 triangles
     * and Q2 quadrilaterals) whose flux expression replaces it, face_coef then holding its
     * ncoef coefficients per face */
    int face_form;
    /* 0: the EnforcedBC faces carry the built-in constraint alpha*u + beta; else the id of
     * a generated constraint form (pyhip_codegen_set_constraint_form, same element
     * families) whose constraint expression c replaces it, enforced_coef then holding its
     * ncoef coefficients per face: R[u_i] += int l psi_i ds, R[l_i] += int c psi_i ds */
    int enforced_form;
} pyhip_subproblem_desc;

typedef struct pyhip_problem pyhip_problem;

/* Builds the sparsity pattern (host), uploads mesh, maps and pattern (stream: hipStream_t
 * or NULL for the default stream). */
int pyhip_problem_create(const pyhip_subproblem_desc *subs, int nsub, void *stream,
                         pyhip_problem **out);
int pyhip_problem_destroy(pyhip_problem *p);
/* ndof, nnz of the structural (column-sorted) pattern, number of element contributions,
 * bytes of device memory held (problem data plus the solver workspace allocated so far). */
int pyhip_problem_info(const pyhip_problem *p, int64_t *ndof, int64_t *nnz,
                       int64_t *ncontrib, int64_t *device_bytes);
/* dof offset of each sub-problem, host [nsub+1] */
int pyhip_problem_dof_offsets(const pyhip_problem *p, int64_t *offsets);
int pyhip_problem_set_params(pyhip_problem *p, int sub, const double *params, int nparams);
/* The residual the following assemblies and solves use: "" the default one, or a named
 * residual table of the generated forms (pyhip_codegen_add_residual_table). Replaces
 * Problem::_set_solved_residual (src/problem.cpp:621-633) -> the element codes'
 * _set_solved_residual (:186-200): element groups without that table contribute nothing
 * (the compiled forms and the boundary faces have the default residual only).
 * *ngroups receives the number of bulk element groups that have it (the reference throws
 * when none has it; pyoomph_amd.problem.Problem._set_solved_residual does). */
int pyhip_problem_set_solved_residual(pyhip_problem *p, const char *name, int *ngroups);
/* Column-sorted structural pattern to host buffers [ndof+1], [nnz]. */
int pyhip_pattern_copy(const pyhip_problem *p, int32_t *row_start, int32_t *col_index);

/* ---- assembly ----------------------------------------------------------------------------
 * Replaces oomph Problem::get_residuals (problem.cc:3810-3900) and Problem::get_jacobian
 * with the vectors_of_pairs method (problem.cc:4038-4170, 5332-5685) for elements whose
 * functions come from the JIT table (src/elements.cpp:5054-5127).
 *   dofs_dev      [ndof] double
 *   flag          0: residual only, 1: residual and Jacobian, 2: residual and mass
 *                 matrix dR/d(du/dt) (jitbridge.h:285 flag; 2 needs time stepping on,
 *                 else the mass matrix is 0)
 *   deriv_param   -1, or p >= 0: d/d(global parameter p) of both (with flag 2: dM/dp,
 *                 for problems whose bulk forms are generated Taylor-Hood forms only)
 *                 (get_derivative_wrt_global_parameter, src/problem.cpp:922;
 *                 fill_in_generic_dresidual_contribution_jit, src/elements.cpp:5005-5052)
 *   precision     PYHIP_F64 / PYHIP_F32: arithmetic of the element kernels
 *   residual_dev  [ndof] double (always FP64 accumulation of FP64 or FP32 element values)
 *   values_dev    [nnz] double (F64) or float (F32), column-sorted structural pattern;
 *                 may be NULL when flag == 0
 * Element-order, atomics-free summation: values are bit-for-bit reproducible. */
int pyhip_assemble(pyhip_problem *p, const double *dofs_dev, int flag, int deriv_param,
                   int precision, double *residual_dev, void *values_dev);

/* The matrix of the last complete FP64 assembly with a matrix output (the Jacobian, the
 * mass matrix, dJ/dp or a Hessian-vector product; values_dev holds that assembly's values)
 * in the reference's CSR form: exact zeros
 * dropped per element contribution (problem.cc:5524, Numerical_zero_for_sparse_assembly
 * = 0.0, :110) and columns in first-touch order (problem.cc:5618-5659).
 * row_start_dev [ndof+1], col_dev / val_dev [nnz] (structural nnz is an upper bound).
 * *nnz_out receives the number of entries written. */
/* ---- time stepping -------------------------------------------------------------------------
 * Time derivative for the assemblies and Newton solves that follow:
 * du/dt = weights[0] u + weights[1] u(t-1) + weights[2] u(t-2), the weak form gaining
 * (du/dt) psi_i (scalar forms) or rho (du/dt) . psi_i (Taylor-Hood momentum). The weights are the BDF weights of the reference's MultiTimeStepper
 * (src/timestepper.cpp:32-59; BDF2, degraded to BDF1 before the first completed step,
 * src/elements.cpp:4583-4636), 3 host values; hist1_dev / hist2_dev [ndof] hold the previous and
 * second-previous dofs (device, owned by the caller, read at every assembly). enabled = 0
 * returns to the steady forms. With time stepping
 * on, pyhip_assemble with flag 2 returns the residual and the mass matrix dR/d(du/dt)
 * (the reference's flag-2 output, src/jitbridge.h:285) in values. */
int pyhip_set_time_stepping(pyhip_problem *p, int enabled, const double *weights,
                            const double *hist1_dev, const double *hist2_dev);

/* Allocation-free variant of pyhip_assemble (same arguments, same results bit for bit).
 * It has no counterpart in the reference: it is for callers that cannot allocate device
 * memory -- code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc. The element vectors and matrices that pyhip_assemble keeps in buffers it
 * allocates on first use live here in one static device array of
 * PYHIP_NOALLOC_SCRATCH_BYTES per GPU (held until the process ends). The call allocates
 * nothing and rejects, with PYHIP_ERR_INVALID, a problem whose element buffers
 * (elements x (slots + slots^2) values of the precision) exceed that size. Every call on
 * one GPU shares that GPU's array: calls that overlap on one GPU (other streams, other
 * threads) race. Nothing in the array has to start at zero (every entry the gather reads
 * is written first). After it, pyhip_export_reference_csr reports that no FP64 Jacobian
 * is held (the element matrices are not in the problem's own buffer). */
#define PYHIP_NOALLOC_SCRATCH_BYTES (64ll << 20)
int pyhip_assemble_noalloc(pyhip_problem *p, const double *dofs_dev, int flag, int deriv_param,
                           int precision, double *residual_dev, void *values_dev);

/* Hessian-vector product of the scalar forms (compiled or generated) and of Taylor-Hood:
 * values_dev [nnz] (structural pattern) receives sum_k dJ_ij/du_k y_k at dofs_dev,
 * residual_dev the residual; the reference's HessianVectorProduct with one vector
 * (src/codegen.cpp:3646-3910, assembled by src/elements.cpp:4964-4995, flag 1; for the
 * compiled scalar forms the transposed product, flag 4, is the same matrix, for a generated
 * form with a non-symmetric Jacobian it is not). Taylor-Hood: the convection block with
 * the velocity replaced by y (the only non-linear term). Several vectors: one call per
 * vector; the transposed product: pyhip_assemble_multi. pyhip_export_reference_csr then
 * gives its reference CSR. */
int pyhip_assemble_hessian_vector(pyhip_problem *p, const double *dofs_dev, const double *y_dev,
                                  double *residual_dev, double *values_dev);

int pyhip_export_reference_csr(pyhip_problem *p, const double *values_dev,
                               int32_t *row_start_dev, int32_t *col_dev, double *val_dev,
                               int64_t *nnz_out);

/* ---- multi-assembly -------------------------------------------------------------------------
 * Several quantities at the same dofs in one call: the reference's
 * Problem::assemble_multiassembly (src/problem.cpp:2326-2368) with the requests of
 * CustomMultiAssembleHandler (src/bifurcation.cpp:5354-5480), bound to Python as
 * _assemble_multiassembly (src/pybind/problem.cpp:631-650) and used through
 * MultiAssembleRequest (pyoomph/generic/bifurcation_tools.py:442-520). A device problem
 * has one residual (the reference's contribution ""). Request kinds (the reference's
 * "what" strings): */
#define PYHIP_MA_RESIDUALS 0                           /* "residuals"                      */
#define PYHIP_MA_JACOBIAN 1                            /* "jacobian"                       */
#define PYHIP_MA_MASS_MATRIX 2                         /* "mass_matrix"                    */
#define PYHIP_MA_DRESIDUALS_DPARAMETER 3               /* "dresiduals_dparameter"          */
#define PYHIP_MA_DJACOBIAN_DPARAMETER 4                /* "djacobian_dparameter"           */
#define PYHIP_MA_DMASS_MATRIX_DPARAMETER 5             /* "dmass_matrix_dparameter"        */
#define PYHIP_MA_HESSIAN_VECTOR_PRODUCT 6              /* "hessian_vector_product[_transposed]" */
#define PYHIP_MA_MASS_MATRIX_HESSIAN_VECTOR_PRODUCT 7  /* "mass_matrix_hessian_vector_product[_transposed]" */

typedef struct {
    int what;        /* PYHIP_MA_* */
    int param;       /* global parameter index of the three parameter kinds (else ignored) */
    int hvec;        /* row of the Hessian vectors for the two Hessian kinds (else ignored) */
    int transposed;  /* Hessian kinds: 1 = entry (i, k) = sum_j y_j d^2 R_j/(du_i du_k)
                        (jitbridge.h:678-691), 0 = sum_j y_j d^2 R_i/(du_j du_k) */
    const char *contribution;  /* the residual table (NULL or "": the default residual; a
                        name: a generated form's table, pyhip_codegen_add_residual_table) */
} pyhip_multi_request;

/* Numbering of the outputs (CustomMultiAssembleHandler, src/bifurcation.cpp:5393-5480):
 * return_indices[i] = k >= 0 for the k-th vector, -1 - m for the m-th matrix, counted per
 * contribution in order of first appearance and within it in request order. Errors as the reference's: the same quantity requested twice, a parameter
 * or Hessian vector missing, a Hessian vector index >= nhvec, transposed and plain Hessian
 * products in one call. Host-only: no device is touched. */
int pyhip_multi_assembly_plan(int nreq, const pyhip_multi_request *req, int nhvec,
                              int *return_indices, int *nvector, int *nmatrix);

/* Runs the requests at dofs_dev and returns when the outputs are written.
 *   hvecs_dev        [nhvec][ndof] the Hessian directions (device)
 *   vectors          host array [nvector] of device pointers, each [ndof]
 *   row_start_dev    host array [nmatrix] of device pointers, each [ndof+1]
 *   col_dev, val_dev host arrays [nmatrix] of device pointers, each [nnz] (structural nnz
 *                    of the problem is an upper bound)
 *   nnz_out          host [nmatrix], entries written per matrix
 * Matrices come as the reference's multi-assembly returns them: entries with at least one
 * non-zero element contribution, each the element-order sum of those contributions
 * (src/problem.cpp:2195-2203), columns ascending within a row (one std::map per row,
 * :2232-2262). The mass matrix is dR/d(du/dt) (flag 2) whether or not time stepping is on;
 * dM/dp and the mass-matrix Hessian products have no non-zero element contribution for
 * these forms (their mass coefficients depend neither on the parameter nor on the dofs):
 * empty matrices; a generated Taylor-Hood form's dM/dp comes from its kernels (its mass
 * coefficient may hold the parameter; refused when other bulk forms share the problem).
 * Hessian products of run-time generated forms whose source defines no Hessian kernels,
 * and the mass-matrix Hessian products of generated Taylor-Hood forms (whose mass
 * coefficient may depend on the dofs), are refused with PYHIP_ERR_INVALID. */
int pyhip_assemble_multi(pyhip_problem *p, const double *dofs_dev, int nreq,
                         const pyhip_multi_request *req, const double *hvecs_dev, int nhvec,
                         double *const *vectors, int32_t *const *row_start_dev,
                         int32_t *const *col_dev, double *const *val_dev, int64_t *nnz_out);

/* ---- linear solver ------------------------------------------------------------------------
 * Preconditioned restarted GMRES on the device, batched over sub-problems (each has its
 * own Krylov scalars and convergence test). Replaces the SuperLU solve
 * (linear_solver.cc:846-1105 -> src/pybind/solver.cpp:102-142 ->
 * pyoomph/solvers/scipy.py:47-104). */
#define PYHIP_PRECOND_JACOBI 0
#define PYHIP_PRECOND_ILU0 1
#define PYHIP_PRECOND_AMG 2    /* smoothed-aggregation multigrid V-cycle (FP64 solves) */
#define PYHIP_PRECOND_BLOCK 3  /* Taylor-Hood: block-triangular, multigrid V-cycle on the velocity
                                  components and the lumped pressure mass / viscosity for the
                                  Schur complement (FP64 solves; scalar forms: multigrid) */
/* ILU(0) row order within a sub-problem (Taylor-Hood: velocity unknowns, then pressure).
 * NATURAL keeps the equation order; MULTICOLOR sorts each group by a greedy colouring,
 * so the level-scheduled triangular solves need one launch per colour instead of one per
 * level of the natural order (O(rows per side)), for a weaker factor. AUTO takes
 * MULTICOLOR for each sub-problem whose natural order has more than 64 lower-solve
 * levels. */
#define PYHIP_ILU_ORDER_AUTO 0
#define PYHIP_ILU_ORDER_NATURAL 1
#define PYHIP_ILU_ORDER_MULTICOLOR 2
#define PYHIP_KRYLOV_GMRES 0
#define PYHIP_KRYLOV_CG 1    /* symmetric positive definite Jacobians (Poisson, Bratu
                                below its fold, reaction-diffusion with u^3)           */
typedef struct {
    int krylov;         /* PYHIP_KRYLOV_*                                             */
    int precond;        /* PYHIP_PRECOND_*                                            */
    int precision;      /* Krylov + preconditioner arithmetic: PYHIP_F64 / PYHIP_F32   */
    int restart;        /* GMRES restart length m (1..1000)                           */
    int max_iters;      /* total inner iterations per solve                           */
    double rtol;        /* relative residual target of an FP64 inner solve            */
    double rtol_f32;    /* relative residual target of an FP32 inner solve            */
    int max_refinements;/* FP64 iterative-refinement sweeps around an FP32 solve       */
    double refine_rtol; /* relative residual target of the refined solution            */
    int ilu_order;      /* PYHIP_ILU_ORDER_*: row order of the ILU(0) factor           */
    int amg_element_level; /* 1 (default): in the multigrid, sub-problems on P2 / Q2
                           elements of a scalar form coarsen first to the linear (P1 / Q1)
                           functions on their vertices, then by aggregation; 0: by
                           aggregation from the first level. Changing it rebuilds the
                           hierarchy.                                                   */
} pyhip_linear_opts;
void pyhip_linear_opts_default(pyhip_linear_opts *o);

/* ---- Newton ---------------------------------------------------------------------------------
 * Replaces oomph Problem::newton_solve (problem.cc:8986-9345) / steady_newton_solve
 * (:9495-9577) with pyoomph's defaults (pyoomph/generic/problem.py:409-411,536). Every
 * sub-problem runs its own Newton iteration (own convergence test and error state); the
 * device work of all of them is batched in shared launches. */
typedef struct {
    double tol;                      /* Newton_solver_tolerance (1e-8)                */
    int max_iter;                    /* Max_newton_iterations (10)                     */
    double max_residuals;            /* Max_residuals (1e10)                           */
    int always_take_one_newton_step; /* (1)                                            */
    double relaxation;               /* Relaxation_factor (1.0)                        */
    int jacobian_precision;          /* element kernels for J: PYHIP_F64 / PYHIP_F32   */
    pyhip_linear_opts linear;
    /* 0 (default): the reference's test max|R| < tol alone. > 0: a sub-problem also needs
     * max|dx| <= correction_tol * max|x| of its last Newton update (FP64 norms). With an
     * FP32 Jacobian Newton converges linearly, and max|R| < tol does not bound the error
     * of the solution; the update does (error ~ contraction * |dx|). No counterpart in
     * the reference, whose Jacobians are FP64.                                         */
    double correction_tol;
} pyhip_newton_opts;
typedef struct {
    int newton_iters;       /* Newton steps taken                                     */
    int linear_iters;       /* inner Krylov iterations, all steps                      */
    int linear_failures;    /* linear solves that stopped above their tolerance        */
    int converged;          /* 1 converged, 0 NewtonSolverError                        */
    double final_maxres;    /* max |R| after the last step                             */
    double history[16];     /* max |R| per residual evaluation (Max_res)               */
    int nhistory;
    double done_ms;         /* wall time from the start of the solve to the Newton step
                               that ended this sub-problem's iteration                  */
} pyhip_newton_stats;
void pyhip_newton_opts_default(pyhip_newton_opts *o);
/* dofs_dev [ndof] in/out. stats: host array [nsub]. Returns PYHIP_ERR_NEWTON if any
 * sub-problem failed (its stats say which). */
int pyhip_newton_solve(pyhip_problem *p, double *dofs_dev, const pyhip_newton_opts *opts,
                       pyhip_newton_stats *stats);

/* One Newton iteration of pyhip_newton_solve for the sub-problems with mask[s] != 0 that
 * have not converged or failed yet; state[s] (host, [nsub]) carries each sub-problem's
 * progress between calls, newton_iters = 0 starting a new solve (the caller puts the
 * start values into its dofs). Convergence and failure follow pyhip_newton_solve, and a
 * sub-problem whose start residual is not finite fails at once. Element kernels, the
 * preconditioner set-up and the solves cover the stepped sub-problems only. Every
 * sub-problem's iteration is its own (per-sub-problem assembly rows, Krylov scalars and
 * tests; the multigrid hierarchy is rebuilt from every call's Jacobian), so a
 * sub-problem gives the same iterates whichever others run in the same calls.
 * Sub-problems can thus start and finish at any call: the slots of a serving loop. No counterpart in the reference, which solves one
 * problem per process (pyoomph/utils/paramscan.py:249-277). */
typedef struct {
    int newton_iters;       /* in/out: Newton steps taken in this solve; 0 starts one      */
    int converged;          /* in/out: 1 once converged (the sub-problem is then skipped)   */
    int failed;             /* in/out: 1 once NewtonSolverError (max_residuals / max_iter) */
    double maxres;          /* out: max |R| after the last step                            */
    double max_update;      /* out: max |dx| of the last step                              */
    int linear_iters;       /* in/out: accumulated inner Krylov iterations                  */
    int linear_failures;    /* in/out: accumulated linear solves above their tolerance      */
} pyhip_newton_step_state;
int pyhip_newton_step(pyhip_problem *p, double *dofs_dev, const pyhip_newton_opts *opts, const int32_t *mask,
                      pyhip_newton_step_state *state);

/* Moves the problem to another HIP stream (waits for the work queued on the old one); its
 * later calls enqueue there. Lets a serving loop run whichever pool steps on one of a few
 * streams, each on its own hardware queue. */
int pyhip_problem_set_stream(pyhip_problem *p, void *stream);

/* Replaces one sub-problem's auxiliary nodal field (reaction-diffusion source f): aux
 * [nnode] host values, nnode the sub-problem's node count. */
int pyhip_problem_set_aux(pyhip_problem *p, int sub, const double *aux, int64_t nnode);

/* ---- arclength continuation ------------------------------------------------------------------
 * One pseudo-arclength step in global parameter `param` (0: the Bratu lambda) for every
 * sub-problem with active[s] != 0, each with its own step length ds[s] and state:
 * oomph's Problem::arc_length_step_solve (problem.cc:10718-11080, no spatial adaptation)
 * with newton_solve_continuation (problem.cc:9599-9917), the
 * block elimination of the bordered system (two solves with J per Newton step,
 * J y = R and J z = dR/dparam), calculate_continuation_derivatives (:9918-10020,
 * Scale_arc_length with Desired_proportion_of_arc_length 0.5) and the step-length rule
 * (Desired_newton_iterations_ds 5: x1.5 below, x2/3 above; a failed Newton solve retries
 * with ds x2/3 down to Minimum_ds 1e-10). Entered by pyoomph's
 * Problem.arclength_continuation (pyoomph/generic/problem.py:3585, src/problem.cpp:1010).
 * The library keeps each sub-problem's Dof_current and Dof_derivative; a state with
 * step_taken = 0 starts from a zero Dof_derivative. The sub-problems' parameters are
 * updated in place (pyhip_problem_set_params values). ds_next[s] is the suggested next
 * step. Linear solves use opts->linear (GMRES: J is indefinite past a fold). */
typedef struct {
    double theta_squared;        /* Theta_squared (1.0 initially)                       */
    double parameter_derivative; /* Parameter_derivative (1.0 initially)                */
    double parameter_current;    /* Parameter_current (set by the step)                 */
    double ds_current;           /* Ds_current (set by the step)                        */
    int step_taken;              /* Arc_length_step_taken (0 initially)                 */
    double parameter;            /* out: the parameter after the step                   */
    int newton_iters;            /* out: Newton iterations of the accepted solve        */
    int rejected;                /* out: retries with a shorter step in this call      */
    int linear_failed;           /* out: 1 if a linear solve of this sub-problem ended
                                    above its tolerance; the step is then not taken (its
                                    dofs, parameter and state are kept) and the call
                                    returns PYHIP_ERR_NOT_CONVERGED after finishing the
                                    other sub-problems (the reference throws,
                                    problem.cc:10905-10915)                              */
    int turning_point;           /* out: 1 if Parameter_derivative changed sign in this
                                    step (0 on a sub-problem's first step): a fold lies
                                    between parameter_current and parameter, where det J
                                    changes sign. oomph's detection reads the sign of det J
                                    from SuperLU (problem.cc:10944-10975) and is off
                                    unless Bifurcation_detection is set (pyoomph never
                                    sets it, and its SuperLU plugin returns sign 0,
                                    pyoomph/solvers/scipy.py:101); the batched Krylov
                                    solves have no determinant, and the bordered system's
                                    determinant, det J (Parameter_derivative -
                                    Theta^2 Dof_derivative.z), keeps its sign along a
                                    branch, so the sign of Parameter_derivative tracks
                                    that of det J. Branch points (det J changing sign
                                    without a fold) are not detected.                    */
} pyhip_arclength_state;
int pyhip_arclength_step(pyhip_problem *p, double *dofs_dev, int param, const pyhip_newton_opts *opts,
                         const int32_t *active, const double *ds, pyhip_arclength_state *state,
                         double *ds_next);

/* ---- kernel timing -----------------------------------------------------------------------
 * Measurement aid (no counterpart in the reference, whose only timers are wall clocks,
 * problem.cc:8990,9149): while enabled, the problem records a pair of HIP events on its
 * stream around every launch of the named kernel families, "element" (element
 * residual / Jacobian kernels), "gather_values", "gather_residual", "spmv" (every sparse
 * matrix-vector launch: the Krylov solver's and, inside the multigrid V-cycle, those of
 * each level's matrix, restriction and prolongation), "precond" (whole preconditioner
 * applies and factorisations, including the spmv launches they make), "vector" (dots,
 * updates).
 * pyhip_timing_read returns the launches and summed device milliseconds so far;
 * pyhip_timing_bytes the algorithmic bytes those launches moved, where the family counts
 * them ("spmv": matrix values and column indices once, row starts, each entry of the input
 * vector once, the output and the added vector once per row, over the rows each launch
 * actually processed). */
int pyhip_timing_enable(pyhip_problem *p, int enable);
int pyhip_timing_read(pyhip_problem *p, const char *family, int64_t *launches, double *total_ms);
int pyhip_timing_bytes(pyhip_problem *p, const char *family, int64_t *bytes);

/* ---- linear-solver plugin entry --------------------------------------------------------------
 * Same arguments and meaning as GenericLinearSystemSolver.solve_serial
 * (pyoomph/solvers/generic.py:77-78) and the extern "C" superlu() wrapper
 * (src/pybind/solver.cpp:105-108): op_flag 1 = factorise (here: upload + preconditioner
 * setup), 2 = solve in place in b, 3 = free. The matrix arrives as oomph passes it to
 * SuperLU: compressed-column arrays (values, rowind, colptr) of J^T, and transpose != 0
 * asks for J x = b. Host arrays; *handle keeps the factorisation between calls. */
int pyhip_solve_serial(int op_flag, int n, int nnz, int nrhs, const double *values,
                       const int32_t *rowind, const int32_t *colptr, double *b, int ldb,
                       int transpose, void **handle);

#ifdef __cplusplus
}
#endif
#endif
