/*
 * The element-function table of pyoomph's JIT ABI, declared with the reference's memory
 * layout so that a host built against src/jitbridge.h can take a table filled by this
 * library (src/jitbridge.h:88-120 JITElementInfo, :139-151 hang info, :169-283
 * JITShapeInfo, :285-292 function types, :294-327 required shapes and callback entries,
 * :329-497 JITFuncSpec_Table_FiniteElement, :499 the init type). Field names, order and
 * types follow the reference (the non-fixed-size shape buffer variant), since the layout
 * is the interface; every pointer here is host memory.
 *
 * pyhip_jit_element_init() fills a table for one weak form on one element family, the
 * way a generated JIT_ELEMENT_init (src/codegen.cpp:6398-7090) does for its element:
 * field counts and names, the residual names, the ResidualAndJacobian /
 * ResidualAndJacobianSteady / ResidualAndJacobian_NoHang slots, ParameterDerivative,
 * global parameters, required shapes, integration order and clean_up. Each slot is a host
 * function with the reference signature that evaluates its element with this library's
 * device element kernel (one element per launch, the same kernel and bits as the batched
 * pyhip_assemble) and adds the result into the caller's zeroed R / J / M buffers by the
 * element's local equations, as generated code does. The per-element slots exist for
 * hosts that walk elements one by one (oomph's element loop, its finite-difference
 * checker); batched assembly goes through pyhip_assemble.
 *
 * Like the reference's generated code (a static table pointer per unit,
 * src/codegen.cpp:4120), the slots of one (form, element family) keep the table they
 * were filled into: one live table per (form, family) at a time, not re-entrant.
 */
#ifndef PYOOMPH_JIT_TABLE_H
#define PYOOMPH_JIT_TABLE_H

#include <stdbool.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct JITElementInfo {
    unsigned int nnode, nnode_C1, nnode_C2, nnode_C1TB, nnode_C2TB, nnode_DL;
    unsigned int nodal_dim;
    double ***nodal_coords;      /* [node][x index][time]                       */
    double ***nodal_data;        /* [node][field][time]                         */
    int **nodal_local_eqn;       /* [node][field], -1 pinned                    */
    int **pos_local_eqn;
    unsigned int ndof;
    bool alloced;
    void *elem_ptr;
    struct JITElementInfo *bulk_eleminfo;
    struct JITElementInfo *opposite_eleminfo;
} JITElementInfo_t;

typedef struct JITHangInfoEntry {
    double weight;
    int *local_eqn;
} JITHangInfoEntry_t;

typedef struct JITHangInfo {
    int nummaster;
    bool *has_contributions;
    JITHangInfoEntry_t *masters;
} JITHangInfo_t;

/* one space's shape data at the current integration point */
#define PYHIP_JIT_SPACE_FIELDS(S)                                                          \
    double *shape_##S;                                                                     \
    double **dx_shape_##S;                                                                 \
    double **dX_shape_##S;                                                                 \
    double **dS_shape_##S;                                                                 \
    double ****d_dx_shape_dcoord_##S;                                                      \
    double ******d2_dx2_shape_dcoord_##S;                                                  \
    double **nodal_shape_##S;

typedef struct JITShapeInfo {
    unsigned int n_int_pt;
    double int_pt_weight, int_pt_weight_Lagrangian, int_pt_weight_unity;
    double **int_pt_weights_d_coords;
    double ****int_pt_weights_d2_coords;
    double elemsize_Eulerian, elemsize_Eulerian_cartesian;
    double elemsize_Lagrangian, elemsize_Lagrangian_cartesian;
    double **elemsize_d_coords;
    double ****elemsize_d2_coords;
    double **elemsize_Cart_d_coords;
    double ****elemsize_Cart_d2_coords;
    PYHIP_JIT_SPACE_FIELDS(C2)
    PYHIP_JIT_SPACE_FIELDS(C2TB)
    PYHIP_JIT_SPACE_FIELDS(C1)
    PYHIP_JIT_SPACE_FIELDS(C1TB)
    PYHIP_JIT_SPACE_FIELDS(DL)
    double ****d_dshape_dx_tensor;
    double *shape_Pos;
    double **dx_shape_Pos;
    double **dX_shape_Pos;
    double **dS_shape_Pos;
    double ****d_dx_shape_dcoord_Pos;
    double ******d2_dx2_shape_dcoord_Pos;
    unsigned int jacobian_size;      /* row stride of J and M                      */
    unsigned int mass_matrix_size;
    double *normal;
    double ***d_normal_dcoord;
    double *****d2_normal_d2coord;
    double *t;
    double *dt;
    unsigned int timestepper_ntstorage;
    double *timestepper_weights_dt_BDF1;
    double *timestepper_weights_dt_BDF2;
    double *timestepper_weights_dt_Newmark2;
    double *timestepper_weights_d2t_Newmark2;
    double *timestepper_weights_dt_BDF2_degr;
    double *timestepper_weights_dt_Newmark2_degr;
    JITHangInfo_t *hanginfo_C1;
    JITHangInfo_t *hanginfo_C1TB;
    JITHangInfo_t *hanginfo_C2;
    JITHangInfo_t *hanginfo_C2TB;
    JITHangInfo_t *hanginfo_Pos;
    JITHangInfo_t *hanginfo_Discont;
    struct JITShapeInfo *bulk_shapeinfo;
    struct JITShapeInfo *opposite_shapeinfo;
    int *opposite_node_index;
} JITShapeInfo_t;
#undef PYHIP_JIT_SPACE_FIELDS

typedef void (*JITFuncSpec_ResidualAndJacobian_FiniteElement)(const JITElementInfo_t *, const JITShapeInfo_t *,
                                                             double *, double *, double *, unsigned);
typedef void (*JITFuncSpec_HessianVectorProduct_FiniteElement)(const JITElementInfo_t *, const JITShapeInfo_t *,
                                                              const double *, double *, double *, unsigned,
                                                              unsigned);
typedef void (*JITFuncSpec_GetZ2Fluxes_FiniteElement)(const JITElementInfo_t *, const JITShapeInfo_t *, double *);
typedef double (*JITFuncSpec_InitialCondition_FiniteElement)(const JITElementInfo_t *, int, double *, double *,
                                                             double *, double, int, double);
typedef double (*JITFuncSpec_DirichletCondition_FiniteElement)(const JITElementInfo_t *, int, double *, double *,
                                                               double *, double, double);
typedef double (*JITFuncSpec_EvalIntegralExpr_FiniteElement)(const JITElementInfo_t *, const JITShapeInfo_t *,
                                                             unsigned);
typedef void (*JITFuncSpec_EvalTracerAdvection_FiniteElement)(const JITElementInfo_t *, const JITShapeInfo_t *,
                                                              unsigned, double, double *);
typedef double (*JITFuncSpec_GeometricJacobian)(const JITElementInfo_t *, const double *);
typedef void (*JITFuncSpec_GeometricJacobianSpatialDerivative)(const JITElementInfo_t *, const double *, double *);

typedef struct JITFuncSpec_RequiredShapes_FiniteElement {
    bool psi_C1, psi_C1TB, psi_C2, psi_C2TB, psi_DL, psi_D0;
    bool dx_psi_C1, dx_psi_C1TB, dx_psi_C2, dx_psi_C2TB, dx_psi_DL, dx_psi_D0;
    bool dX_psi_C1, dX_psi_C1TB, dX_psi_C2, dX_psi_C2TB, dX_psi_DL, dX_psi_D0;
    bool psi_Pos, dx_psi_Pos, dX_psi_Pos;
    bool normal_Pos;
    bool elemsize_Eulerian_Pos, elemsize_Lagrangian_Pos;
    bool elemsize_Eulerian_cartesian_Pos, elemsize_Lagrangian_cartesian_Pos;
    struct JITFuncSpec_RequiredShapes_FiniteElement *bulk_shapes;
    struct JITFuncSpec_RequiredShapes_FiniteElement *opposite_shapes;
} JITFuncSpec_RequiredShapes_FiniteElement_t;

typedef struct JITFuncSpec_Callback_Entry {
    char *idname;
    unsigned unique_id;
    int is_deriv_of;
    int deriv_index;
    void *cb_obj;
} JITFuncSpec_Callback_Entry_t;

typedef struct JITFuncSpec_MultiRet_Entry {
    char *idname;
    unsigned unique_id;
    void *cb_obj;
} JITFuncSpec_MultiRet_Entry_t;

/* numfields of one space: total, on the bulk mesh, at the base bulk level, new here */
#define PYHIP_JIT_NUMFIELDS(S)                                                              \
    unsigned int numfields_##S, numfields_##S##_bulk, numfields_##S##_basebulk, numfields_##S##_new; \
    char **fieldnames_##S;

typedef struct JITFuncSpec_Table_FiniteElement {
    unsigned int nodal_dim, lagr_dim;
    PYHIP_JIT_NUMFIELDS(C1)
    PYHIP_JIT_NUMFIELDS(C2)
    PYHIP_JIT_NUMFIELDS(C2TB)
    PYHIP_JIT_NUMFIELDS(C1TB)
    PYHIP_JIT_NUMFIELDS(D1)
    PYHIP_JIT_NUMFIELDS(D1TB)
    PYHIP_JIT_NUMFIELDS(D2)
    PYHIP_JIT_NUMFIELDS(D2TB)
    unsigned int numfields_Pos;
    char **fieldnames_Pos;
    bool bulk_position_space_to_C1;
    int hangindex_C1, hangindex_C2, hangindex_C1TB, hangindex_C2TB, hangindex_Pos;
    unsigned int numfields_DL;
    char **fieldnames_DL;
    unsigned int numfields_D0;
    char **fieldnames_D0;
    unsigned int numfields_ED0;
    char **fieldnames_ED0;
    unsigned int nodal_offset_C2TB_basebulk, nodal_offset_C2_basebulk, nodal_offset_C1TB_basebulk,
        nodal_offset_C1_basebulk;
    unsigned int buffer_offset_C2TB_basebulk, buffer_offset_C2_basebulk, buffer_offset_C1TB_basebulk,
        buffer_offset_C1_basebulk;
    unsigned int buffer_offset_C2TB_interf, buffer_offset_C2_interf, buffer_offset_C1TB_interf,
        buffer_offset_C1_interf;
    unsigned int internal_offset_D2TB_new, internal_offset_D2_new, internal_offset_D1TB_new,
        internal_offset_D1_new;
    unsigned int buffer_offset_D2TB_basebulk, buffer_offset_D2_basebulk, buffer_offset_D1TB_basebulk,
        buffer_offset_D1_basebulk;
    unsigned int buffer_offset_D2TB_interf, buffer_offset_D2_interf, buffer_offset_D1TB_interf,
        buffer_offset_D1_interf;
    unsigned int external_offset_D2TB_bulk, external_offset_D2_bulk, external_offset_D1TB_bulk,
        external_offset_D1_bulk;
    unsigned int internal_offset_DL, internal_offset_D0;
    unsigned int buffer_offset_DL, buffer_offset_D0;
    unsigned int buffer_offset_ED0, external_offset_ED0;
    double *discontinuous_refinement_exponents;
    double *temporal_error_scales;
    bool has_temporal_estimators;
    unsigned num_res_jacs;
    int current_res_jac;
    char **res_jac_names;
    JITFuncSpec_RequiredShapes_FiniteElement_t *shapes_required_ResJac;
    JITFuncSpec_RequiredShapes_FiniteElement_t *shapes_required_Hessian;
    JITFuncSpec_RequiredShapes_FiniteElement_t merged_required_shapes;
    unsigned numglobal_params;
    unsigned *global_paramindices;
    double **global_parameters;       /* set by the host: pointers to its values      */
    JITFuncSpec_ResidualAndJacobian_FiniteElement **ParameterDerivative;   /* [res][param] */
    unsigned numintegral_expressions;
    char **integral_expressions_names;
    unsigned numlocal_expressions;
    char **local_expressions_names;
    unsigned numextremum_expressions;
    char **extremum_expressions_names;
    unsigned numtracer_advections;
    char **tracer_advection_names;
    char *dominant_space;
    int max_dt_order;
    bool fd_jacobian;
    bool fd_position_jacobian;
    double debug_jacobian_epsilon;
    bool with_adaptivity;
    bool stop_on_jacobian_difference;
    int integration_order;
    bool moving_nodes;
    bool use_shared_shape_buffer_during_multi_assemble, during_shared_multi_assembling;
    void *handle;
    JITFuncSpec_ResidualAndJacobian_FiniteElement *ResidualAndJacobian;
    JITFuncSpec_ResidualAndJacobian_FiniteElement *ResidualAndJacobianSteady;
    JITFuncSpec_ResidualAndJacobian_FiniteElement *ResidualAndJacobian_NoHang;
    bool *missing_residual_assembly;
    JITFuncSpec_HessianVectorProduct_FiniteElement *HessianVectorProduct;
    bool hessian_generated;
    unsigned num_Z2_flux_terms, num_Z2_flux_terms_for_eigen;
    JITFuncSpec_GetZ2Fluxes_FiniteElement GetZ2Fluxes, GetZ2FluxesForEigen;
    JITFuncSpec_RequiredShapes_FiniteElement_t shapes_required_Z2Fluxes;
    JITFuncSpec_InitialCondition_FiniteElement *InitialConditionFunc;
    unsigned num_ICs;
    char **IC_names;
    JITFuncSpec_DirichletCondition_FiniteElement DirichletConditionFunc;
    bool *Dirichlet_set;
    unsigned Dirichlet_set_size;
    char **Dirichlet_names;
    double (*invoke_callback)(void *, int, double *, int);
    void (*invoke_multi_ret)(void *, int, int, double *, double *, double *, int, int);
    JITFuncSpec_EvalIntegralExpr_FiniteElement EvalIntegralExpression;
    JITFuncSpec_RequiredShapes_FiniteElement_t shapes_required_IntegralExprs;
    JITFuncSpec_EvalIntegralExpr_FiniteElement EvalLocalExpression;
    JITFuncSpec_EvalIntegralExpr_FiniteElement EvalExtremumExpression;
    JITFuncSpec_RequiredShapes_FiniteElement_t shapes_required_LocalExprs;
    JITFuncSpec_RequiredShapes_FiniteElement_t shapes_required_ExtremumExprs;
    JITFuncSpec_EvalTracerAdvection_FiniteElement EvalTracerAdvection;
    JITFuncSpec_RequiredShapes_FiniteElement_t shapes_required_TracerAdvection;
    unsigned numcallbacks;
    JITFuncSpec_Callback_Entry_t *callback_infos;
    unsigned num_multi_rets;
    JITFuncSpec_MultiRet_Entry_t *multi_ret_infos;
    JITFuncSpec_GeometricJacobian GeometricJacobian;
    JITFuncSpec_GeometricJacobian JacobianForElementSize;
    JITFuncSpec_GeometricJacobianSpatialDerivative JacobianForElementSizeSpatialDerivative;
    JITFuncSpec_GeometricJacobianSpatialDerivative JacobianForElementSizeSecondSpatialDerivative;
    char *domain_name;
    bool *has_constant_mass_matrix_for_sure;
    void (*check_compiler_size)(unsigned long long, unsigned long long, char *);
    double (*get_element_size)(void *);
    void (*fill_shape_buffer_for_point)(unsigned, JITFuncSpec_RequiredShapes_FiniteElement_t *, int);
    void (*clean_up)(struct JITFuncSpec_Table_FiniteElement *functable);
} JITFuncSpec_Table_FiniteElement_t;
#undef PYHIP_JIT_NUMFIELDS

typedef void (*JIT_ELEMENT_init_SPEC)(JITFuncSpec_Table_FiniteElement_t *functable);

/* Fills a zeroed table (the host's `new` + zero, src/problem.cpp:110-111) for weak form
 * `form` (PYHIP_FORM_*, or a run-time generated form id from pyhip_codegen_compile, on the
 * family it was generated for) on element family `elem_kind` (PYHIP_ELEM_*), like a
 * generated JIT_ELEMENT_init. Fields: nodal_dim; the velocity / pressure or scalar field
 * counts and names in the space of the family (Taylor-Hood: "velocity_x", "velocity_y" in
 * C2, "pressure" in C1; scalar forms: "u", reaction-diffusion and generated forms with an
 * auxiliary field add it second; generated forms: the names pyhip_codegen_set_fields
 * gave); one residual "" with its ResidualAndJacobian, ResidualAndJacobianSteady and
 * ResidualAndJacobian_NoHang slots; Bratu and generated forms with dR/dp: one global
 * parameter with ParameterDerivative[0][0]; the required shapes; integration order 0 (the
 * family's default scheme); clean_up, which frees what the init allocated. The slots of a
 * generated form run its hiprtc-compiled kernel on the element (at most 16 generated forms
 * have tables at a time).
 * The host points global_parameters[i] at its parameter values before the slots run.
 * Returns PYHIP_OK or PYHIP_ERR_INVALID (unknown form / family). */
int pyhip_jit_element_init(int form, int elem_kind, JITFuncSpec_Table_FiniteElement_t *functable);

/* Form parameters the reference compiles into the generated code and this library keeps
 * per table: the Poisson Gaussian source (n, then cx, cy, w, a per Gaussian) and the
 * Navier-Stokes (mu, rho). Up to 64 values; Bratu's lambda comes from
 * global_parameters[0] instead. */
int pyhip_jit_set_form_params(JITFuncSpec_Table_FiniteElement_t *functable, const double *params, int nparams);

/* The reference's serial SuperLU entry (src/pybind/solver.cpp:105-108, called by oomph's
 * SuperLUSolver, linear_solver.cc:2297-2309, 2604-2616): op_flag 1 factorises (values,
 * rowind, colptr: the compressed-row Jacobian handed over as compressed columns of J^T,
 * with a non-null transpose pointer asking for J x = b, as the reference's wrapper reads
 * it), 2 solves in place in b (nrhs right-hand sides, leading dimension *ldb, or n when
 * ldb is null), 3 frees. *f_factors keeps the factorisation between calls. *info is 0 on
 * success, else the positive code of the failure (oomph then throws). The return value of
 * op_flag 1 is oomph's sign-of-determinant slot: 0, as the reference's SuperLU plugin
 * returns (pyoomph/solvers/scipy.py:103; a Krylov solve has no determinant); on failure
 * the negative PYHIP status. Arguments oomph leaves null (nrhs and b for op 1; nnz,
 * values, rowind, colptr for op 2) are not read. */
int pyhip_superlu(int *op_flag, int *n, int *nnz, int *nrhs, double *values, int *rowind, int *colptr, double *b,
                  int *ldb, int *transpose, int *doc, void **f_factors, int *info);

#ifdef __cplusplus
}
#endif
#endif
