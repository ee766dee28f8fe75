// Device-side element helpers shared by the compiled element kernels (element_kernels.h)
// and the kernels generated at run time (pyoomph_amd/codegen, compiled with hiprtc, which
// receives this file as a header): the reference-element tables, the kernel arguments,
// geometry per integration point and small vector helpers. Device code only, no host
// library headers.
#pragma once
#if defined(__HIPCC_RTC__)
typedef __hip_internal::int8_t int8_t;
typedef __hip_internal::int32_t int32_t;
typedef __hip_internal::int64_t int64_t;
#else
#include <hip/hip_runtime.h>
#include <cstdint>
#endif

namespace pyhip {

// element families (the values of PYHIP_ELEM_*)
enum ElemKind : int {
    ELEM_TRI_P1 = 1,    // BulkElementTri2dC1, TGauss<2,2>
    ELEM_TRI_P2 = 2,    // BulkElementTri2dC2, TGauss<2,3>
    ELEM_QUAD_Q2 = 3,   // BulkElementQuad2dC2, Gauss<2,3>
    ELEM_HEX_Q2 = 4,    // BulkElementBrick3dC2, Gauss<3,3>
};

template <typename T>
struct DevTables {
    T w[27];
    T psi[27 * 27];
    T dpsi[27 * 27 * 3];
    T psi1[27 * 4];
};

template <typename T>
struct ElemArgs {
    int nel;                  // elements in this launch (one element group)
    const int32_t *conn;      // [nnode_el][nel] global node (slot-major)
    const int32_t *seqn;      // [nslot][nel] global equation, -1 pinned
    const int32_t *sub;       // [nel] sub-problem id
    const double *coords;     // [nnode_total][dim]
    const double *dofs;       // [ndof_total]
    const double *pinned;     // [nnode_total][nvalue]
    const double *aux;        // [nnode_total] or nullptr
    const double *params;     // [nsub][pstride]; Poisson: n, then (cx, cy, w, a) * n
    int pstride;
    int nvalue;
    const DevTables<T> *tab;
    T *Er;                    // [nel][nslot]
    T *Ej;                    // [nel][nslot][nslot] or nullptr
    int flag;                 // 0 residual, 1 residual + Jacobian, 2 residual + mass matrix
    int deriv;                // 1: d/dlambda (Bratu)
    // time derivative of the scalar forms: du/dt = tw0 u + (tw1 u(t-1) + tw2 u(t-2)),
    // history dofs [ndof_total] (pinned nodes keep their boundary value)
    int transient;
    double tw0, tw1, tw2;
    const double *hist1, *hist2;
    // Hessian-vector product: Ej receives sum_k dJ/du_k y_k (y dof-indexed, 0 at pinned);
    // htr = 1: the transposed product, entry (i, k) = sum_j y_j d^2 R_j / du_i du_k
    const double *hvec;
    int htr;
    // boundary faces: [nel][2] alpha, beta of each face's flux alpha*u + beta
    const double *fcoef;
    // elements [e0, e1) of the group are computed by this launch (the slot-major arrays
    // keep the stride nel): the assembly of a subset of the sub-problems launches once per
    // contiguous run of their elements
    int e0, e1;
};

// Line faces of the 2D element families (the interface elements of the boundary-face
// forms): TElement<1,2> / TElement<1,3> lines on s in [0, 1] with TGauss<1,2> / TGauss<1,3>
// (faces of P1 / P2 triangles; integral.cc:303-316) and the QElement<1,3> line on [-1, 1]
// with Gauss<1,3> (faces of Q2 quadrilaterals; integral.cc:50-53): knot and weight of point
// g, Lagrange shapes and their s-derivatives.
template <int NN> struct FaceRule;
template <> struct FaceRule<2> {
    __device__ static void point(int g, double &s, double &w) {
        s = g == 0 ? 0.5 * (-0.577350269189626 + 1.0) : 0.5 * (0.577350269189626 + 1.0);
        w = 0.5;
    }
    template <typename T> __device__ static void shape(T s, T (&psi)[2], T (&dpsi)[2]) {
        psi[0] = T(1.0) - s;
        psi[1] = s;
        dpsi[0] = T(-1.0);
        dpsi[1] = T(1.0);
    }
};
template <> struct FaceRule<3> {
    __device__ static void point(int g, double &s, double &w) {
        s = g == 0 ? 0.5 * (-0.774596669241483 + 1.0) : (g == 1 ? 0.5 : 0.5 * (0.774596669241483 + 1.0));
        w = g == 1 ? (8.0 / 18.0) : (5.0 / 18.0);
    }
    template <typename T> __device__ static void shape(T s, T (&psi)[3], T (&dpsi)[3]) {
        psi[0] = T(2.0) * (s - T(1.0)) * (s - T(0.5));
        psi[1] = T(4.0) * (T(1.0) - s) * s;
        psi[2] = T(2.0) * (s - T(0.5)) * s;
        dpsi[0] = T(4.0) * s - T(3.0);
        dpsi[1] = T(4.0) - T(8.0) * s;
        dpsi[2] = T(4.0) * s - T(1.0);
    }
};

// Q2-quadrilateral faces: QElement<1,3>, Gauss<1,3> on [-1, 1]
struct FaceRuleQ3 {
    __device__ static void point(int g, double &s, double &w) {
        s = g == 0 ? -0.774596669241483 : (g == 1 ? 0.0 : 0.774596669241483);
        w = g == 1 ? (8.0 / 9.0) : (5.0 / 9.0);
    }
    template <typename T> __device__ static void shape(T s, T (&psi)[3], T (&dpsi)[3]) {
        psi[0] = T(0.5) * s * (s - T(1.0));
        psi[1] = T(1.0) - s * s;
        psi[2] = T(0.5) * s * (s + T(1.0));
        dpsi[0] = s - T(0.5);
        dpsi[1] = T(-2.0) * s;
        dpsi[2] = s + T(0.5);
    }
};

template <int KIND> struct ElemTraits;
template <> struct ElemTraits<ELEM_TRI_P1> { static constexpr int NN = 3, DIM = 2, NQ = 3; };
template <> struct ElemTraits<ELEM_TRI_P2> { static constexpr int NN = 6, DIM = 2, NQ = 7; };
template <> struct ElemTraits<ELEM_QUAD_Q2> { static constexpr int NN = 9, DIM = 2, NQ = 9; };
template <> struct ElemTraits<ELEM_HEX_Q2> { static constexpr int NN = 27, DIM = 3, NQ = 27; };

__device__ __forceinline__ double texp(double x) { return ::exp(x); }
__device__ __forceinline__ float texp(float x) { return ::expf(x); }
__device__ __forceinline__ double tsqrt(double x) { return ::sqrt(x); }
__device__ __forceinline__ float tsqrt(float x) { return ::sqrtf(x); }

// Metric inverse and g^{ab} g_a (src/elements.cpp:3655-3700 / 3800-3840), from the
// tangents t[a][i]. Returns w*sqrt(det).
template <typename T, int DIM>
__device__ __forceinline__ T metric(const T (&t)[DIM][DIM], T wq, T (&g)[DIM][DIM]) {
    T am[DIM][DIM];
#pragma unroll
    for (int a = 0; a < DIM; a++)
#pragma unroll
        for (int b = 0; b < DIM; b++) {
            T s = T(0);
#pragma unroll
            for (int i = 0; i < DIM; i++) s = s + t[a][i] * t[b][i];
            am[a][b] = s;
        }
    T au[DIM][DIM];
    T det_a;
    if constexpr (DIM == 2) {
        det_a = am[0][0] * am[1][1] - am[0][1] * am[1][0];
        au[0][0] = am[1][1] / det_a;
        au[0][1] = -am[0][1] / det_a;
        au[1][0] = -am[1][0] / det_a;
        au[1][1] = am[0][0] / det_a;
    } else {
        det_a = am[0][0] * am[1][1] * am[2][2] + am[0][1] * am[1][2] * am[2][0] +
                am[0][2] * am[1][0] * am[2][1] - am[0][0] * am[1][2] * am[2][1] -
                am[0][1] * am[1][0] * am[2][2] - am[0][2] * am[1][1] * am[2][0];
        au[0][0] = (am[1][1] * am[2][2] - am[1][2] * am[2][1]) / det_a;
        au[0][1] = -(am[0][1] * am[2][2] - am[0][2] * am[2][1]) / det_a;
        au[0][2] = (am[0][1] * am[1][2] - am[0][2] * am[1][1]) / det_a;
        au[1][0] = -(am[1][0] * am[2][2] - am[1][2] * am[2][0]) / det_a;
        au[1][1] = (am[0][0] * am[2][2] - am[0][2] * am[2][0]) / det_a;
        au[1][2] = -(am[0][0] * am[1][2] - am[0][2] * am[1][0]) / det_a;
        au[2][0] = (am[1][0] * am[2][1] - am[1][1] * am[2][0]) / det_a;
        au[2][1] = -(am[0][0] * am[2][1] - am[0][1] * am[2][0]) / det_a;
        au[2][2] = (am[0][0] * am[1][1] - am[0][1] * am[1][0]) / det_a;
    }
#pragma unroll
    for (int b = 0; b < DIM; b++)
#pragma unroll
        for (int i = 0; i < DIM; i++) {
            T s = au[0][b] * t[0][i] + au[1][b] * t[1][i];
            if constexpr (DIM == 3) s = s + au[2][b] * t[2][i];
            g[b][i] = s;
        }
    return wq * tsqrt(det_a);
}

template <typename T, int DIM>
__device__ __forceinline__ T dotv(const T *a, const T *b) {
    T s = a[0] * b[0] + a[1] * b[1];
    if constexpr (DIM == 3) s = s + a[2] * b[2];
    return s;
}

// Per-integration-point geometry and field values, computed once per (element, point)
// by one thread and shared through LDS: tangents, metric inverse, w*det, dx_shape of
// every node, and the interpolated fields. Same operation order as the CPU restatement
// of the generated code (fill_shape_info_at_s, then the interpolation loops).
template <typename T, int NN, int DIM>
__device__ __forceinline__ T point_geometry(const T (*X)[DIM], const T *__restrict__ D, T wq,
                                            T (*dx)[DIM]) {
    T t[DIM][DIM];
#pragma unroll
    for (int i = 0; i < DIM; i++)
#pragma unroll
        for (int j = 0; j < DIM; j++) t[i][j] = T(0);
#pragma unroll
    for (int l = 0; l < NN; l++)
#pragma unroll
        for (int i = 0; i < DIM; i++)
#pragma unroll
            for (int j = 0; j < DIM; j++) t[j][i] = t[j][i] + X[l][i] * D[l * DIM + j];
    T g[DIM][DIM];
    const T W = metric<T, DIM>(t, wq, g);
#pragma unroll
    for (int l = 0; l < NN; l++)
#pragma unroll
        for (int i = 0; i < DIM; i++) {
            T s = T(0);
#pragma unroll
            for (int b = 0; b < DIM; b++) s = s + g[b][i] * D[l * DIM + b];
            dx[l][i] = s;
        }
    return W;
}

// Elements per 256-thread workgroup of the scalar-form kernels: every (element, slot row)
// pair of phase B and every (element, point) pair of a phase-A chunk gets a thread.
template <int KIND> struct ScalarChunk { static constexpr int QC = ElemTraits<KIND>::NQ; };
template <> struct ScalarChunk<ELEM_HEX_Q2> { static constexpr int QC = 9; };
template <int KIND>
constexpr int scalar_epb() {
    return 256 / (ElemTraits<KIND>::NN > ScalarChunk<KIND>::QC ? ElemTraits<KIND>::NN : ScalarChunk<KIND>::QC);
}

}  // namespace pyhip
