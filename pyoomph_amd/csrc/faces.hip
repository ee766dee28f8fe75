// Boundary-face (interface) elements of the scalar forms: the flux term of pyoomph's
// NeumannBC(u=flux) @ boundary (pyoomph/meshes/bcs.py:71-89: residual weak(flux, test)),
// flux = alpha*u + beta with alpha, beta per face, on the faces of P1 / P2 triangles, Q2
// quadrilaterals and Q2 bricks.
//
// A face element of a triangle is the reference's InterfaceTElementLine1dC1 / C2
// (src/mesh.cpp:627-635): a TElement<1,2> / TElement<1,3> line (nodes at s = 0, 1 or
// 0, 1/2, 1; shapes Telements.h TElementShape<1,2>, <1,3>) with its default TGauss<1,2> /
// TGauss<1,3> scheme (integral.cc:303-316); a face element of a Q2 quadrilateral is
// InterfaceElementLine1dC2 (src/mesh.cpp:618-621): a QElement<1,3> line (nodes at
// s = -1, 0, 1; shape.h OneDimLagrange shape<3>, dshape<3>) with Gauss<1,3>
// (integral.cc:50-53). The functable's integration order is 0, so the element default
// scheme stays (src/mesh.cpp:664-667). Per point: tangent t = sum_l x_l dpsi_l/ds,
// W = w |t|, u = sum_l u_l psi_l;
//   R_i += W * (flux * psi_i),   J_ij += W * ((alpha * psi_j) * psi_i).
// The flux does not depend on lambda, on du/dt or on second derivatives: dR/dlambda, the
// mass matrix and the Hessian-vector product of a face are zero. One thread per face;
// the element vector / matrix layout is the bulk kernels' ([face][i], [face][i][j]).
#include <cmath>

#include "element_kernels.h"
#include "problem.h"

namespace pyhip {

// face rules (FaceRule<2>, FaceRule<3>, FaceRuleQ3): element_common.h

template <typename T, int NN, typename RULE>
__global__ __launch_bounds__(256) void flux_face_kernel(ElemArgs<T> a) {
    const int e = a.e0 + (int)(blockIdx.x * 256 + threadIdx.x);
    if (e >= a.e1) return;
    const T alpha = (T)a.fcoef[2 * (int64_t)e], beta = (T)a.fcoef[2 * (int64_t)e + 1];
    T X[NN][2], U[NN];
#pragma unroll
    for (int l = 0; l < NN; l++) {
        const int32_t n = a.conn[(int64_t)l * a.nel + e];
        X[l][0] = (T)a.coords[(int64_t)n * 3];
        X[l][1] = (T)a.coords[(int64_t)n * 3 + 1];
        const int32_t q = a.seqn[(int64_t)l * a.nel + e];
        U[l] = (T)(q >= 0 ? a.dofs[q] : a.pinned[(int64_t)n * a.nvalue]);
    }
    T R[NN], J[NN][NN];
#pragma unroll
    for (int i = 0; i < NN; i++) {
        R[i] = T(0);
#pragma unroll
        for (int j = 0; j < NN; j++) J[i][j] = T(0);
    }
#pragma unroll
    for (int g = 0; g < NN; g++) {
        double sd, wd;
        RULE::point(g, sd, wd);
        const T sq = (T)sd, wq = (T)wd;
        T psi[NN], dpsi[NN];
        RULE::shape(sq, psi, dpsi);
        T tx = T(0), ty = T(0), u = T(0);
#pragma unroll
        for (int l = 0; l < NN; l++) {
            tx = tx + X[l][0] * dpsi[l];
            ty = ty + X[l][1] * dpsi[l];
            u = u + U[l] * psi[l];
        }
        const T W = wq * sqrt(tx * tx + ty * ty);
        const T flux = alpha * u + beta;
#pragma unroll
        for (int i = 0; i < NN; i++) {
            R[i] = R[i] + W * (flux * psi[i]);
#pragma unroll
            for (int j = 0; j < NN; j++) J[i][j] = J[i][j] + W * ((alpha * psi[j]) * psi[i]);
        }
    }
    // dR/dlambda, the mass matrix and the Hessian-vector product of a face vanish
    const bool zero_r = a.deriv != 0;
    const bool zero_j = a.deriv != 0 || a.flag == 2 || a.hvec != nullptr;
#pragma unroll
    for (int i = 0; i < NN; i++) {
        a.Er[(int64_t)e * NN + i] = zero_r ? T(0) : R[i];
        if (a.Ej) {
#pragma unroll
            for (int j = 0; j < NN; j++) a.Ej[((int64_t)e * NN + i) * NN + j] = zero_j ? T(0) : J[i][j];
        }
    }
}

// Faces of Q2 bricks: InterfaceElementQuad2dC2 (src/mesh.cpp:650-653), a QElement<2,3>
// with its Gauss<2,3> scheme and shapes -- the Q2 quadrilateral tables (a.tab), mistyped
// knot included -- on a face in 3D: t_d = sum_l x_l dpsi_l/ds_d, W = w |t_0 x t_1|. Per
// face the point weights and fluxes come first; then every residual entry and Jacobian
// entry sums its nine point terms in point order:
//   R_i = sum_q W_q (flux_q psi_qi),   J_ij = sum_q W_q ((alpha psi_qj) psi_qi).
template <typename T>
__global__ __launch_bounds__(256) void flux_face_quad9_kernel(ElemArgs<T> a) {
    constexpr int NN = 9, NQ = 9;
    const int e = a.e0 + (int)(blockIdx.x * 256 + threadIdx.x);
    if (e >= a.e1) return;
    const T alpha = (T)a.fcoef[2 * (int64_t)e], beta = (T)a.fcoef[2 * (int64_t)e + 1];
    T X[NN][3], U[NN];
#pragma unroll
    for (int l = 0; l < NN; l++) {
        const int32_t n = a.conn[(int64_t)l * a.nel + e];
#pragma unroll
        for (int d = 0; d < 3; d++) X[l][d] = (T)a.coords[(int64_t)n * 3 + d];
        const int32_t q = a.seqn[(int64_t)l * a.nel + e];
        U[l] = (T)(q >= 0 ? a.dofs[q] : a.pinned[(int64_t)n * a.nvalue]);
    }
    T W[NQ], F[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const T *P = a.tab->psi + q * NN;
        const T *D = a.tab->dpsi + q * NN * 2;
        T t0[3] = {T(0), T(0), T(0)}, t1[3] = {T(0), T(0), T(0)}, u = T(0);
#pragma unroll
        for (int l = 0; l < NN; l++) {
#pragma unroll
            for (int d = 0; d < 3; d++) {
                t0[d] = t0[d] + X[l][d] * D[l * 2];
                t1[d] = t1[d] + X[l][d] * D[l * 2 + 1];
            }
            u = u + U[l] * P[l];
        }
        const T n0 = t0[1] * t1[2] - t0[2] * t1[1];
        const T n1 = t0[2] * t1[0] - t0[0] * t1[2];
        const T n2 = t0[0] * t1[1] - t0[1] * t1[0];
        W[q] = a.tab->w[q] * sqrt((n0 * n0 + n1 * n1) + n2 * n2);
        F[q] = alpha * u + beta;
    }
    const bool zero_r = a.deriv != 0;
    const bool zero_j = a.deriv != 0 || a.flag == 2 || a.hvec != nullptr;
    for (int i = 0; i < NN; i++) {
        T r = T(0);
#pragma unroll
        for (int q = 0; q < NQ; q++) r = r + W[q] * (F[q] * a.tab->psi[q * NN + i]);
        a.Er[(int64_t)e * NN + i] = zero_r ? T(0) : r;
        if (a.Ej) {
            for (int j = 0; j < NN; j++) {
                T v = T(0);
#pragma unroll
                for (int q = 0; q < NQ; q++)
                    v = v + W[q] * ((alpha * a.tab->psi[q * NN + j]) * a.tab->psi[q * NN + i]);
                a.Ej[((int64_t)e * NN + i) * NN + j] = zero_j ? T(0) : v;
            }
        }
    }
}

// Traction faces of Taylor-Hood Q2 quadrilaterals: NeumannBC(velocity_x=f_x,
// velocity_y=f_y) with f_c = alpha*u_c + beta_c (coefficients alpha, beta_x, beta_y per
// face) on QElement<1,3> faces with Gauss<1,3>; the face's slots are (node l, component c)
// = 2l + c. Per point, in the order of flux_face_kernel:
//   R_{lc} += W * (f_c * psi_l),   J_{lc,mc} += W * ((alpha * psi_m) * psi_l).
template <typename T>
__global__ __launch_bounds__(256) void traction_face_kernel(ElemArgs<T> a) {
    constexpr int NN = 3, NS = 6;
    const int e = a.e0 + (int)(blockIdx.x * 256 + threadIdx.x);
    if (e >= a.e1) return;
    const T alpha = (T)a.fcoef[3 * (int64_t)e];
    const T beta[2] = {(T)a.fcoef[3 * (int64_t)e + 1], (T)a.fcoef[3 * (int64_t)e + 2]};
    T X[NN][2], U[NN][2];
#pragma unroll
    for (int l = 0; l < NN; l++) {
        const int32_t n = a.conn[(int64_t)l * a.nel + e];
        X[l][0] = (T)a.coords[(int64_t)n * 3];
        X[l][1] = (T)a.coords[(int64_t)n * 3 + 1];
#pragma unroll
        for (int c = 0; c < 2; c++) {
            const int32_t q = a.seqn[(int64_t)(2 * l + c) * a.nel + e];
            U[l][c] = (T)(q >= 0 ? a.dofs[q] : a.pinned[(int64_t)n * a.nvalue + c]);
        }
    }
    T R[NS], J[NN][NN];
#pragma unroll
    for (int i = 0; i < NS; i++) R[i] = T(0);
#pragma unroll
    for (int i = 0; i < NN; i++)
#pragma unroll
        for (int j = 0; j < NN; j++) J[i][j] = T(0);
#pragma unroll
    for (int g = 0; g < NN; g++) {
        double sd, wd;
        FaceRuleQ3::point(g, sd, wd);
        const T sq = (T)sd, wq = (T)wd;
        T psi[NN], dpsi[NN];
        FaceRuleQ3::shape(sq, psi, dpsi);
        T tx = T(0), ty = T(0), u0 = T(0), u1 = T(0);
#pragma unroll
        for (int l = 0; l < NN; l++) {
            tx = tx + X[l][0] * dpsi[l];
            ty = ty + X[l][1] * dpsi[l];
            u0 = u0 + U[l][0] * psi[l];
            u1 = u1 + U[l][1] * psi[l];
        }
        const T W = wq * sqrt(tx * tx + ty * ty);
        const T f[2] = {alpha * u0 + beta[0], alpha * u1 + beta[1]};
#pragma unroll
        for (int l = 0; l < NN; l++) {
#pragma unroll
            for (int c = 0; c < 2; c++) R[2 * l + c] = R[2 * l + c] + W * (f[c] * psi[l]);
#pragma unroll
            for (int m = 0; m < NN; m++) J[l][m] = J[l][m] + W * ((alpha * psi[m]) * psi[l]);
        }
    }
    const bool zero_r = a.deriv != 0;
    const bool zero_j = a.deriv != 0 || a.flag == 2 || a.hvec != nullptr;
#pragma unroll
    for (int i = 0; i < NS; i++) {
        a.Er[(int64_t)e * NS + i] = zero_r ? T(0) : R[i];
        if (a.Ej) {
#pragma unroll
            for (int j = 0; j < NS; j++) {
                // velocity components do not couple through the face
                const T v = (i % 2 == j % 2) ? J[i / 2][j / 2] : T(0);
                a.Ej[((int64_t)e * NS + i) * NS + j] = zero_j ? T(0) : v;
            }
        }
    }
}

// Faces of pyoomph's EnforcedBC(u=constraint) @ boundary (pyoomph/meshes/bcs.py:129-253)
// with constraint = alpha*u + beta per face: the Lagrange multiplier l is value 1 of the
// face nodes, in u's space; slots (u_l, l_l) per face node l (slot 2l, 2l+1, the node-major
// order of the nodes' values). The line elements, rules and geometry are the flux faces'.
// Per point: u = sum u_l psi_l, l = sum l_l psi_l, c = alpha*u + beta;
//   R[u_i] += W * (l * psi_i),               R[l_i] += W * (c * psi_i),
//   J[u_i][l_j] += W * (psi_j * psi_i),      J[l_i][u_j] += W * ((alpha * psi_j) * psi_i),
// (the residuals weak(l, u_test) and weak(constraint, l_test), bcs.py:185-189); the u-u and
// l-l blocks are zero. No time derivative, no dependence on the global parameter and none
// of second order: the mass matrix, dR/dp and the Hessian-vector products vanish.
template <typename T, int NN, typename RULE>
__global__ __launch_bounds__(256) void enforced_face_kernel(ElemArgs<T> a) {
    constexpr int NS = 2 * NN;
    const int e = a.e0 + (int)(blockIdx.x * 256 + threadIdx.x);
    if (e >= a.e1) return;
    const T alpha = (T)a.fcoef[2 * (int64_t)e], beta = (T)a.fcoef[2 * (int64_t)e + 1];
    T X[NN][2], U[NN], Lm[NN];
#pragma unroll
    for (int l = 0; l < NN; l++) {
        const int32_t n = a.conn[(int64_t)l * a.nel + e];
        X[l][0] = (T)a.coords[(int64_t)n * 3];
        X[l][1] = (T)a.coords[(int64_t)n * 3 + 1];
        const int32_t qu = a.seqn[(int64_t)(2 * l) * a.nel + e];
        const int32_t ql = a.seqn[(int64_t)(2 * l + 1) * a.nel + e];
        U[l] = (T)(qu >= 0 ? a.dofs[qu] : a.pinned[(int64_t)n * a.nvalue]);
        Lm[l] = (T)(ql >= 0 ? a.dofs[ql] : a.pinned[(int64_t)n * a.nvalue + 1]);
    }
    T R[NS], J[NS][NS];
#pragma unroll
    for (int i = 0; i < NS; i++) {
        R[i] = T(0);
#pragma unroll
        for (int j = 0; j < NS; j++) J[i][j] = T(0);
    }
#pragma unroll
    for (int g = 0; g < NN; g++) {
        double sd, wd;
        RULE::point(g, sd, wd);
        const T sq = (T)sd, wq = (T)wd;
        T psi[NN], dpsi[NN];
        RULE::shape(sq, psi, dpsi);
        T tx = T(0), ty = T(0), u = T(0), lm = T(0);
#pragma unroll
        for (int l = 0; l < NN; l++) {
            tx = tx + X[l][0] * dpsi[l];
            ty = ty + X[l][1] * dpsi[l];
            u = u + U[l] * psi[l];
            lm = lm + Lm[l] * psi[l];
        }
        const T W = wq * sqrt(tx * tx + ty * ty);
        const T c = alpha * u + beta;
#pragma unroll
        for (int i = 0; i < NN; i++) {
            R[2 * i] = R[2 * i] + W * (lm * psi[i]);
            R[2 * i + 1] = R[2 * i + 1] + W * (c * psi[i]);
#pragma unroll
            for (int j = 0; j < NN; j++) {
                J[2 * i][2 * j + 1] = J[2 * i][2 * j + 1] + W * (psi[j] * psi[i]);
                J[2 * i + 1][2 * j] = J[2 * i + 1][2 * j] + W * ((alpha * psi[j]) * psi[i]);
            }
        }
    }
    const bool zero_r = a.deriv != 0;
    const bool zero_j = a.deriv != 0 || a.flag == 2 || a.hvec != nullptr;
#pragma unroll
    for (int i = 0; i < NS; i++) {
        a.Er[(int64_t)e * NS + i] = zero_r ? T(0) : R[i];
        if (a.Ej) {
#pragma unroll
            for (int j = 0; j < NS; j++) a.Ej[((int64_t)e * NS + i) * NS + j] = zero_j ? T(0) : J[i][j];
        }
    }
}

template <typename T>
static int launch_faces(int kind, const ElemArgs<T> &a, hipStream_t st) {
    if (a.e1 <= a.e0) return PYHIP_OK;
    const dim3 g((a.e1 - a.e0 + 255) / 256), t(256);
    if (kind == kKindLine2) hipLaunchKernelGGL((flux_face_kernel<T, 2, FaceRule<2>>), g, t, 0, st, a);
    else if (kind == kKindLine3) hipLaunchKernelGGL((flux_face_kernel<T, 3, FaceRule<3>>), g, t, 0, st, a);
    else if (kind == kKindLine3Q) hipLaunchKernelGGL((flux_face_kernel<T, 3, FaceRuleQ3>), g, t, 0, st, a);
    else if (kind == kKindQuad9) hipLaunchKernelGGL((flux_face_quad9_kernel<T>), g, t, 0, st, a);
    else if (kind == kKindLine3TH) hipLaunchKernelGGL((traction_face_kernel<T>), g, t, 0, st, a);
    else if (kind == kKindLine2E) hipLaunchKernelGGL((enforced_face_kernel<T, 2, FaceRule<2>>), g, t, 0, st, a);
    else if (kind == kKindLine3E) hipLaunchKernelGGL((enforced_face_kernel<T, 3, FaceRule<3>>), g, t, 0, st, a);
    else if (kind == kKindLine3QE) hipLaunchKernelGGL((enforced_face_kernel<T, 3, FaceRuleQ3>), g, t, 0, st, a);
    else {
        set_error("unknown face element kind");
        return PYHIP_ERR_INVALID;
    }
    PYHIP_CHECK_HIP(hipGetLastError());
    return PYHIP_OK;
}

int launch_faces_f64(int kind, const ElemArgs<double> &a, hipStream_t st) { return launch_faces<double>(kind, a, st); }
int launch_faces_f32(int kind, const ElemArgs<float> &a, hipStream_t st) { return launch_faces<float>(kind, a, st); }

}  // namespace pyhip
