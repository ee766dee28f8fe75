// Batched element residual / Jacobian / parameter-derivative kernels.
//
// Included by elements_f64.hip (built with -ffp-contract=off: the FP64 kernels execute
// exactly the written operation order, the same order as the CPU restatement of the
// generated element code) and elements_f32.hip (FP32, contraction allowed).
//
// What each thread computes follows the code pyoomph generates per element
// (src/codegen.cpp:3912-4109, 2015-2257; accumulation macros src/jitbridge_nohang.h:
// 23-50): integration-point loop, geometry as BulkElementBase::fill_shape_info_at_s
// (src/elements.cpp:3593-3700, 3800-3840, 3960-3990) with weight*det
// (src/elements.cpp:4564-4575), field interpolation, then the test-function row and its
// shape-function columns, skipping pinned equations (local_eqn < 0).
//
// Mapping (one workgroup = a block of elements, 64-wide waves):
//   thread = (element, slot row i) in phase B and (element, integration point) in
//   phase A; EPB = 256 / max(NS, points per chunk) elements per workgroup.
//   Per element the nodal coordinates, slot values, equations and auxiliary field are
//   staged once in LDS (coalesced slot-major gathers). Per integration point, one thread
//   computes the geometry (tangents, metric inverse, w*det, dx_shape of every node) and
//   the interpolated fields once into LDS; then every thread runs its residual entry and
//   Jacobian row over the points, reading them as LDS broadcasts.
//   Reference-element tables (weights, psi, dpsi/ds) are read with uniform indices, so
//   they come through the scalar cache.
// Output: element vectors Er[e][i] and matrices Ej[e][i][j] in slot numbering; the
// deterministic gather (scatter.hip) sums them into the residual and the CSR values.
#pragma once
#include "common.h"
#include "element_common.h"

namespace pyhip {

// Taylor-Hood Q2/Q1 slot layout = oomph local dof order (node-major, value-minor):
// node 0: ux uy p | 1: ux uy | 2: ux uy p | 3,4,5: ux uy | 6: ux uy p | 7: ux uy | 8: ux uy p
__device__ __constant__ const int8_t kTHSlotNode[22] = {0, 0, 0, 1, 1, 2, 2, 2, 3, 3, 4,
                                                        4, 5, 5, 6, 6, 6, 7, 7, 8, 8, 8};
__device__ __constant__ const int8_t kTHSlotKind[22] = {0, 1, 2, 0, 1, 0, 1, 2, 0, 1, 0,
                                                        1, 0, 1, 0, 1, 2, 0, 1, 0, 1, 2};
__device__ __constant__ const int8_t kTHSlotP[22] = {-1, -1, 0,  -1, -1, -1, -1, 1,  -1, -1, -1,
                                                     -1, -1, -1, -1, -1, 2,  -1, -1, -1, -1, 3};
// velocity-x slot of each node (velocity-y is the next slot)
__device__ __constant__ const int8_t kTHNodeSlot[9] = {0, 3, 5, 8, 10, 12, 14, 17, 19};

// Scalar weak forms (Poisson / Bratu / reaction-diffusion): slot = node.
// Phase A (per chunk of integration points): thread t of an element computes points
// q = t, t + NS, ... of the chunk: geometry, dx_shape of all nodes, u, grad u and the
// source term and its derivative. Phase B: thread = test row i, accumulates its residual
// and Jacobian row over the chunk's points in increasing point order.

// TR: the form has a time derivative (a.transient); HV: the matrix output is the
// Hessian-vector product sum_k dJ_ij/du_k y_k = int d2src (sum_k y_k psi_k) psi_j psi_i
// (d2src: Bratu -lambda e^u, reaction-diffusion 6u, Poisson 0), symmetric in i, j and k,
// so the transposed product (a.htr) is the same matrix. Separate instantiations keep the
// steady kernels' LDS and registers.
template <typename T, int KIND, int FORM, bool TR, bool HV>
__global__ __launch_bounds__(256) void scalar_element_kernel(ElemArgs<T> a) {
    using Tr = ElemTraits<KIND>;
    constexpr int NN = Tr::NN, DIM = Tr::DIM, NQ = Tr::NQ, NS = NN;
    constexpr int EPB = scalar_epb<KIND>();
    constexpr int QC = ScalarChunk<KIND>::QC;
    // per point: src, dsrc, grad u[, du/dt][, d2src, y]
    constexpr int NQV = 2 + DIM + (TR ? 1 : 0) + (HV ? 2 : 0);
    constexpr int QT = 2 + DIM, QH = 2 + DIM + (TR ? 1 : 0);
    __shared__ T sX[EPB][NN][DIM];
    __shared__ T sU[EPB][NN];
    __shared__ T sH[TR ? EPB : 1][NN];   // history combination tw1 u(t-1) + tw2 u(t-2) per node
    __shared__ T sY[HV ? EPB : 1][NN];   // Hessian-vector product direction per node
    __shared__ T sF[FORM == PYHIP_FORM_REACTION_DIFFUSION ? EPB : 1][NN];
    __shared__ int sEq[EPB][NS];
    __shared__ T sW[EPB][QC];
    __shared__ T sDx[EPB][QC][NN][DIM];
    __shared__ T sQ[EPB][QC][NQV];

    const int le = threadIdx.x / NS;        // element within block
    const int row = threadIdx.x % NS;       // slot row of this thread
    const int e = a.e0 + (int)blockIdx.x * EPB + le;
    const bool active = (le < EPB) && (e < a.e1);

    if (active) {
        const int node = a.conn[(int64_t)row * a.nel + e];
        const int eq = a.seqn[(int64_t)row * a.nel + e];
#pragma unroll
        for (int d = 0; d < DIM; d++) sX[le][row][d] = T(a.coords[(int64_t)node * 3 + d]);
        sU[le][row] = T(eq >= 0 ? a.dofs[eq] : a.pinned[(int64_t)node * a.nvalue]);
        if constexpr (FORM == PYHIP_FORM_REACTION_DIFFUSION) sF[le][row] = T(a.aux[node]);
        if constexpr (TR) {
            const double pin = a.pinned[(int64_t)node * a.nvalue];
            const double u1 = eq >= 0 ? a.hist1[eq] : pin, u2 = eq >= 0 ? a.hist2[eq] : pin;
            sH[le][row] = T(a.tw1 * u1 + a.tw2 * u2);
        }
        if constexpr (HV) sY[le][row] = T(eq >= 0 ? a.hvec[eq] : 0.0);
        sEq[le][row] = eq;
    }
    __syncthreads();

    // phase-A thread: point qa of the current chunk of element la
    const int la = threadIdx.x / QC, qa = threadIdx.x % QC;
    const int ea = a.e0 + (int)blockIdx.x * EPB + la;
    const bool active_a = (la < EPB) && (ea < a.e1);
    T par0 = T(0);
    const double *par = nullptr;
    if (active_a) {
        par = a.params + (int64_t)a.sub[ea] * a.pstride;
        par0 = T(par[0]);
    }
    const bool myrow = active && sEq[le][row] >= 0;
    T Racc = T(0);
    T Jacc[NS];
#pragma unroll
    for (int j = 0; j < NS; j++) Jacc[j] = T(0);

    for (int q0 = 0; q0 < NQ; q0 += QC) {
        // ---- phase A: per-point quantities (element la, point q0 + qa)
        if (active_a && q0 + qa < NQ) {
            {
                const int le = la, qq = qa;
                const int q = q0 + qq;
                const T *D = a.tab->dpsi + q * NN * DIM;
                const T *P = a.tab->psi + q * NN;
                const T W = point_geometry<T, NN, DIM>(sX[le], D, a.tab->w[q], sDx[le][qq]);
                sW[le][qq] = W;
                T u = T(0);
#pragma unroll
                for (int l = 0; l < NN; l++) u = u + sU[le][l] * P[l];
#pragma unroll
                for (int i = 0; i < DIM; i++) {
                    T sacc = T(0);
#pragma unroll
                    for (int l = 0; l < NN; l++) sacc = sacc + sU[le][l] * sDx[le][qq][l][i];
                    sQ[le][qq][2 + i] = sacc;
                }
                T src = T(0), dsrc = T(0);
                if constexpr (FORM == PYHIP_FORM_POISSON) {
                    T x = T(0), y = T(0);
#pragma unroll
                    for (int l = 0; l < NN; l++) x = x + sX[le][l][0] * P[l];
#pragma unroll
                    for (int l = 0; l < NN; l++) y = y + sX[le][l][1] * P[l];
                    T f = T(0);
                    const int ng = (int)par[0];  // Gaussians follow as (cx, cy, w, a)
                    for (int k = 1; k < 1 + 4 * ng; k += 4) {
                        T ddx = x - T(par[k + 0]), ddy = y - T(par[k + 1]);
                        T r2 = ddx * ddx + ddy * ddy;
                        T wk = T(par[k + 2]);
                        f = f + T(par[k + 3]) * texp(-r2 / (wk * wk));
                    }
                    src = -f;
                } else if constexpr (FORM == PYHIP_FORM_BRATU) {
                    T ev = texp(u);
                    if (a.deriv) {
                        src = -ev;
                        dsrc = -ev;
                    } else {
                        src = -par0 * ev;
                        dsrc = -par0 * ev;
                    }
                } else {
                    T f = T(0);
#pragma unroll
                    for (int l = 0; l < NN; l++) f = f + sF[le][l] * P[l];
                    src = u * u * u - f;
                    dsrc = T(3) * (u * u);
                }
                sQ[le][qq][0] = src;
                sQ[le][qq][1] = dsrc;
                if constexpr (TR) {
                    T hq = T(0);
#pragma unroll
                    for (int l = 0; l < NN; l++) hq = hq + sH[le][l] * P[l];
                    sQ[le][qq][QT] = T(a.tw0) * u + hq;
                }
                if constexpr (HV) {
                    T yq = T(0);
#pragma unroll
                    for (int l = 0; l < NN; l++) yq = yq + sY[le][l] * P[l];
                    T d2 = T(0);
                    if constexpr (FORM == PYHIP_FORM_BRATU) d2 = -par0 * texp(u);
                    else if constexpr (FORM == PYHIP_FORM_REACTION_DIFFUSION) d2 = T(6) * u;
                    sQ[le][qq][QH] = d2;
                    sQ[le][qq][QH + 1] = yq;
                }
            }
        }
        __syncthreads();
        // ---- phase B: this thread's residual entry and Jacobian row
        if (myrow) {
            for (int qq = 0; qq < QC && q0 + qq < NQ; qq++) {
                const int q = q0 + qq;
                const T *P = a.tab->psi + q * NN;
                const T W = sW[le][qq];
                const T *dxi = sDx[le][qq][row];
                const T src = sQ[le][qq][0], dsrc = sQ[le][qq][1];
                const bool tr = TR && !a.deriv;
                if (a.deriv)
                    Racc = Racc + W * (src * P[row]);
                else if (tr)
                    Racc = Racc + W * ((dotv<T, DIM>(&sQ[le][qq][2], dxi) + src * P[row]) +
                                       sQ[le][qq][QT] * P[row]);
                else
                    Racc = Racc + W * (dotv<T, DIM>(&sQ[le][qq][2], dxi) + src * P[row]);
                if (a.flag == 2 && !TR) continue;   // no time derivative: mass matrix 0
                if (a.flag) {
                    const T tw0 = T(a.tw0);
#pragma unroll
                    for (int j = 0; j < NS; j++) {
                        if (sEq[le][j] < 0) continue;
                        const T *dxj = sDx[le][qq][j];
                        T v;
                        if constexpr (HV) {
                            v = W * (((sQ[le][qq][QH] * sQ[le][qq][QH + 1]) * P[j]) * P[row]);
                        } else if (a.flag == 2) {
                            v = W * (P[j] * P[row]);
                        } else if (a.deriv) {
                            v = W * ((dsrc * P[j]) * P[row]);
                        } else {
                            T inner;
                            if constexpr (FORM == PYHIP_FORM_POISSON)
                                inner = dotv<T, DIM>(dxj, dxi);
                            else
                                inner = dotv<T, DIM>(dxj, dxi) + (dsrc * P[j]) * P[row];
                            v = tr ? W * (inner + (tw0 * P[j]) * P[row]) : W * inner;
                        }
                        Jacc[j] = Jacc[j] + v;
                    }
                }
            }
        }
        if (q0 + QC < NQ) __syncthreads();
    }
    if (myrow) {
        a.Er[(int64_t)e * NS + row] = Racc;
        if (a.flag) {
            T *dst = a.Ej + ((int64_t)e * NS + row) * NS;
#pragma unroll
            for (int j = 0; j < NS; j++) dst[j] = Jacc[j];
        }
    }
}

// Taylor-Hood Q2/Q1 steady Navier-Stokes (22 slots, 9 geometric nodes). Phase A: thread
// t < 9 of an element computes integration point t (geometry, dx_shape, u, grad u, p);
// phase B: thread = slot row.
// TR: time derivative of the velocities (a.transient): du/dt = tw0 u + sum_l H psi_l with
// H = tw1 u(t-1) + tw2 u(t-2) per slot; momentum rows gain (rho du/dt) psi_i, the diagonal
// velocity blocks tw0 rho psi_j psi_i; flag 2 writes the mass matrix rho psi_j psi_i there.
template <typename T, bool TR, bool HV = false>
__global__ __launch_bounds__(256) void th_element_kernel(ElemArgs<T> a) {
    constexpr int NN = 9, DIM = 2, NQ = 9, NS = 22;
    constexpr int EPB = 256 / NS;  // 11 elements, 242 threads
    // per point: u0, u1, du00, du01, du10, du11, p[, du0/dt, du1/dt][, y0, y1, dy00, dy01, dy10, dy11]
    constexpr int QY = TR ? 9 : 7;
    __shared__ T sX[EPB][NN][DIM];
    __shared__ T sV[EPB][NS];
    __shared__ int sEq[EPB][NS];
    __shared__ T sW[EPB][NQ];
    __shared__ T sDx[EPB][NQ][NN][DIM];
    __shared__ T sQ[EPB][NQ][QY + (HV ? 6 : 0)];
    __shared__ T sH[TR ? EPB : 1][NS];
    __shared__ T sY[HV ? EPB : 1][NS];   // Hessian-vector product direction per slot

    const int le = threadIdx.x / NS;
    const int slot = threadIdx.x % NS;
    const int e = a.e0 + (int)blockIdx.x * EPB + le;
    const bool active = (le < EPB) && (e < a.e1);
    if (active) {
        const int ln = kTHSlotNode[slot];
        const int kind = kTHSlotKind[slot];
        const int node = a.conn[(int64_t)ln * a.nel + e];
        const int eq = a.seqn[(int64_t)slot * a.nel + e];
        if (slot < NN) {
            const int gn = a.conn[(int64_t)slot * a.nel + e];
            sX[le][slot][0] = T(a.coords[(int64_t)gn * 3 + 0]);
            sX[le][slot][1] = T(a.coords[(int64_t)gn * 3 + 1]);
        }
        sV[le][slot] = T(eq >= 0 ? a.dofs[eq] : a.pinned[(int64_t)node * a.nvalue + kind]);
        if constexpr (TR) {
            const double pin = a.pinned[(int64_t)node * a.nvalue + kind];
            const double u1 = eq >= 0 ? a.hist1[eq] : pin, u2 = eq >= 0 ? a.hist2[eq] : pin;
            sH[le][slot] = T(a.tw1 * u1 + a.tw2 * u2);
        }
        if constexpr (HV) sY[le][slot] = T(eq >= 0 ? a.hvec[eq] : 0.0);
        sEq[le][slot] = eq;
    }
    __syncthreads();
    T mu = T(0), rho = T(0);
    if (active) {
        const double *par = a.params + (int64_t)a.sub[e] * a.pstride;
        mu = T(par[0]);
        rho = T(par[1]);
    }
    if (active && slot < NQ) {
        const int q = slot;
        const T *D = a.tab->dpsi + q * NN * DIM;
        const T *P = a.tab->psi + q * NN;
        const T *C = a.tab->psi1 + q * 4;
        sW[le][q] = point_geometry<T, NN, DIM>(sX[le], D, a.tab->w[q], sDx[le][q]);
        T u0 = T(0), u1 = T(0), p = T(0);
#pragma unroll
        for (int l = 0; l < NN; l++) u0 = u0 + sV[le][kTHNodeSlot[l]] * P[l];
#pragma unroll
        for (int l = 0; l < NN; l++) u1 = u1 + sV[le][1 + kTHNodeSlot[l]] * P[l];
#pragma unroll
        for (int b = 0; b < 2; b++) {
            T s0 = T(0), s1 = T(0);
#pragma unroll
            for (int l = 0; l < NN; l++) s0 = s0 + sV[le][kTHNodeSlot[l]] * sDx[le][q][l][b];
#pragma unroll
            for (int l = 0; l < NN; l++) s1 = s1 + sV[le][kTHNodeSlot[l] + 1] * sDx[le][q][l][b];
            sQ[le][q][2 + b] = s0;       // du[0][b]
            sQ[le][q][4 + b] = s1;       // du[1][b]
        }
        // pressure slots 2, 7, 16, 21 hold corner nodes 0, 2, 6, 8
        p = p + sV[le][2] * C[0];
        p = p + sV[le][7] * C[1];
        p = p + sV[le][16] * C[2];
        p = p + sV[le][21] * C[3];
        sQ[le][q][0] = u0;
        sQ[le][q][1] = u1;
        sQ[le][q][6] = p;
        if constexpr (TR) {
            T h0 = T(0), h1 = T(0);
#pragma unroll
            for (int l = 0; l < NN; l++) h0 = h0 + sH[le][kTHNodeSlot[l]] * P[l];
#pragma unroll
            for (int l = 0; l < NN; l++) h1 = h1 + sH[le][kTHNodeSlot[l] + 1] * P[l];
            sQ[le][q][7] = T(a.tw0) * u0 + h0;
            sQ[le][q][8] = T(a.tw0) * u1 + h1;
        }
        if constexpr (HV) {
            T y0 = T(0), y1 = T(0);
#pragma unroll
            for (int l = 0; l < NN; l++) y0 = y0 + sY[le][kTHNodeSlot[l]] * P[l];
#pragma unroll
            for (int l = 0; l < NN; l++) y1 = y1 + sY[le][1 + kTHNodeSlot[l]] * P[l];
#pragma unroll
            for (int b = 0; b < 2; b++) {
                T s0 = T(0), s1 = T(0);
#pragma unroll
                for (int l = 0; l < NN; l++) s0 = s0 + sY[le][kTHNodeSlot[l]] * sDx[le][q][l][b];
#pragma unroll
                for (int l = 0; l < NN; l++) s1 = s1 + sY[le][kTHNodeSlot[l] + 1] * sDx[le][q][l][b];
                sQ[le][q][QY + 2 + b] = s0;   // dy[0][b]
                sQ[le][q][QY + 4 + b] = s1;   // dy[1][b]
            }
            sQ[le][q][QY] = y0;
            sQ[le][q][QY + 1] = y1;
        }
    }
    __syncthreads();
    const bool myrow = active && sEq[le][slot] >= 0;
    if (!myrow) return;
    const int kt = kTHSlotKind[slot], lt = kTHSlotNode[slot];
    T Racc = T(0);
    T Jacc[NS];
#pragma unroll
    for (int j = 0; j < NS; j++) Jacc[j] = T(0);
    for (int q = 0; q < NQ; q++) {
        const T *P = a.tab->psi + q * NN;
        const T *C = a.tab->psi1 + q * 4;
        const T W = sW[le][q];
        const T u0 = sQ[le][q][0], u1 = sQ[le][q][1], p = sQ[le][q][6];
        T du[2][2];
        du[0][0] = sQ[le][q][2];
        du[0][1] = sQ[le][q][3];
        du[1][0] = sQ[le][q][4];
        du[1][1] = sQ[le][q][5];
        T sig[2][2];
#pragma unroll
        for (int aa = 0; aa < 2; aa++)
#pragma unroll
            for (int b = 0; b < 2; b++) {
                T s = mu * (du[aa][b] + du[b][aa]);
                if (aa == b) s = s - p;
                sig[aa][b] = s;
            }
        const T *dxl = sDx[le][q][lt];
        if (kt < 2) {
            const T conv = rho * (u0 * du[kt][0] + u1 * du[kt][1]);
            if constexpr (TR)
                Racc = Racc + W * (((sig[kt][0] * dxl[0] + sig[kt][1] * dxl[1]) + conv * P[lt]) +
                                   (rho * sQ[le][q][7 + kt]) * P[lt]);
            else
                Racc = Racc + W * ((sig[kt][0] * dxl[0] + sig[kt][1] * dxl[1]) + conv * P[lt]);
        } else {
            const int pidx = kTHSlotP[slot];
            Racc = Racc + W * ((du[0][0] + du[1][1]) * C[pidx]);
        }
        if (a.flag == 2 && !TR) continue;   // no time derivative: mass matrix 0
        if (a.flag) {
#pragma unroll
            for (int ss = 0; ss < NS; ss++) {
                if (sEq[le][ss] < 0) continue;
                const int ks = kTHSlotKind[ss], ls = kTHSlotNode[ss];
                T v;
                if constexpr (HV) {
                    if (kt < 2 && ks < 2) {
                        const T *dxs = sDx[le][q][ls];
                        if (a.htr) {
                            // Hessian of y . R (row kt,lt; column ks,ls): only the convection
                            // term is not linear, rho (psi_lt d_kt psi_ls y_ks + psi_ls d_ks psi_lt y_kt)
                            const T cvt = P[lt] * dxs[kt] * sQ[le][q][QY + ks] + P[ls] * dxl[ks] * sQ[le][q][QY + kt];
                            v = W * (rho * cvt);
                        } else {
                            // d/du of the convection block in the direction y: the block with u -> y
                            T cvy = P[ls] * sQ[le][q][QY + 2 + 2 * kt + ks];
                            if (kt == ks) cvy = cvy + (sQ[le][q][QY] * dxs[0] + sQ[le][q][QY + 1] * dxs[1]);
                            v = W * ((rho * cvy) * P[lt]);
                        }
                    } else {
                        continue;
                    }
                } else if (a.flag == 2) {
                    if (kt < 2 && kt == ks) v = W * ((rho * P[ls]) * P[lt]);
                    else continue;
                } else if (kt < 2 && ks < 2) {
                    const T *dxs = sDx[le][q][ls];
                    T visc = dxs[kt] * dxl[ks];
                    T cv = P[ls] * du[kt][ks];
                    if (kt == ks) {
                        visc = (dxs[0] * dxl[0] + dxs[1] * dxl[1]) + visc;
                        cv = cv + (u0 * dxs[0] + u1 * dxs[1]);
                    }
                    if (TR && kt == ks)
                        v = W * ((mu * visc + (rho * cv) * P[lt]) + ((rho * T(a.tw0)) * P[ls]) * P[lt]);
                    else
                        v = W * (mu * visc + (rho * cv) * P[lt]);
                } else if (kt < 2 && ks == 2) {
                    const int pidx = kTHSlotP[ss];
                    v = W * (-(C[pidx] * dxl[kt]));
                } else if (kt == 2 && ks < 2) {
                    const int pidx = kTHSlotP[slot];
                    v = W * (sDx[le][q][ls][ks] * C[pidx]);
                } else {
                    continue;
                }
                Jacc[ss] = Jacc[ss] + v;
            }
        }
    }
    a.Er[(int64_t)e * NS + slot] = Racc;
    if (a.flag) {
        T *dst = a.Ej + ((int64_t)e * NS + slot) * NS;
#pragma unroll
        for (int j = 0; j < NS; j++) dst[j] = Jacc[j];
    }
}

}  // namespace pyhip
