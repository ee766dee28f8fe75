/*
 * CPU oracle: element residual / Jacobian / parameter-derivative functions and the
 * vectors_of_pairs CSR builder.
 *
 * TEST INFRASTRUCTURE ONLY. Used by tests/, __graft_entry__.smoke() and bench.py's CPU
 * baseline, as the checker. The product library never links it.
 *
 * Element functions restate what pyoomph's code generator emits for each weak form
 * (src/codegen.cpp:3912-4109 qp loop + test/shape loops 2015-2257, accumulation macros
 * src/jitbridge_nohang.h:23-50): a loop over integration points that evaluates the
 * geometry the way BulkElementBase::fill_shape_info_at_s does (src/elements.cpp:
 * 3593-3700 tangents and metric, 3800-3840 3D inverse, 3960-3990 dx_shape, 4564-4575
 * weight*det), interpolates fields, then loops l_test with local_eqn >= 0 and, when
 * flag >= 1, l_shape with local_unknown >= 0, accumulating with "+=".
 *
 * Arithmetic order is fixed and written out operation by operation; this file is built
 * with -ffp-contract=off so that no fused multiply-add changes it.
 *
 * Weak forms (steady):
 *   form 1  Poisson      R_i = sum_q W (grad u . grad psi_i - f(x) psi_i)
 *           (pyoomph/equations/poisson.py:70-74), f = sum_k a_k exp(-|x-c_k|^2/w_k^2)
 *   form 2  Bratu        R_i = sum_q W (grad u . grad psi_i - lambda e^u psi_i)
 *   form 3  reaction-diffusion  R_i = sum_q W (grad u . grad psi_i + (u^3 - f) psi_i),
 *           f a nodal field
 *   form 4  Navier-Stokes Taylor-Hood (pyoomph/equations/navier_stokes.py:303-347,
 *           481-503): momentum  sum_q W (sigma : grad v + rho (u.grad u) . v),
 *           sigma = mu (grad u + grad u^T) - p I ; continuity sum_q W (div u) q.
 * Parity: pinned by tests/test_oracle_elements.py (textbook P1/P2/Q2 stiffness matrices
 * on affine elements, FD of the residual vs the Jacobian, FD in lambda vs dR/dlambda,
 * manufactured-solution convergence orders, brute-force dense assembly).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define MAXN 27
#define MAXD 3
#define MAXSLOT 27

typedef struct {
    int dim, nnode, nqp, nnode1;
    const double *w;     /* [nqp] */
    const double *psi;   /* [nqp][nnode] */
    const double *dpsi;  /* [nqp][nnode][dim] */
    const double *psi1;  /* [nqp][nnode1] pressure sub-basis (Taylor-Hood) */
} ref_t;

/* Geometry at one integration point, as BulkElementBase::fill_shape_info_at_s. */
static double geometry(const ref_t *r, int q, const double *X, double dx[MAXN][MAXD])
{
    const int dim = r->dim, nn = r->nnode;
    const double *D = r->dpsi + (size_t)q * nn * dim;
    double t[MAXD][MAXD];
    for (int a = 0; a < dim; a++)
        for (int i = 0; i < dim; i++) t[a][i] = 0.0;
    for (int l = 0; l < nn; l++)
        for (int i = 0; i < dim; i++)
            for (int j = 0; j < dim; j++) t[j][i] = t[j][i] + X[l * dim + i] * D[l * dim + j];
    double am[MAXD][MAXD] = {{0.0}};
    for (int a = 0; a < dim; a++)
        for (int b = 0; b < dim; b++) {
            double s = 0.0;
            for (int i = 0; i < dim; i++) s = s + t[a][i] * t[b][i];
            am[a][b] = s;
        }
    double au[MAXD][MAXD], det_a;
    if (dim == 2) {
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
    double g[MAXD][MAXD]; /* g[b][i] = [g^{ab} g_a]_i */
    for (int b = 0; b < dim; b++)
        for (int i = 0; i < dim; i++) {
            double s = au[0][b] * t[0][i] + au[1][b] * t[1][i];
            if (dim == 3) s = s + au[2][b] * t[2][i];
            g[b][i] = s;
        }
    for (int l = 0; l < nn; l++)
        for (int i = 0; i < dim; i++) {
            double s = 0.0;
            for (int b = 0; b < dim; b++) s = s + g[b][i] * D[l * dim + b];
            dx[l][i] = s;
        }
    return r->w[q] * sqrt(det_a);
}

static double dot(int dim, const double *a, const double *b)
{
    double s = a[0] * b[0] + a[1] * b[1];
    if (dim == 3) s = s + a[2] * b[2];
    return s;
}

/* Scalar forms 1-3. Slots are the nodes.
 * Time derivative (H non-NULL): d/dt u = tw0 * u + H, H the history combination of each
 * node (tw1 * u(t-1) + tw2 * u(t-2), BDF weights of MultiTimeStepper::set_weights,
 * src/timestepper.cpp:32-59, selected as in prepare_shape_buffer_for_integration,
 * src/elements.cpp:4583-4636); the weak form gains (du/dt) * psi_i. flag 2 writes the mass
 * matrix dR_i/d(du_j/dt) = int psi_j psi_i into J instead of the Jacobian (zero without a
 * time derivative). */
/* Hessian-vector product (Y non-NULL, the element's slot values of the vector):
 * J receives sum_k d^2 R_i/(du_j du_k) Y_k = int d2src (sum_k Y_k psi_k) psi_j psi_i, the
 * product of the reference's HessianVectorProduct (src/codegen.cpp:3646-3910, called from
 * src/elements.cpp:4964-4995 with flag 1); d2src is the second u-derivative of the
 * source term (Bratu -lambda e^u, reaction-diffusion 6u, Poisson 0). flag 4 asks for the
 * transposed product sum_j Y_j d^2 R_j/(du_i du_k) (jitbridge.h:678-691) = int d2src
 * (sum_j Y_j psi_j) psi_i psi_k: the same integrand, so the same matrix. */
static void scalar_element(int form, const ref_t *r, const double *X, const double *U,
                           const double *aux, const long *leqn, const double *par, int npar,
                           int flag, int deriv, int nvar, const double *H, double tw0, const double *Y,
                           double *R, double *J)
{
    const int dim = r->dim, nn = r->nnode;
    double dx[MAXN][MAXD];
    for (int q = 0; q < r->nqp; q++) {
        const double W = geometry(r, q, X, dx);
        const double *P = r->psi + (size_t)q * nn;
        double u = 0.0, gu[MAXD] = {0.0, 0.0, 0.0};
        for (int l = 0; l < nn; l++) u = u + U[l] * P[l];
        for (int i = 0; i < dim; i++) {
            double s = 0.0;
            for (int l = 0; l < nn; l++) s = s + U[l] * dx[l][i];
            gu[i] = s;
        }
        double src = 0.0, dsrc = 0.0; /* residual source term and its derivative in u */
        if (form == 1) {
            double x = 0.0, y = 0.0;
            for (int l = 0; l < nn; l++) x = x + X[l * dim + 0] * P[l];
            for (int l = 0; l < nn; l++) y = y + X[l * dim + 1] * P[l];
            double f = 0.0;
            for (int k = 0; k + 3 < npar; k += 4) {
                double ddx = x - par[k + 0], ddy = y - par[k + 1];
                double r2 = ddx * ddx + ddy * ddy;
                f = f + par[k + 3] * exp(-r2 / (par[k + 2] * par[k + 2]));
            }
            src = -f;
        } else if (form == 2) {
            double e = exp(u);
            src = -par[0] * e;
            dsrc = -par[0] * e;
            if (deriv) src = -e;
            if (deriv) dsrc = -e;
        } else if (form == 3) {
            double f = 0.0;
            for (int l = 0; l < nn; l++) f = f + aux[l] * P[l];
            src = u * u * u - f;
            dsrc = 3.0 * (u * u);
        }
        double d2src = 0.0, yq = 0.0; /* Hessian-vector product */
        if (Y) {
            for (int l = 0; l < nn; l++) yq = yq + Y[l] * P[l];
            if (form == 2) d2src = -par[0] * exp(u);
            else if (form == 3) d2src = 6.0 * u;
        }
        double ud = 0.0; /* du/dt at the point */
        if (H) {
            double hq = 0.0;
            for (int l = 0; l < nn; l++) hq = hq + H[l] * P[l];
            ud = tw0 * u + hq;
        }
        for (int lt = 0; lt < nn; lt++) {
            long ie = leqn[lt];
            if (ie < 0) continue;
            if (deriv)
                R[ie] = R[ie] + W * (src * P[lt]);
            else if (H)
                R[ie] = R[ie] + W * ((dot(dim, gu, dx[lt]) + src * P[lt]) + ud * P[lt]);
            else
                R[ie] = R[ie] + W * (dot(dim, gu, dx[lt]) + src * P[lt]);
            if (!flag) continue;
            if (flag == 2 && !H) continue;
            for (int ls = 0; ls < nn; ls++) {
                long iu = leqn[ls];
                if (iu < 0) continue;
                double v;
                if (Y)
                    v = W * (((d2src * yq) * P[ls]) * P[lt]);
                else if (flag == 2)
                    v = W * (P[ls] * P[lt]);
                else if (deriv)
                    v = W * ((dsrc * P[ls]) * P[lt]);
                else {
                    double inner;
                    if (form == 1)
                        inner = dot(dim, dx[ls], dx[lt]);
                    else
                        inner = dot(dim, dx[ls], dx[lt]) + (dsrc * P[ls]) * P[lt];
                    if (H)
                        v = W * (inner + (tw0 * P[ls]) * P[lt]);
                    else
                        v = W * inner;
                }
                J[ie * nvar + iu] = J[ie * nvar + iu] + v;
            }
        }
    }
}

/* Taylor-Hood Q2/Q1. Slot order (oomph local dof order, node-major, value-minor):
 * node 0: ux uy p | 1: ux uy | 2: ux uy p | 3,4,5: ux uy | 6: ux uy p | 7: ux uy | 8: ux uy p */
static const int TH_slot_node[22] = {0, 0, 0, 1, 1, 2, 2, 2, 3, 3, 4, 4, 5, 5, 6, 6, 6, 7, 7, 8, 8, 8};
static const int TH_slot_kind[22] = {0, 1, 2, 0, 1, 0, 1, 2, 0, 1, 0, 1, 0, 1, 0, 1, 2, 0, 1, 0, 1, 2};
static const int TH_slot_pidx[22] = {-1, -1, 0, -1, -1, -1, -1, 1, -1, -1, -1, -1, -1, -1, -1, -1, 2, -1, -1, -1, -1, 3};

/* Time derivative (H non-NULL, H[slot] = w1 v(t-1) + w2 v(t-2)): du_a/dt = w0 u_a + sum_l
 * H_a[l] psi_l; the momentum rows gain (rho du_a/dt) psi_i and the Jacobian w0 rho psi_j psi_i
 * on the diagonal velocity blocks; flag 2 writes the mass matrix rho psi_j psi_i there.
 * Hessian-vector product (Y non-NULL, slot values of the vector, flag 1): only the
 * convection term rho (u . grad) u_a is not linear, so J receives its Jacobian with the
 * velocity u replaced by y: rho (psi_j d_c y_a + delta_ac (y . grad psi_j)) psi_i on the
 * velocity-velocity blocks, nothing elsewhere (src/codegen.cpp:3646-3910). flag 4 (with Y):
 * the transposed product, entry (i, k) = sum_j Y_j d^2 R_j/(du_i du_k) (jitbridge.h:678-691),
 * the Hessian of int rho ((u . grad) u) . y: for row (a, psi_i) and column (c, psi_j)
 * rho (psi_i d_a psi_j y_c + psi_j d_c psi_i y_a). */
static void th_element(const ref_t *r, const double *X, const double *V, const long *leqn,
                       const double *par, int flag, int nvar, const double *H, double tw0, const double *Y,
                       double *R, double *J)
{
    const double mu = par[0], rho = par[1];
    const int nn = 9;
    double dx[MAXN][MAXD];
    double Ux[9], Uy[9], Pp[4], Hx[9] = {0}, Hy[9] = {0}, Yx[9] = {0}, Yy[9] = {0};
    for (int s = 0; s < 22; s++) {
        int k = TH_slot_kind[s], n = TH_slot_node[s];
        if (k == 0) Ux[n] = V[s];
        else if (k == 1) Uy[n] = V[s];
        else Pp[TH_slot_pidx[s]] = V[s];
        if (H && k == 0) Hx[n] = H[s];
        if (H && k == 1) Hy[n] = H[s];
        if (Y && k == 0) Yx[n] = Y[s];
        if (Y && k == 1) Yy[n] = Y[s];
    }
    for (int q = 0; q < r->nqp; q++) {
        const double W = geometry(r, q, X, dx);
        const double *P = r->psi + (size_t)q * nn;
        const double *C = r->psi1 + (size_t)q * 4;
        double u[2] = {0.0, 0.0}, du[2][2] = {{0.0, 0.0}, {0.0, 0.0}}, p = 0.0;
        for (int l = 0; l < nn; l++) u[0] = u[0] + Ux[l] * P[l];
        for (int l = 0; l < nn; l++) u[1] = u[1] + Uy[l] * P[l];
        for (int b = 0; b < 2; b++) {
            double s0 = 0.0, s1 = 0.0;
            for (int l = 0; l < nn; l++) s0 = s0 + Ux[l] * dx[l][b];
            for (int l = 0; l < nn; l++) s1 = s1 + Uy[l] * dx[l][b];
            du[0][b] = s0;
            du[1][b] = s1;
        }
        for (int c = 0; c < 4; c++) p = p + Pp[c] * C[c];
        double sig[2][2];
        for (int a = 0; a < 2; a++)
            for (int b = 0; b < 2; b++) {
                double s = mu * (du[a][b] + du[b][a]);
                if (a == b) s = s - p;
                sig[a][b] = s;
            }
        double conv[2];
        for (int a = 0; a < 2; a++) conv[a] = rho * (u[0] * du[a][0] + u[1] * du[a][1]);
        const double divu = du[0][0] + du[1][1];
        double yv[2] = {0.0, 0.0}, dy[2][2] = {{0.0, 0.0}, {0.0, 0.0}};
        if (Y) {
            for (int l = 0; l < nn; l++) yv[0] = yv[0] + Yx[l] * P[l];
            for (int l = 0; l < nn; l++) yv[1] = yv[1] + Yy[l] * P[l];
            for (int b = 0; b < 2; b++) {
                double s0 = 0.0, s1 = 0.0;
                for (int l = 0; l < nn; l++) s0 = s0 + Yx[l] * dx[l][b];
                for (int l = 0; l < nn; l++) s1 = s1 + Yy[l] * dx[l][b];
                dy[0][b] = s0;
                dy[1][b] = s1;
            }
        }
        double ud[2] = {0.0, 0.0};
        if (H) {
            double h0 = 0.0, h1 = 0.0;
            for (int l = 0; l < nn; l++) h0 = h0 + Hx[l] * P[l];
            for (int l = 0; l < nn; l++) h1 = h1 + Hy[l] * P[l];
            ud[0] = tw0 * u[0] + h0;
            ud[1] = tw0 * u[1] + h1;
        }
        for (int st = 0; st < 22; st++) {
            long ie = leqn[st];
            if (ie < 0) continue;
            int kt = TH_slot_kind[st], lt = TH_slot_node[st];
            if (kt < 2) {
                int a = kt;
                if (H)
                    R[ie] = R[ie] + W * (((sig[a][0] * dx[lt][0] + sig[a][1] * dx[lt][1]) + conv[a] * P[lt]) +
                                         (rho * ud[a]) * P[lt]);
                else
                    R[ie] = R[ie] + W * ((sig[a][0] * dx[lt][0] + sig[a][1] * dx[lt][1]) + conv[a] * P[lt]);
            } else {
                R[ie] = R[ie] + W * (divu * C[TH_slot_pidx[st]]);
            }
            if (!flag) continue;
            if (flag == 2 && !H) continue;
            for (int ss = 0; ss < 22; ss++) {
                long iu = leqn[ss];
                if (iu < 0) continue;
                int ks = TH_slot_kind[ss], ls = TH_slot_node[ss];
                double v;
                if (Y && flag == 4) {
                    if (kt < 2 && ks < 2) {
                        int a = kt, c = ks;
                        v = W * (rho * (P[lt] * dx[ls][a] * yv[c] + P[ls] * dx[lt][c] * yv[a]));
                    } else {
                        continue;
                    }
                } else if (Y) {
                    if (kt < 2 && ks < 2) {
                        int a = kt, c = ks;
                        double cvy = P[ls] * dy[a][c];
                        if (a == c) cvy = cvy + (yv[0] * dx[ls][0] + yv[1] * dx[ls][1]);
                        v = W * ((rho * cvy) * P[lt]);
                    } else {
                        continue;
                    }
                } else if (flag == 2) {
                    if (kt < 2 && kt == ks) v = W * ((rho * P[ls]) * P[lt]);
                    else continue;
                } else if (kt < 2 && ks < 2) {
                    int a = kt, c = ks;
                    double visc = dx[ls][a] * dx[lt][c];
                    double cv = P[ls] * du[a][c];
                    if (a == c) {
                        visc = (dx[ls][0] * dx[lt][0] + dx[ls][1] * dx[lt][1]) + visc;
                        cv = cv + (u[0] * dx[ls][0] + u[1] * dx[ls][1]);
                    }
                    if (H && a == c)
                        v = W * ((mu * visc + (rho * cv) * P[lt]) + ((rho * tw0) * P[ls]) * P[lt]);
                    else
                        v = W * (mu * visc + (rho * cv) * P[lt]);
                } else if (kt < 2 && ks == 2) {
                    v = W * (-(C[TH_slot_pidx[ss]] * dx[lt][kt]));
                } else if (kt == 2 && ks < 2) {
                    v = W * (dx[ls][ks] * C[TH_slot_pidx[st]]);
                } else {
                    continue; /* pressure-pressure block: no term */
                }
                J[ie * nvar + iu] = J[ie * nvar + iu] + v;
            }
        }
    }
}

/*
 * Element loop over a batch. Inputs per element e:
 *   X    [nel][nnode][dim]   nodal coordinates
 *   V    [nel][nslot]        nodal values per slot
 *   aux  [nel][nnode]        nodal auxiliary field (form 3 source) or NULL
 *   leqn [nel][nslot]        local equation of the slot, -1 if pinned
 * Outputs (caller-zeroed, as oomph zeroes the element buffers, elements.h:990-999):
 *   R [nel][maxdof], J [nel][maxdof][maxdof] (local numbering, row stride nvar[e]).
 * flag: 0 residual, 1 residual + Jacobian. deriv: 1 evaluates dR/dparam (and dJ/dparam).
 */
int oracle_elements(int form, int nel, int dim, int nnode, int nqp, int nnode1,
                    const double *w, const double *psi, const double *dpsi, const double *psi1,
                    int nslot, const double *X, const double *V, const double *aux,
                    const long *leqn, const long *nvar, const double *par, int npar,
                    int flag, int deriv, int maxdof, const double *Hist, double tw0,
                    const double *Yv, double *R, double *J)
{
    ref_t r = {dim, nnode, nqp, nnode1, w, psi, dpsi, psi1};
    if (nnode > MAXN || dim > MAXD || nslot > 22 + 5) return -1;
    for (int e = 0; e < nel; e++) {
        const double *Xe = X + (size_t)e * nnode * dim;
        const double *Ve = V + (size_t)e * nslot;
        const long *Le = leqn + (size_t)e * nslot;
        double *Re = R + (size_t)e * maxdof;
        double *Je = J ? J + (size_t)e * maxdof * maxdof : NULL;
        if (form == 4) {
            if (deriv) return -2;
            const double *He = Hist ? Hist + (size_t)e * nslot : NULL;
            const double *Ye = Yv ? Yv + (size_t)e * nslot : NULL;
            th_element(&r, Xe, Ve, Le, par, flag, (int)nvar[e], He, tw0, Ye, Re, Je);
        } else {
            const double *Ae = aux ? aux + (size_t)e * nnode : NULL;
            const double *He = Hist ? Hist + (size_t)e * nslot : NULL;
            const double *Ye = Yv ? Yv + (size_t)e * nslot : NULL;
            scalar_element(form, &r, Xe, Ve, Ae, Le, par, npar, flag, deriv, (int)nvar[e], He, tw0, Ye, Re, Je);
        }
    }
    return 0;
}

/*
 * oomph Problem::sparse_assemble_row_or_column_compressed_with_vectors_of_pairs
 * (problem.cc:5459-5659), compressed-row branch: element order; residuals[eqn] +=
 * el_res[i]; for each (i, j) with fabs(value) > 0.0 a linear search of the row's pair
 * vector, then "+=" or push_back; rows are emitted in first-touch (unsorted) order.
 *
 * Inputs: per element nvar[e], geqn[e][maxdof] (global equation of local dof i),
 * R [nel][maxdof], J [nel][maxdof][maxdof] (row stride nvar[e]).
 * Outputs: residual[ndof] (caller-zeroed); row_start[ndof+1]; *col / *val are malloc'ed
 * here (free with oracle_free). Returns nnz, or -1 on allocation failure.
 */
typedef struct { long *col; double *val; long n, cap; } pairvec_t;

long oracle_vectors_of_pairs(long ndof, int nel, int maxdof, const long *nvar, const long *geqn,
                             const double *R, const double *J, double *residual,
                             long *row_start, long **col_out, double **val_out)
{
    pairvec_t *rows = (pairvec_t *)calloc((size_t)ndof, sizeof(pairvec_t));
    if (!rows) return -1;
    for (int e = 0; e < nel; e++) {
        const long nv = nvar[e];
        const long *ge = geqn + (size_t)e * maxdof;
        const double *Re = R + (size_t)e * maxdof;
        const double *Je = J + (size_t)e * maxdof * maxdof;
        for (long i = 0; i < nv; i++) {
            long eq = ge[i];
            residual[eq] += Re[i];
            pairvec_t *pv = &rows[eq];
            for (long j = 0; j < nv; j++) {
                long unk = ge[j];
                double v = Je[i * nv + j];
                if (!(fabs(v) > 0.0)) continue;
                long k;
                for (k = 0; k < pv->n; k++)
                    if (pv->col[k] == unk) break;
                if (k < pv->n) {
                    pv->val[k] += v;
                } else {
                    if (pv->n == pv->cap) {
                        long nc = pv->cap ? 2 * pv->cap : 16;
                        long *c2 = (long *)realloc(pv->col, nc * sizeof(long));
                        double *v2 = (double *)realloc(pv->val, nc * sizeof(double));
                        if (!c2 || !v2) return -1;
                        pv->col = c2; pv->val = v2; pv->cap = nc;
                    }
                    pv->col[pv->n] = unk;
                    pv->val[pv->n] = v;
                    pv->n++;
                }
            }
        }
    }
    row_start[0] = 0;
    for (long i = 0; i < ndof; i++) row_start[i + 1] = row_start[i] + rows[i].n;
    long nnz = row_start[ndof];
    long *col = (long *)malloc((nnz ? nnz : 1) * sizeof(long));
    double *val = (double *)malloc((nnz ? nnz : 1) * sizeof(double));
    if (!col || !val) return -1;
    for (long i = 0; i < ndof; i++) {
        memcpy(col + row_start[i], rows[i].col, rows[i].n * sizeof(long));
        memcpy(val + row_start[i], rows[i].val, rows[i].n * sizeof(double));
        free(rows[i].col);
        free(rows[i].val);
    }
    free(rows);
    *col_out = col;
    *val_out = val;
    return nnz;
}

void oracle_free(void *p) { free(p); }

/* oomph Problem::get_residuals (problem.cc:3810-3900): element order, residuals[eqn] +=
 * el_res[i]. residual[] caller-zeroed. */
void oracle_assemble_residual(int nel, int maxdof, const long *nvar, const long *geqn,
                              const double *R, double *residual)
{
    for (int e = 0; e < nel; e++)
        for (long i = 0; i < nvar[e]; i++)
            residual[geqn[(size_t)e * maxdof + i]] += R[(size_t)e * maxdof + i];
}
