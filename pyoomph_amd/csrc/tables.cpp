// Reference-element tables: integration schemes and Lagrange shape functions.
//
// Knots and weights are the reference's data, value for value (a table that has to be
// alike: the truncated decimals, and the mistyped positive knot 0.774596662941483 of
// Gauss<2,3>, change every element integral in the last digits):
//   TGauss<2,2> oomph integral.cc:344-349, TGauss<2,3> :355-369,
//   Gauss<2,3> :84-102, Gauss<3,3> :169-206.
// The element keeps its default scheme (Telements.h:1515, Qelements.h:845,1257) because
// pyoomph leaves integration_order at 0 (src/codegen.cpp:7059-7061).
// Shape functions: TElement<2,2> Telements.h:523-542, TElement<2,3> :627-662,
// OneDimLagrange shape.h:607-650 with the tensor products of Qelements.cc:320-372 and
// 621-658; the Q1 pressure basis on Q2 quads src/elements.cpp:9288-9305.
#include "common.h"

namespace pyhip {

static const double kTG22K[3][2] = {{0.1666666666667, 0.1666666666667},
                                    {0.6666666666667, 0.1666666666667},
                                    {0.1666666666667, 0.6666666666667}};
static const double kTG22W[3] = {0.1666666666667, 0.1666666666667, 0.1666666666667};
static const double kTG23K[7][2] = {{0.1012865073235, 0.1012865073235},
                                    {0.7974269853531, 0.1012865073235},
                                    {0.1012865073235, 0.7974269853531},
                                    {0.4701420641051, 0.0597158717898},
                                    {0.4701420641051, 0.4701420641051},
                                    {0.0597158717898, 0.4701420641051},
                                    {0.3333333333333, 0.3333333333333}};
static const double kTG23W[7] = {0.5 * 0.1259391805448, 0.5 * 0.1259391805448,
                                 0.5 * 0.1259391805448, 0.5 * 0.1323941527885,
                                 0.5 * 0.1323941527885, 0.5 * 0.1323941527885, 0.5 * 0.225};
static const double kG23Neg = -0.774596669241483, kG23Pos = 0.774596662941483;
static const double kG33 = 0.77459666924148;
static const double kG33W[27] = {
    0.17146776406035, 0.27434842249657, 0.17146776406035, 0.27434842249657, 0.43895747599451,
    0.27434842249657, 0.17146776406035, 0.27434842249657, 0.17146776406035, 0.27434842249657,
    0.43895747599451, 0.27434842249657, 0.43895747599451, 0.70233196159122, 0.43895747599451,
    0.27434842249657, 0.43895747599451, 0.27434842249657, 0.17146776406035, 0.27434842249657,
    0.17146776406035, 0.27434842249657, 0.43895747599451, 0.27434842249657, 0.17146776406035,
    0.27434842249657, 0.17146776406035};

static void lag3(double s, double *p) {
    p[0] = 0.5 * s * (s - 1.0);
    p[1] = 1.0 - s * s;
    p[2] = 0.5 * s * (s + 1.0);
}
static void dlag3(double s, double *d) {
    d[0] = s - 0.5;
    d[1] = -2.0 * s;
    d[2] = s + 0.5;
}

int fill_ref_tables(int kind, RefTables *t) {
    *t = RefTables();
    if (kind == ELEM_TRI_P1) {
        t->dim = 2; t->nnode = 3; t->nqp = 3;
        for (int q = 0; q < 3; q++) {
            const double s0 = kTG22K[q][0], s1 = kTG22K[q][1];
            t->w[q] = kTG22W[q];
            double *P = t->psi + q * 3, *D = t->dpsi + q * 3 * 2;
            P[0] = s0;
            P[1] = s1;
            P[2] = 1.0 - s0 - s1;
            D[0] = 1.0; D[1] = 0.0; D[2] = 0.0; D[3] = 1.0; D[4] = -1.0; D[5] = -1.0;
        }
        return 0;
    }
    if (kind == ELEM_TRI_P2) {
        t->dim = 2; t->nnode = 6; t->nqp = 7;
        for (int q = 0; q < 7; q++) {
            const double s0 = kTG23K[q][0], s1 = kTG23K[q][1];
            const double s2 = 1.0 - s0 - s1;
            t->w[q] = kTG23W[q];
            double *P = t->psi + q * 6, *D = t->dpsi + q * 6 * 2;
            P[0] = 2.0 * s0 * (s0 - 0.5);
            P[1] = 2.0 * s1 * (s1 - 0.5);
            P[2] = 2.0 * s2 * (s2 - 0.5);
            P[3] = 4.0 * s0 * s1;
            P[4] = 4.0 * s1 * s2;
            P[5] = 4.0 * s2 * s0;
            const double c = 2.0 * (2.0 * s0 - 1.5 + 2.0 * s1);
            D[0] = 4.0 * s0 - 1.0; D[1] = 0.0;
            D[2] = 0.0; D[3] = 4.0 * s1 - 1.0;
            D[4] = c; D[5] = c;
            D[6] = 4.0 * s1; D[7] = 4.0 * s0;
            D[8] = -4.0 * s1; D[9] = 4.0 * (1.0 - s0 - 2.0 * s1);
            D[10] = 4.0 * (1.0 - 2.0 * s0 - s1); D[11] = -4.0 * s0;
        }
        return 0;
    }
    if (kind == ELEM_QUAD_Q2) {
        t->dim = 2; t->nnode = 9; t->nqp = 9; t->nnode1 = 4;
        const double k1[3] = {kG23Neg, 0.0, kG23Pos};
        const double w1[3] = {25.0 / 81.0, 40.0 / 81.0, 64.0 / 81.0};
        for (int qa = 0; qa < 3; qa++)
            for (int qb = 0; qb < 3; qb++) {
                const int q = qa * 3 + qb;
                // oomph Gauss<2,3>: first coordinate outer; weights 25/81, 40/81, 64/81
                const double s0 = k1[qa], s1 = k1[qb];
                const int ca = (qa == 1), cb = (qb == 1);
                t->w[q] = (ca && cb) ? w1[2] : ((ca || cb) ? w1[1] : w1[0]);
                double P0[3], P1[3], D0[3], D1[3];
                lag3(s0, P0); lag3(s1, P1); dlag3(s0, D0); dlag3(s1, D1);
                double *P = t->psi + q * 9, *D = t->dpsi + q * 9 * 2;
                int idx = 0;
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) {
                        P[idx] = P1[i] * P0[j];
                        D[idx * 2 + 0] = P1[i] * D0[j];
                        D[idx * 2 + 1] = D1[i] * P0[j];
                        idx++;
                    }
                const double p1[2] = {0.5 * (1.0 - s0), 0.5 * (1.0 + s0)};
                const double p2[2] = {0.5 * (1.0 - s1), 0.5 * (1.0 + s1)};
                for (int i = 0; i < 2; i++)
                    for (int j = 0; j < 2; j++) t->psi1[q * 4 + 2 * i + j] = p2[i] * p1[j];
            }
        return 0;
    }
    if (kind == ELEM_HEX_Q2) {
        t->dim = 3; t->nnode = 27; t->nqp = 27;
        const double k1[3] = {-kG33, 0.0, kG33};
        for (int qa = 0; qa < 3; qa++)
            for (int qb = 0; qb < 3; qb++)
                for (int qc = 0; qc < 3; qc++) {
                    const int q = (qa * 3 + qb) * 3 + qc;
                    const double s0 = k1[qa], s1 = k1[qb], s2 = k1[qc];
                    t->w[q] = kG33W[q];
                    double P0[3], P1[3], P2[3], D0[3], D1[3], D2[3];
                    lag3(s0, P0); lag3(s1, P1); lag3(s2, P2);
                    dlag3(s0, D0); dlag3(s1, D1); dlag3(s2, D2);
                    double *P = t->psi + q * 27, *D = t->dpsi + q * 27 * 3;
                    int idx = 0;
                    for (int i = 0; i < 3; i++)
                        for (int j = 0; j < 3; j++)
                            for (int k = 0; k < 3; k++) {
                                P[idx] = P2[i] * P1[j] * P0[k];
                                D[idx * 3 + 0] = P2[i] * P1[j] * D0[k];
                                D[idx * 3 + 1] = P2[i] * D1[j] * P0[k];
                                D[idx * 3 + 2] = D2[i] * P1[j] * P0[k];
                                idx++;
                            }
                }
        return 0;
    }
    return -1;
}

}  // namespace pyhip
