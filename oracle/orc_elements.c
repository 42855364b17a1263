/*
 * orc_elements.c -- per-element energy, force and Hessian (oracle, FP64).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Written from the closed forms of the energies, not from the generated code:
 *
 *  inertial  (ComputeInertial.cpp:2-136, called at Forces.cpp:390)
 *     A = signed material area, f_v = (rho A / 3) g, M = (rho A / 12)(1+d_vw) I3,
 *     W = -rho A g.(xa+xb+xc)/3.
 *  membrane  (ComputeMembrane.cpp:2-320, called at Forces.cpp:389)
 *     R = Q^T P (2x3), G = R Dx DX^-1, E = G - I, mu = e/(2(1+nu)),
 *     lambda = e nu/((1+nu)(1-2nu)), W = A (mu |E|^2 + lambda/2 tr(E)^2).
 *     With d_v = DX^-T c_v (c_a=(-1,-1), c_b=(1,0), c_c=(0,1)) and
 *     S = dW/dG = A(2 mu E + lambda tr(E) I):
 *        dW/dx_v = R^T S d_v,
 *        K_vw   = A [ 2 mu (d_v.d_w) R^T R + lambda (R^T d_v)(R^T d_w)^T ].
 *  P, Q      (Forces.cpp:357-372, poldec Forces.cpp:33-52)
 *  bending   (ComputeBending.cpp:16-738, called at Forces.cpp:741)
 *     n0 = (x1-x0)x(x2-x0), n1 = (x0-x1)x(x3-x1), phi = n0^.n1^,
 *     w = 3/2 beta |X1-X0|^2 / (A0+A1), W = w (1 - phi),
 *     f = w grad(phi), K = -w Hess(phi) (exact, indefinite in general).
 *
 * Parity pinned by: golden vectors from the reference's compiled element
 * kernels (tests/golden/elements.npz), finite differences, invariants.
 */
#include "oracle.h"
#include <math.h>
#include <string.h>

static double signed_area2(const double *Xa, const double *Xb, const double *Xc)
{
    /* twice the signed material-space area (ComputeInertial.cpp:33) */
    return (Xb[0] - Xa[0]) * (Xc[1] - Xa[1]) - (Xb[1] - Xa[1]) * (Xc[0] - Xa[0]);
}

void orc_inertial(const double *xa, const double *xb, const double *xc,
                  const double *Xa, const double *Xb, const double *Xc,
                  const double *g, double rho, double *W, double *f, double *M)
{
    double A = 0.5 * signed_area2(Xa, Xb, Xc);
    double m = rho * A;
    double s = 0.0;
    for (int k = 0; k < 3; ++k) s += g[k] * (xa[k] + xb[k] + xc[k]);
    W[0] = -m * s / 3.0;
    for (int v = 0; v < 3; ++v)
        for (int k = 0; k < 3; ++k) f[3 * v + k] = m * g[k] / 3.0;
    memset(M, 0, 81 * sizeof(double));
    for (int v = 0; v < 3; ++v)
        for (int w = 0; w < 3; ++w)
            for (int k = 0; k < 3; ++k)
                M[(3 * v + k) * 9 + 3 * w + k] = (v == w) ? m / 6.0 : m / 12.0;
}

/* poldec(M): Forces.cpp:33-52.  Q = (M + s*cof(M)) / |col0|, s = sign(det M). */
void orc_poldec(const double *Mc, double *Q)
{
    double m11 = Mc[0], m21 = Mc[1], m12 = Mc[2], m22 = Mc[3];
    double det = m11 * m22 - m12 * m21;
    int s = 1;
    if (det < 0) s = -1;
    else if (det == 0) s = 0;
    /* MM = [[m22, -m21], [-m12, m11]] (row-major comma initialiser) */
    double q00 = m11 + s * m22;
    double q01 = m12 + s * (-m21);
    double q10 = m21 + s * (-m12);
    double q11 = m22 + s * m11;
    double clen = sqrt(q00 * q00 + q10 * q10);
    Q[0] = q00 / clen; Q[1] = q10 / clen; Q[2] = q01 / clen; Q[3] = q11 / clen;
}

static void cross3(const double *a, const double *b, double *c)
{
    c[0] = a[1] * b[2] - a[2] * b[1];
    c[1] = a[2] * b[0] - a[0] * b[2];
    c[2] = a[0] * b[1] - a[1] * b[0];
}
static double dot3(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }

/* P and Q of one face: Forces.cpp:357-372 */
void orc_face_PQ(const double *xa, const double *xb, const double *xc,
                 const double *Xa, const double *Xb, const double *Xc,
                 double *P, double *Q)
{
    double e1[3], e2[3], nrm[3], px[3], py[3];
    for (int k = 0; k < 3; ++k) { e1[k] = xb[k] - xa[k]; e2[k] = xc[k] - xa[k]; }
    cross3(e1, e2, nrm);
    double l1 = sqrt(dot3(e1, e1));
    for (int k = 0; k < 3; ++k) px[k] = e1[k] / l1;
    cross3(nrm, px, py);
    double l2 = sqrt(dot3(py, py));
    for (int k = 0; k < 3; ++k) py[k] /= l2;
    /* P (2x3) rows px, py; stored column-major */
    for (int k = 0; k < 3; ++k) { P[2 * k] = px[k]; P[2 * k + 1] = py[k]; }
    /* DX = [Xb-Xa, Xc-Xa]; F = Dx DX^-1 (3x2); Fbar = P F (2x2) */
    double a = Xb[0] - Xa[0], b = Xc[0] - Xa[0], c = Xb[1] - Xa[1], d = Xc[1] - Xa[1];
    double det = a * d - b * c;
    double Di[4] = { d / det, -c / det, -b / det, a / det }; /* col-major 2x2 inverse */
    double Fm[6];                                            /* col-major 3x2 */
    for (int k = 0; k < 3; ++k) {
        Fm[k] = e1[k] * Di[0] + e2[k] * Di[1];
        Fm[3 + k] = e1[k] * Di[2] + e2[k] * Di[3];
    }
    double Fb[4];
    for (int r = 0; r < 2; ++r)
        for (int cc = 0; cc < 2; ++cc) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += P[2 * k + r] * Fm[3 * cc + k];
            Fb[2 * cc + r] = s;
        }
    orc_poldec(Fb, Q);
}

void orc_membrane(const double *xa, const double *xb, const double *xc,
                  const double *Xa, const double *Xb, const double *Xc,
                  double e, double nu, const double *P, const double *Q,
                  double *W, double *f, double *K)
{
    double A2 = signed_area2(Xa, Xb, Xc);
    double A = 0.5 * A2;
    double mu = e / (2.0 * (1.0 + nu));
    double lam = e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));
    /* R = Q^T P, 2x3; R[a][k] */
    double R[2][3];
    for (int a = 0; a < 2; ++a)
        for (int k = 0; k < 3; ++k)
            R[a][k] = Q[2 * a + 0] * P[2 * k + 0] + Q[2 * a + 1] * P[2 * k + 1];
    /* DX^-1 (row-major Di[m][b]); d_v[b] = sum_m c_v[m] Di[m][b] */
    double a11 = Xb[0] - Xa[0], a12 = Xc[0] - Xa[0], a21 = Xb[1] - Xa[1], a22 = Xc[1] - Xa[1];
    double Di[2][2] = { { a22 / A2, -a12 / A2 }, { -a21 / A2, a11 / A2 } };
    double d[3][2];
    for (int b = 0; b < 2; ++b) {
        d[1][b] = Di[0][b];
        d[2][b] = Di[1][b];
        d[0][b] = -Di[0][b] - Di[1][b];
    }
    const double *xs[3] = { xa, xb, xc };
    /* G = sum_v (R x_v) d_v^T */
    double G[2][2] = { { 0, 0 }, { 0, 0 } };
    double Rx[3][2];
    for (int v = 0; v < 3; ++v)
        for (int a = 0; a < 2; ++a) Rx[v][a] = dot3(R[a], xs[v]);
    for (int v = 0; v < 3; ++v)
        for (int a = 0; a < 2; ++a)
            for (int b = 0; b < 2; ++b) G[a][b] += Rx[v][a] * d[v][b];
    double E[2][2] = { { G[0][0] - 1.0, G[0][1] }, { G[1][0], G[1][1] - 1.0 } };
    double tr = E[0][0] + E[1][1];
    double fro = E[0][0] * E[0][0] + E[0][1] * E[0][1] + E[1][0] * E[1][0] + E[1][1] * E[1][1];
    W[0] = A * (mu * fro + 0.5 * lam * tr * tr);
    double S[2][2];
    for (int a = 0; a < 2; ++a)
        for (int b = 0; b < 2; ++b) S[a][b] = A * (2.0 * mu * E[a][b] + (a == b ? lam * tr : 0.0));
    /* f_v = -R^T S d_v */
    for (int v = 0; v < 3; ++v) {
        double Sd[2] = { S[0][0] * d[v][0] + S[0][1] * d[v][1], S[1][0] * d[v][0] + S[1][1] * d[v][1] };
        for (int k = 0; k < 3; ++k) f[3 * v + k] = -(R[0][k] * Sd[0] + R[1][k] * Sd[1]);
    }
    /* K_vw = A [2 mu (d_v.d_w) R^T R + lambda (R^T d_v)(R^T d_w)^T] */
    double RtR[3][3], Rd[3][3];
    for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) RtR[k][l] = R[0][k] * R[0][l] + R[1][k] * R[1][l];
    for (int v = 0; v < 3; ++v)
        for (int k = 0; k < 3; ++k) Rd[v][k] = R[0][k] * d[v][0] + R[1][k] * d[v][1];
    for (int v = 0; v < 3; ++v)
        for (int w = 0; w < 3; ++w) {
            double dd = d[v][0] * d[w][0] + d[v][1] * d[w][1];
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l)
                    K[(3 * v + k) * 9 + 3 * w + l] = A * (2.0 * mu * dd * RtR[k][l] + lam * Rd[v][k] * Rd[w][l]);
        }
}

/* skew matrix [a]x : [a]x b = a x b, row-major 3x3 */
static void skew(const double *a, double S[3][3])
{
    S[0][0] = 0;     S[0][1] = -a[2]; S[0][2] = a[1];
    S[1][0] = a[2];  S[1][1] = 0;     S[1][2] = -a[0];
    S[2][0] = -a[1]; S[2][1] = a[0];  S[2][2] = 0;
}

void orc_bending(const double *x0, const double *x1, const double *x2, const double *x3,
                 const double *X0, const double *X1, const double *X2, const double *X3,
                 double beta, double *W, double *f, double *K)
{
    /* stiffness weight: W00 of ComputeBending.cpp:98 is 3/4 beta |E|^2/(A0+A1) (2 - 2 phi) */
    double ex = X1[0] - X0[0], ey = X1[1] - X0[1];
    double A0 = 0.5 * ((X1[0] - X0[0]) * (X2[1] - X0[1]) - (X1[1] - X0[1]) * (X2[0] - X0[0]));
    double A1 = 0.5 * ((X0[0] - X1[0]) * (X3[1] - X1[1]) - (X0[1] - X1[1]) * (X3[0] - X1[0]));
    double w = 1.5 * beta * (ex * ex + ey * ey) / (A0 + A1);

    double u[3], v[3], s[3], t[3], n0[3], n1[3];
    for (int k = 0; k < 3; ++k) {
        u[k] = x1[k] - x0[k]; v[k] = x2[k] - x0[k];
        s[k] = x0[k] - x1[k]; t[k] = x3[k] - x1[k];
    }
    cross3(u, v, n0);
    cross3(s, t, n1);
    double r0 = sqrt(dot3(n0, n0)), r1 = sqrt(dot3(n1, n1));
    double h0[3], h1[3];
    for (int k = 0; k < 3; ++k) { h0[k] = n0[k] / r0; h1[k] = n1[k] / r1; }
    double phi = dot3(h0, h1);
    W[0] = w * (1.0 - phi);
    double g0[3], g1[3];
    for (int k = 0; k < 3; ++k) {
        g0[k] = (h1[k] - phi * h0[k]) / r0;
        g1[k] = (h0[k] - phi * h1[k]) / r1;
    }
    /* coefficient of each vertex in u, v (n0 = u x v) and s, t (n1 = s x t) */
    const double cu[4] = { -1, 1, 0, 0 }, cv[4] = { -1, 0, 1, 0 };
    const double cs[4] = { 1, -1, 0, 0 }, ct[4] = { 0, -1, 0, 1 };
    /* J0_j = dn0/dx_j = -cu_j [v]x + cv_j [u]x ; J1_j = -cs_j [t]x + ct_j [s]x */
    double Su[3][3], Sv[3][3], Ss[3][3], St[3][3];
    skew(u, Su); skew(v, Sv); skew(s, Ss); skew(t, St);
    double J[2][4][3][3];
    for (int j = 0; j < 4; ++j)
        for (int a = 0; a < 3; ++a)
            for (int b = 0; b < 3; ++b) {
                J[0][j][a][b] = -cu[j] * Sv[a][b] + cv[j] * Su[a][b];
                J[1][j][a][b] = -cs[j] * St[a][b] + ct[j] * Ss[a][b];
            }
    /* gradient: dphi/dx_j = J0_j^T g0 + J1_j^T g1; f = w grad(phi) */
    for (int j = 0; j < 4; ++j)
        for (int b = 0; b < 3; ++b) {
            double acc = 0.0;
            for (int a = 0; a < 3; ++a) acc += J[0][j][a][b] * g0[a] + J[1][j][a][b] * g1[a];
            f[3 * j + b] = w * acc;
        }
    /* Hessian blocks of phi w.r.t. (n0, n1) */
    double Hn[2][2][3][3];
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) {
            double I = (a == b) ? 1.0 : 0.0;
            Hn[0][0][a][b] = -(g0[a] * h0[b] + h0[a] * g0[b]) / r0 - phi * (I - h0[a] * h0[b]) / (r0 * r0);
            Hn[1][1][a][b] = -(g1[a] * h1[b] + h1[a] * g1[b]) / r1 - phi * (I - h1[a] * h1[b]) / (r1 * r1);
            Hn[0][1][a][b] = (I - h0[a] * h0[b] - h1[a] * h1[b] + phi * h0[a] * h1[b]) / (r0 * r1);
        }
    for (int a = 0; a < 3; ++a)
        for (int b = 0; b < 3; ++b) Hn[1][0][a][b] = Hn[0][1][b][a];
    /* C_{p,j} = sum_q Hn[p][q] J_q,j */
    double C[2][4][3][3];
    for (int p = 0; p < 2; ++p)
        for (int j = 0; j < 4; ++j)
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    double acc = 0.0;
                    for (int q = 0; q < 2; ++q)
                        for (int c = 0; c < 3; ++c) acc += Hn[p][q][a][c] * J[q][j][c][b];
                    C[p][j][a][b] = acc;
                }
    double G0[3][3], G1[3][3];
    skew(g0, G0); skew(g1, G1);
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double c0 = cv[i] * cu[j] - cu[i] * cv[j];
            double c1 = ct[i] * cs[j] - cs[i] * ct[j];
            for (int a = 0; a < 3; ++a)
                for (int b = 0; b < 3; ++b) {
                    double acc = 0.0;
                    for (int p = 0; p < 2; ++p)
                        for (int c = 0; c < 3; ++c) acc += J[p][i][c][a] * C[p][j][c][b];
                    acc += c0 * G0[a][b] + c1 * G1[a][b];
                    K[(3 * i + a) * 12 + 3 * j + b] = -w * acc;
                }
        }
}
