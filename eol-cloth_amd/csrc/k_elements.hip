// Element evaluation kernels (gfx950, FP64): one thread per triangle / hinge.
//
// face:  faceBasedF (Forces.cpp:331-520) -- P and Q (Forces.cpp:357-372, poldec
//        Forces.cpp:33-52), ComputeMembrane (ComputeMembrane.cpp:2-320) and
//        ComputeInertial (ComputeInertial.cpp:2-136), from their closed forms:
//          A  = signed material area, R = Q^T P (2x3), d_v = DX^-T c_v,
//          G  = sum_v (R x_v) d_v^T, E = G - I, S = A (2 mu E + lambda tr(E) I),
//          f_v = -R^T S d_v + (rho A / 3) g,
//          K_vw = A [2 mu (d_v.d_w) R^T R + lambda (R^T d_v)(R^T d_w)^T].
//        Writes dampingB h^2 K_vw of the six local blocks (aa, bb, cc, ab, ac, bc --
//        the order of fillxMI/fillxxMI, Forces.cpp:103-125), their mass scalars
//        (M_vw = m I3; the assembly adds them back, giving the reference's
//        M_vw + dampingB h^2 K_vw triplet values) and f_v.
// hinge: edgeBasedF (Forces.cpp:685-910) -- ComputeBending
//        (ComputeBending.cpp:16-738): W = w (1 - n0^.n1^), with
//        n0 = (x1-x0)x(x2-x0), n1 = (x0-x1)x(x3-x1), w = 3/2 beta |X1-X0|^2/(A0+A1).
//        dn_q/dx_j = [a_qj]x, so grad_j phi = sum_q g_q x a_qj and
//        Hess_ij phi = sum_p C_pj x a_pi (column-wise) + (c0 [g0]x + c1 [g1]x),
//        C_pj = sum_q Hn_pq [a_qj]x.  K = -w Hess phi, written as dampingB h^2 K for
//        the 10 blocks (aa,bb,cc,dd,ab,ac,ad,bc,bd,cd -- fillxB/fillxxB order,
//        Forces.cpp:522-539).  Optional SPD projection: the Hessian is restricted
//        to the orthonormal complement of the translations (which it annihilates),
//        diagonalised by cyclic Jacobi and its negative eigenvalues set to zero.
//
// Staging layout (slot-major, 9 doubles per 3x3 block, row-major in the block):
//   stageK[(6 f + p) * 9 + c], stageK[(6 F + 10 h + p) * 9 + c]
//   stageF[(3 f + v) * 3 + k], stageF[(3 F + 4 h + v) * 3 + k]
//   stageM[6 f + p]   (mass scalar: M block = m I3)
#include "eolc_internal.h"

namespace {

struct V3 { double x, y, z; };
__device__ __forceinline__ V3 mk(double a, double b, double c) { return V3{a, b, c}; }
__device__ __forceinline__ V3 ld3(const double *p) { return V3{p[0], p[1], p[2]}; }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return V3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return V3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ __forceinline__ V3 operator*(double s, V3 a) { return V3{s * a.x, s * a.y, s * a.z}; }
__device__ __forceinline__ double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b)
{
    return V3{a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x};
}
__device__ __forceinline__ double comp(V3 a, int k) { return k == 0 ? a.x : (k == 1 ? a.y : a.z); }

// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256)
k_face_eval(int F, const int *__restrict__ tri, const double *__restrict__ x, const double *__restrict__ X,
            double e_all, const double *__restrict__ e_face, double nu, double rho, double gx, double gy, double gz,
            double s, double *__restrict__ stageK, double *__restrict__ stageF, double *__restrict__ stageM,
            int *__restrict__ bad)
{
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const double e = e_face ? e_face[f] : e_all;   // per-face Young's modulus (batched cloths)
    const int ia = tri[3 * f], ib = tri[3 * f + 1], ic = tri[3 * f + 2];
    const V3 xa = ld3(x + 3 * ia), xb = ld3(x + 3 * ib), xc = ld3(x + 3 * ic);
    const double Xax = X[2 * ia], Xay = X[2 * ia + 1];
    const double a11 = X[2 * ib] - Xax, a21 = X[2 * ib + 1] - Xay;
    const double a12 = X[2 * ic] - Xax, a22 = X[2 * ic + 1] - Xay;
    const double A2 = a11 * a22 - a21 * a12;   // twice the signed material area
    if (!(A2 > 0.0)) atomicOr(bad, 1);
    const double A = 0.5 * A2;
    const double mu = e / (2.0 * (1.0 + nu));
    const double lam = e * nu / ((1.0 + nu) * (1.0 - 2.0 * nu));

    // P: rows px = e1/|e1|, py = normalize(n x px), n = e1 x e2
    const V3 e1 = xb - xa, e2 = xc - xa;
    const V3 nrm = cross(e1, e2);
    const V3 px = (1.0 / sqrt(dot(e1, e1))) * e1;
    V3 py = cross(nrm, px);
    py = (1.0 / sqrt(dot(py, py))) * py;
    // DX^-1 (row-major Di[m][b])
    const double i00 = a22 / A2, i01 = -a12 / A2, i10 = -a21 / A2, i11 = a11 / A2;
    // F = Dx DX^-1 (3x2), Fbar = P F (2x2)
    const V3 F0 = i00 * e1 + i10 * e2;   // column 0
    const V3 F1 = i01 * e1 + i11 * e2;   // column 1
    const double m11 = dot(px, F0), m12 = dot(px, F1), m21 = dot(py, F0), m22 = dot(py, F1);
    // poldec
    const double det = m11 * m22 - m12 * m21;
    const double sg = det < 0.0 ? -1.0 : (det == 0.0 ? 0.0 : 1.0);
    double q00 = m11 + sg * m22, q01 = m12 - sg * m21, q10 = m21 - sg * m12, q11 = m22 + sg * m11;
    const double cl = sqrt(q00 * q00 + q10 * q10);
    q00 /= cl; q01 /= cl; q10 /= cl; q11 /= cl;
    // R = Q^T P: row a = Q[0][a] px + Q[1][a] py
    const V3 R0 = q00 * px + q10 * py;
    const V3 R1 = q01 * px + q11 * py;
    // d_v (rows of DX^-T c_v): d_b = (i00, i01), d_c = (i10, i11), d_a = -d_b - d_c
    const double db0 = i00, db1 = i01, dc0 = i10, dc1 = i11;
    const double da0 = -db0 - dc0, da1 = -db1 - dc1;
    const double d[3][2] = {{da0, da1}, {db0, db1}, {dc0, dc1}};
    // G = sum_v (R x_v) d_v^T
    const V3 xs[3] = {xa, xb, xc};
    double G00 = 0, G01 = 0, G10 = 0, G11 = 0;
#pragma unroll
    for (int vv = 0; vv < 3; ++vv) {
        double r0 = dot(R0, xs[vv]), r1 = dot(R1, xs[vv]);
        G00 += r0 * d[vv][0]; G01 += r0 * d[vv][1];
        G10 += r1 * d[vv][0]; G11 += r1 * d[vv][1];
    }
    const double E00 = G00 - 1.0, E11 = G11 - 1.0;
    const double tr = E00 + E11;
    const double S00 = A * (2.0 * mu * E00 + lam * tr), S11 = A * (2.0 * mu * E11 + lam * tr);
    const double S01 = A * (2.0 * mu * G01), S10 = A * (2.0 * mu * G10);
    const double m = rho * A;
    const double mg = m / 3.0;
    double *sf = stageF + (size_t)9 * f;
#pragma unroll
    for (int vv = 0; vv < 3; ++vv) {
        double sd0 = S00 * d[vv][0] + S01 * d[vv][1];
        double sd1 = S10 * d[vv][0] + S11 * d[vv][1];
        V3 fm = (-sd0) * R0 + (-sd1) * R1;
        sf[3 * vv + 0] = fm.x + mg * gx;
        sf[3 * vv + 1] = fm.y + mg * gy;
        sf[3 * vv + 2] = fm.z + mg * gz;
    }
    // K blocks
    double RtR[3][3];
#pragma unroll
    for (int k = 0; k < 3; ++k)
#pragma unroll
        for (int l = 0; l < 3; ++l) RtR[k][l] = comp(R0, k) * comp(R0, l) + comp(R1, k) * comp(R1, l);
    V3 Rd[3];
#pragma unroll
    for (int vv = 0; vv < 3; ++vv) Rd[vv] = d[vv][0] * R0 + d[vv][1] * R1;
    const int pv[6] = {0, 1, 2, 0, 0, 1}, pw[6] = {0, 1, 2, 1, 2, 2};
    double *sk = stageK + (size_t)54 * f;
    double *sm = stageM + (size_t)6 * f;
#pragma unroll
    for (int p = 0; p < 6; ++p) {
        const int vv = pv[p], ww = pw[p];
        const double dd = d[vv][0] * d[ww][0] + d[vv][1] * d[ww][1];
        const double mass = (vv == ww) ? m / 6.0 : m / 12.0;
        sm[p] = mass;
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                // (Rd_v,j * Rd_w,k) is formed first so that diagonal blocks come out exactly symmetric
                double kk = A * (2.0 * mu * dd * RtR[j][k] + lam * (comp(Rd[vv], j) * comp(Rd[ww], k)));
                sk[9 * p + 3 * j + k] = s * kk;
            }
    }
}

// ---------------------------------------------------------------------------
// Shared pieces of the hinge evaluation.
struct HingeGeom {
    V3 a[2][4];     // dn_q/dx_j = [a_qj]x
    V3 g[2];        // dphi/dn_q
    double Hn[2][2][3][3];
    double w;
    double phi;
};

__device__ __forceinline__ void hinge_geom(const double *__restrict__ x, const double *__restrict__ X,
                                           const int *q, double beta, HingeGeom &G)
{
    const V3 x0 = ld3(x + 3 * q[0]), x1 = ld3(x + 3 * q[1]), x2 = ld3(x + 3 * q[2]), x3 = ld3(x + 3 * q[3]);
    const double X0x = X[2 * q[0]], X0y = X[2 * q[0] + 1], X1x = X[2 * q[1]], X1y = X[2 * q[1] + 1];
    const double X2x = X[2 * q[2]], X2y = X[2 * q[2] + 1], X3x = X[2 * q[3]], X3y = X[2 * q[3] + 1];
    const double ex = X1x - X0x, ey = X1y - X0y;
    const double A0 = 0.5 * ((X1x - X0x) * (X2y - X0y) - (X1y - X0y) * (X2x - X0x));
    const double A1 = 0.5 * ((X0x - X1x) * (X3y - X1y) - (X0y - X1y) * (X3x - X1x));
    G.w = 1.5 * beta * (ex * ex + ey * ey) / (A0 + A1);
    const V3 u = x1 - x0, v = x2 - x0, s = x0 - x1, t = x3 - x1;
    const V3 n0 = cross(u, v), n1 = cross(s, t);
    const double r0 = sqrt(dot(n0, n0)), r1 = sqrt(dot(n1, n1));
    const V3 h0 = (1.0 / r0) * n0, h1 = (1.0 / r1) * n1;
    const double phi = dot(h0, h1);
    G.phi = phi;
    G.g[0] = (1.0 / r0) * (h1 - phi * h0);
    G.g[1] = (1.0 / r1) * (h0 - phi * h1);
    // a_0j = -cu_j v + cv_j u ; a_1j = -cs_j t + ct_j s
    G.a[0][0] = v - u; G.a[0][1] = (-1.0) * v; G.a[0][2] = u; G.a[0][3] = mk(0, 0, 0);
    G.a[1][0] = (-1.0) * t; G.a[1][1] = t - s; G.a[1][2] = mk(0, 0, 0); G.a[1][3] = s;
    const double ir00 = 1.0 / (r0 * r0), ir11 = 1.0 / (r1 * r1), ir01 = 1.0 / (r0 * r1);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) {
            const double I = (i == j) ? 1.0 : 0.0;
            const double g0i = comp(G.g[0], i), g0j = comp(G.g[0], j), g1i = comp(G.g[1], i), g1j = comp(G.g[1], j);
            const double h0i = comp(h0, i), h0j = comp(h0, j), h1i = comp(h1, i), h1j = comp(h1, j);
            G.Hn[0][0][i][j] = -(g0i * h0j + h0i * g0j) / r0 - phi * (I - h0i * h0j) * ir00;
            G.Hn[1][1][i][j] = -(g1i * h1j + h1i * g1j) / r1 - phi * (I - h1i * h1j) * ir11;
            G.Hn[0][1][i][j] = (I - h0i * h0j - h1i * h1j + phi * h0i * h1j) * ir01;
        }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) G.Hn[1][0][i][j] = G.Hn[0][1][j][i];
}

// Hessian block (i, j) of phi (3x3, row-major in out)
__device__ __forceinline__ void hinge_hess_block(const HingeGeom &G, int i, int j, double out[9])
{
    const double cu[4] = {-1, 1, 0, 0}, cv[4] = {-1, 0, 1, 0}, cs[4] = {1, -1, 0, 0}, ct[4] = {0, -1, 0, 1};
    const double c0 = cv[i] * cu[j] - cu[i] * cv[j];
    const double c1 = ct[i] * cs[j] - cs[i] * ct[j];
#pragma unroll
    for (int b = 0; b < 3; ++b) {
        const V3 eb = mk(b == 0, b == 1, b == 2);
        V3 col = mk(0, 0, 0);
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            // C_pj e_b = sum_q Hn_pq (a_qj x e_b)
            V3 cpj = mk(0, 0, 0);
#pragma unroll
            for (int qq = 0; qq < 2; ++qq) {
                const V3 ax = cross(G.a[qq][j], eb);
                cpj.x += G.Hn[p][qq][0][0] * ax.x + G.Hn[p][qq][0][1] * ax.y + G.Hn[p][qq][0][2] * ax.z;
                cpj.y += G.Hn[p][qq][1][0] * ax.x + G.Hn[p][qq][1][1] * ax.y + G.Hn[p][qq][1][2] * ax.z;
                cpj.z += G.Hn[p][qq][2][0] * ax.x + G.Hn[p][qq][2][1] * ax.y + G.Hn[p][qq][2][2] * ax.z;
            }
            // [a_pi]x^T C = -(a_pi x C) = C x a_pi
            col = col + cross(cpj, G.a[p][i]);
        }
        // c0 [g0]x e_b + c1 [g1]x e_b
        col = col + c0 * cross(G.g[0], eb) + c1 * cross(G.g[1], eb);
        out[0 * 3 + b] = col.x;
        out[1 * 3 + b] = col.y;
        out[2 * 3 + b] = col.z;
    }
}

__global__ void __launch_bounds__(256)
k_hinge_eval(int H, int F, const int *__restrict__ hinge, const double *__restrict__ x,
             const double *__restrict__ X, double beta_all, const double *__restrict__ beta_hinge, double s,
             double *__restrict__ stageK, double *__restrict__ stageF)
{
    int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= H) return;
    int q[4] = {hinge[4 * h], hinge[4 * h + 1], hinge[4 * h + 2], hinge[4 * h + 3]};
    HingeGeom G;
    hinge_geom(x, X, q, beta_hinge ? beta_hinge[h] : beta_all, G);
    // force f_j = w sum_q g_q x a_qj
    double *sf = stageF + (size_t)9 * F + (size_t)12 * h;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        V3 gj = cross(G.g[0], G.a[0][j]) + cross(G.g[1], G.a[1][j]);
        sf[3 * j] = G.w * gj.x; sf[3 * j + 1] = G.w * gj.y; sf[3 * j + 2] = G.w * gj.z;
    }
    const int pv[10] = {0, 1, 2, 3, 0, 0, 0, 1, 1, 2}, pw[10] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};
    const double scale = -G.w * s;
    double *sk = stageK + (size_t)54 * F + (size_t)90 * h;
    for (int p = 0; p < 10; ++p) {
        double blk[9];
        hinge_hess_block(G, pv[p], pw[p], blk);
        if (pv[p] == pw[p]) {  // keep diagonal blocks exactly symmetric (upper triangle mirrored)
            blk[3] = blk[1]; blk[6] = blk[2]; blk[7] = blk[5];
        }
#pragma unroll
        for (int c = 0; c < 9; ++c) sk[9 * p + c] = scale * blk[c];
    }
}

// ---------------------------------------------------------------------------
// SPD-projected variant.  Q = orthonormal basis (12x9) of the complement of the
// translations: contrasts c1 = (1,-1,0,0)/sqrt2, c2 = (1,1,-2,0)/sqrt6,
// c3 = (1,1,1,-3)/sqrt12 (each times I3).  K9 = Q^T K Q is diagonalised with
// cyclic Jacobi (rotation order (0,1), (0,2), ..., (7,8); at most 30 sweeps, until
// the off-diagonal mass falls below 1e-30 of the Frobenius mass), negative
// eigenvalues are zeroed and K+ = Q V L+ V^T Q^T.
// K9 is kept as its 45 upper entries and V as 81 values in registers; every index is
// a compile-time constant (all loops over matrix entries are unrolled), so nothing
// goes through LDS (one wave per SIMD, no LDS occupancy limit, no scratch).
namespace proj {
constexpr double r2 = 0.70710678118654752440, r6 = 0.40824829046386301637, r12 = 0.28867513459481288225;
constexpr double cw[3][4] = {{r2, -r2, 0, 0}, {r6, r6, -2 * r6, 0}, {r12, r12, r12, -3 * r12}};
constexpr int ix(int P, int Q) { return P <= Q ? P * 9 - P * (P - 1) / 2 + (Q - P) : Q * 9 - Q * (Q - 1) / 2 + (P - Q); }
}  // namespace proj

__global__ void __launch_bounds__(64)
k_hinge_eval_projr(int H, int F, const int *__restrict__ hinge, const double *__restrict__ x,
                   const double *__restrict__ X, double beta_all, const double *__restrict__ beta_hinge, double s,
                   double *__restrict__ stageK, double *__restrict__ stageF)
{
    using proj::cw;
    using proj::ix;
    const int h = blockIdx.x * blockDim.x + threadIdx.x;
    if (h >= H) return;
    int q[4] = {hinge[4 * h], hinge[4 * h + 1], hinge[4 * h + 2], hinge[4 * h + 3]};
    HingeGeom G;
    hinge_geom(x, X, q, beta_hinge ? beta_hinge[h] : beta_all, G);
    {
        double *sf = stageF + (size_t)9 * F + (size_t)12 * h;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            V3 gj = cross(G.g[0], G.a[0][j]) + cross(G.g[1], G.a[1][j]);
            sf[3 * j] = G.w * gj.x; sf[3 * j + 1] = G.w * gj.y; sf[3 * j + 2] = G.w * gj.z;
        }
    }
    double A[45];
#pragma unroll
    for (int e = 0; e < 45; ++e) A[e] = 0.0;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = i; j < 4; ++j) {
            double blk[9];
            hinge_hess_block(G, i, j, blk);
            if (i == j) { blk[3] = blk[1]; blk[6] = blk[2]; blk[7] = blk[5]; }
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int dd = 0; dd < 3; ++dd) {
                    const double wgt = cw[c][i] * cw[dd][j];
                    const double wgt2 = (i != j) ? cw[c][j] * cw[dd][i] : 0.0;
                    if (wgt == 0.0 && wgt2 == 0.0) continue;
#pragma unroll
                    for (int a = 0; a < 3; ++a)
#pragma unroll
                        for (int b = 0; b < 3; ++b) {
                            const int P = 3 * c + a, Qc = 3 * dd + b;
                            if (P > Qc) continue;   // upper triangle (the matrix is symmetric)
                            const double val = wgt * blk[3 * a + b] + wgt2 * blk[3 * b + a];
                            A[ix(P, Qc)] += -G.w * val;
                        }
                }
        }
    double V[81];
#pragma unroll
    for (int e = 0; e < 81; ++e) V[e] = (e % 10 == 0) ? 1.0 : 0.0;
    double fro = 0.0;
#pragma unroll
    for (int P = 0; P < 9; ++P)
#pragma unroll
        for (int Qc = 0; Qc < 9; ++Qc) fro += A[ix(P, Qc)] * A[ix(P, Qc)];
    for (int sweep = 0; sweep < 30; ++sweep) {
        double off = 0.0;
#pragma unroll
        for (int p = 0; p < 9; ++p)
#pragma unroll
            for (int r = p + 1; r < 9; ++r) off += A[ix(p, r)] * A[ix(p, r)];
        if (!(off > 1e-30 * fro)) break;
#pragma unroll
        for (int p = 0; p < 8; ++p)
#pragma unroll
            for (int r = p + 1; r < 9; ++r) {
                const double apr = A[ix(p, r)];
                if (apr != 0.0) {
                    const double app = A[ix(p, p)], arr = A[ix(r, r)];
                    const double theta = (arr - app) / (2.0 * apr);
                    const double tt = (theta >= 0.0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
                    const double c = 1.0 / sqrt(tt * tt + 1.0), sn = tt * c;
                    // columns p, r then rows p, r of the symmetric matrix
#pragma unroll
                    for (int k = 0; k < 9; ++k) {
                        if (k == p || k == r) continue;
                        const double akp = A[ix(k, p)], akr = A[ix(k, r)];
                        A[ix(k, p)] = c * akp - sn * akr;
                        A[ix(k, r)] = sn * akp + c * akr;
                    }
                    // the 2x2 block {p, r}: column step, then row step
                    const double cpp = c * app - sn * apr, cpr = sn * app + c * apr;
                    const double crp = c * apr - sn * arr, crr = sn * apr + c * arr;
                    A[ix(p, p)] = c * cpp - sn * crp;
                    A[ix(p, r)] = c * cpr - sn * crr;
                    A[ix(r, r)] = sn * cpr + c * crr;
#pragma unroll
                    for (int k = 0; k < 9; ++k) {
                        const double vkp = V[k * 9 + p], vkr = V[k * 9 + r];
                        V[k * 9 + p] = c * vkp - sn * vkr;
                        V[k * 9 + r] = sn * vkp + c * vkr;
                    }
                }
            }
    }
    double lam[9];
#pragma unroll
    for (int k = 0; k < 9; ++k) lam[k] = fmax(A[ix(k, k)], 0.0);
    // K9+ = V diag(lam) V^T (upper triangle, into A)
#pragma unroll
    for (int a = 0; a < 9; ++a)
#pragma unroll
        for (int b = a; b < 9; ++b) {
            double acc = 0.0;
#pragma unroll
            for (int k = 0; k < 9; ++k) acc += V[a * 9 + k] * lam[k] * V[b * 9 + k];
            A[ix(a, b)] = acc;
        }
    const int pv[10] = {0, 1, 2, 3, 0, 0, 0, 1, 1, 2}, pw[10] = {0, 1, 2, 3, 1, 2, 3, 2, 3, 3};
    double *sk = stageK + (size_t)54 * F + (size_t)90 * h;
#pragma unroll
    for (int p = 0; p < 10; ++p) {
        const int i = pv[p], j = pw[p];
#pragma unroll
        for (int a = 0; a < 3; ++a)
#pragma unroll
            for (int b = 0; b < 3; ++b) {
                double acc = 0.0;
#pragma unroll
                for (int c = 0; c < 3; ++c)
#pragma unroll
                    for (int dd = 0; dd < 3; ++dd) acc += cw[c][i] * cw[dd][j] * A[ix(3 * c + a, 3 * dd + b)];
                sk[9 * p + 3 * a + b] = s * acc;
            }
    }
}

// x += h v
__global__ void k_update_x(int n, double h, const double *__restrict__ v, double *__restrict__ x)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = x[i] + h * v[i];
}

inline int blocks_for(int64_t n, int t = 256) { return (int)((n + t - 1) / t); }

}  // namespace

namespace eolc {

hipError_t launch_face_eval(eolc_ctx *c, double h, const double *grav)
{
    const int F = c->topo.F;
    if (F == 0) return hipSuccess;
    const double s = c->dampB * h * h;
    hipLaunchKernelGGL(k_face_eval, dim3(blocks_for(F)), dim3(256), 0, c->stream, F, c->tri, c->x, c->X, c->E,
                       (const double *)c->e_face, c->nu, c->rho, grav[0], grav[1], grav[2], s, c->stageK, c->stageF,
                       c->stageM, (int *)(c->counter + 1));
    return hipGetLastError();
}

hipError_t launch_hinge_eval(eolc_ctx *c, double h)
{
    const int H = c->topo.H;
    if (H == 0) return hipSuccess;
    const double s = c->dampB * h * h;
    if (c->flags & EOLC_OPT_SPD_PROJECT_BENDING) {
        hipLaunchKernelGGL(k_hinge_eval_projr, dim3(blocks_for(H, 64)), dim3(64), 0, c->stream, H, c->topo.F,
                           c->hinge, c->x, c->X, c->beta, (const double *)c->beta_hinge, s, c->stageK, c->stageF);
    } else {
        hipLaunchKernelGGL(k_hinge_eval, dim3(blocks_for(H)), dim3(256), 0, c->stream, H, c->topo.F, c->hinge, c->x,
                           c->X, c->beta, (const double *)c->beta_hinge, s, c->stageK, c->stageF);
    }
    return hipGetLastError();
}

hipError_t launch_update_x(eolc_ctx *c, double h)
{
    int n = 3 * c->topo.N;
    hipLaunchKernelGGL(k_update_x, dim3(blocks_for(n)), dim3(256), 0, c->stream, n, h, c->v, c->x);
    return hipGetLastError();
}

}  // namespace eolc
