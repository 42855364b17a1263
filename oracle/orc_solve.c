/*
 * orc_solve.c -- solver "none" and one time step (oracle, FP64).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Cloth::solve (Cloth.cpp:343-357) forms b = -(M v + h f) and hands
 * (MDK, b, Aeq, beq) to GeneralizedSolver::velocitySolve
 * (GeneralizedSolver.cpp:107-172), whose QP minimises 1/2 v^T MDK v + b^T v
 * subject to the fixed-DOF rows Aeq v = beq.  With only single-DOF equality
 * rows (Constraints.cpp:114-119, 472-496: beq = 0.99 v_node + scripted) the QP
 * solution is the Dirichlet-reduced linear solve
 *      MDK_FF v_F = (M v + h f)_F - MDK_FD v_D ,   v_D = fixed values,
 * which is what the dead ConjugateGradient branch (GeneralizedSolver.cpp:120-137)
 * intended.  Solved here with Jacobi PCG (Eigen's default preconditioner) or
 * 3x3 block-Jacobi PCG, stopping at ||r||_2 <= tol ||b_red||_2, or with dense
 * Gaussian elimination for small systems.
 *
 * orc_step: Cloth::step (Cloth.cpp:359-413) without the MATLAB dumps:
 * forces, rhs, solve, then x += h v (Cloth.cpp:394-410).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static double now_s(void)
{
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

static double dotp(const double *a, const double *b, int n, int nt)
{
    double s = 0.0;
#pragma omp parallel for num_threads(nt) reduction(+ : s) schedule(static)
    for (int i = 0; i < n; ++i) s += a[i] * b[i];
    return s;
}

static double csr_get(const orc_csr *A, int r, int c)
{
    for (int64_t k = A->rowptr[r]; k < A->rowptr[r + 1]; ++k)
        if (A->col[k] == c) return A->val[k];
    return 0.0;
}

static int inv3(const double *a, double *o)
{
    double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
    double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    if (!(det != 0.0) || !isfinite(det)) return 1;
    double id = 1.0 / det;
    o[0] = c00 * id; o[1] = (a[2] * a[7] - a[1] * a[8]) * id; o[2] = (a[1] * a[5] - a[2] * a[4]) * id;
    o[3] = c01 * id; o[4] = (a[0] * a[8] - a[2] * a[6]) * id; o[5] = (a[2] * a[3] - a[0] * a[5]) * id;
    o[6] = c02 * id; o[7] = (a[1] * a[6] - a[0] * a[7]) * id; o[8] = (a[0] * a[4] - a[1] * a[3]) * id;
    return 0;
}

int orc_velocity_solve(const orc_csr *A, const double *rhs,
                       int nfixed, const int *fixed_dof, const double *fixed_val,
                       const double *v0, double tol, int maxit, int precond,
                       double *v, orc_solve_stats *st, int nt)
{
    int n = A->n;
    if (nt < 1) nt = 1;
    memset(st, 0, sizeof(*st));
    if (n % 3 != 0) { st->status = 3; return 3; }
    unsigned char *isfix = (unsigned char *)calloc((size_t)n, 1);
    for (int i = 0; i < nfixed; ++i) {
        if (fixed_dof[i] < 0 || fixed_dof[i] >= n) { free(isfix); st->status = 3; return 3; }
        isfix[fixed_dof[i]] = 1;
    }
    double *vd = (double *)calloc((size_t)n, sizeof(double));
    for (int i = 0; i < nfixed; ++i) vd[fixed_dof[i]] = fixed_val[i];
    double *b = (double *)malloc(sizeof(double) * n), *r = (double *)malloc(sizeof(double) * n);
    double *z = (double *)malloc(sizeof(double) * n), *p = (double *)malloc(sizeof(double) * n);
    double *q = (double *)malloc(sizeof(double) * n);
    /* b_red = rhs - A v_D on free DOFs */
    orc_csr_matvec(A, vd, q, nt);
    for (int i = 0; i < n; ++i) b[i] = isfix[i] ? 0.0 : rhs[i] - q[i];
    /* preconditioner */
    double *Dinv = (double *)malloc(sizeof(double) * 9 * (n / 3));
    int N = n / 3;
    for (int vtx = 0; vtx < N; ++vtx) {
        double blk[9];
        for (int a = 0; a < 3; ++a)
            for (int c = 0; c < 3; ++c) {
                int ra = 3 * vtx + a, cc = 3 * vtx + c;
                double val;
                if (isfix[ra] || isfix[cc]) val = (a == c) ? 1.0 : 0.0;
                else if (precond == ORC_PC_JACOBI && a != c) val = 0.0;
                else val = csr_get(A, ra, cc);
                blk[3 * a + c] = val;
            }
        if (inv3(blk, Dinv + 9 * vtx)) {
            for (int k = 0; k < 9; ++k) Dinv[9 * vtx + k] = (k % 4 == 0) ? 1.0 : 0.0;
        }
    }
    /* initial guess */
    for (int i = 0; i < n; ++i) v[i] = isfix[i] ? 0.0 : (v0 ? v0[i] : 0.0);
    orc_csr_matvec(A, v, q, nt);
    for (int i = 0; i < n; ++i) r[i] = isfix[i] ? 0.0 : b[i] - q[i];
    double bn = sqrt(dotp(b, b, n, nt));
    if (bn == 0.0) bn = 1.0;
    int status = 1, it = 0;
    double rn = sqrt(dotp(r, r, n, nt));
    double rz = 0.0;
    if (rn <= tol * bn) status = 0;
    else {
        for (int vtx = 0; vtx < N; ++vtx)
            for (int a = 0; a < 3; ++a) {
                const double *D = Dinv + 9 * vtx + 3 * a;
                z[3 * vtx + a] = D[0] * r[3 * vtx] + D[1] * r[3 * vtx + 1] + D[2] * r[3 * vtx + 2];
            }
        memcpy(p, z, sizeof(double) * n);
        rz = dotp(r, z, n, nt);
        for (it = 1; it <= maxit; ++it) {
            orc_csr_matvec(A, p, q, nt);
            for (int i = 0; i < n; ++i) if (isfix[i]) q[i] = 0.0;
            double pq = dotp(p, q, n, nt);
            if (!(pq > 0.0)) { status = 2; break; }
            double alpha = rz / pq;
            for (int i = 0; i < n; ++i) { v[i] += alpha * p[i]; r[i] -= alpha * q[i]; }
            rn = sqrt(dotp(r, r, n, nt));
            if (rn <= tol * bn) { status = 0; break; }
            for (int vtx = 0; vtx < N; ++vtx)
                for (int a = 0; a < 3; ++a) {
                    const double *D = Dinv + 9 * vtx + 3 * a;
                    z[3 * vtx + a] = D[0] * r[3 * vtx] + D[1] * r[3 * vtx + 1] + D[2] * r[3 * vtx + 2];
                }
            double rz2 = dotp(r, z, n, nt);
            double beta = rz2 / rz;
            rz = rz2;
            for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
        }
        if (it > maxit) it = maxit;
    }
    for (int i = 0; i < n; ++i) if (isfix[i]) v[i] = vd[i];
    /* true residual of the reduced system */
    orc_csr_matvec(A, v, q, nt);
    double tr = 0.0;
    for (int i = 0; i < n; ++i) if (!isfix[i]) { double d = rhs[i] - q[i]; tr += d * d; }
    st->iters = it;
    st->relres = rn / bn;
    st->true_relres = sqrt(tr) / bn;
    st->status = status;
    free(isfix); free(vd); free(b); free(r); free(z); free(p); free(q); free(Dinv);
    return status;
}

int orc_dense_velocity_solve(const orc_csr *A, const double *rhs,
                             int nfixed, const int *fixed_dof, const double *fixed_val,
                             double *v)
{
    int n = A->n;
    int *map = (int *)malloc(sizeof(int) * n);
    unsigned char *isfix = (unsigned char *)calloc((size_t)n, 1);
    double *vd = (double *)calloc((size_t)n, sizeof(double));
    for (int i = 0; i < nfixed; ++i) { isfix[fixed_dof[i]] = 1; vd[fixed_dof[i]] = fixed_val[i]; }
    int m = 0;
    for (int i = 0; i < n; ++i) map[i] = isfix[i] ? -1 : m++;
    double *D = (double *)calloc((size_t)m * m, sizeof(double));
    double *bb = (double *)calloc((size_t)m, sizeof(double));
    for (int r = 0; r < n; ++r) {
        if (isfix[r]) continue;
        double s = rhs[r];
        for (int64_t k = A->rowptr[r]; k < A->rowptr[r + 1]; ++k) {
            int c = A->col[k];
            if (isfix[c]) s -= A->val[k] * vd[c];
            else D[(size_t)map[r] * m + map[c]] = A->val[k];
        }
        bb[map[r]] = s;
    }
    /* Gaussian elimination with partial pivoting */
    int rc = 0;
    for (int k = 0; k < m; ++k) {
        int piv = k;
        double best = fabs(D[(size_t)k * m + k]);
        for (int i = k + 1; i < m; ++i) {
            double a = fabs(D[(size_t)i * m + k]);
            if (a > best) { best = a; piv = i; }
        }
        if (best == 0.0) { rc = 2; break; }
        if (piv != k) {
            for (int j = 0; j < m; ++j) {
                double t = D[(size_t)k * m + j]; D[(size_t)k * m + j] = D[(size_t)piv * m + j]; D[(size_t)piv * m + j] = t;
            }
            double t = bb[k]; bb[k] = bb[piv]; bb[piv] = t;
        }
        double dk = D[(size_t)k * m + k];
        for (int i = k + 1; i < m; ++i) {
            double l = D[(size_t)i * m + k] / dk;
            if (l == 0.0) continue;
            for (int j = k; j < m; ++j) D[(size_t)i * m + j] -= l * D[(size_t)k * m + j];
            bb[i] -= l * bb[k];
        }
    }
    if (rc == 0) {
        for (int k = m - 1; k >= 0; --k) {
            double s = bb[k];
            for (int j = k + 1; j < m; ++j) s -= D[(size_t)k * m + j] * bb[j];
            bb[k] = s / D[(size_t)k * m + k];
        }
        for (int i = 0; i < n; ++i) v[i] = isfix[i] ? vd[i] : bb[map[i]];
    }
    free(map); free(isfix); free(vd); free(D); free(bb);
    return rc;
}

int orc_step(int N, double *x, double *v, const double *X,
             int F, const int *tri, int H, const int *hinge,
             const orc_material *mat, const double *grav, double h, int flags,
             int nfixed_nodes, const int *fixed_nodes, const double *fixed_scripted_vel,
             double tol, int maxit, int precond, orc_solve_stats *st, int nt,
             double *ph)
{
    int n = 3 * N;
    double t0 = now_s();
    double *f = (double *)malloc(sizeof(double) * n);
    orc_csr M, MDK;
    int rc = orc_forces_fill(N, x, X, F, tri, H, hinge, mat, grav, h, flags, f, &M, &MDK, nt);
    double t1 = now_s();
    double *rhs = (double *)malloc(sizeof(double) * n);
    orc_csr_matvec(&M, v, rhs, nt);
    for (int i = 0; i < n; ++i) rhs[i] = rhs[i] + h * f[i];
    int nf = 3 * nfixed_nodes;
    int *fd = (int *)malloc(sizeof(int) * (nf ? nf : 1));
    double *fv = (double *)malloc(sizeof(double) * (nf ? nf : 1));
    for (int i = 0; i < nfixed_nodes; ++i)
        for (int k = 0; k < 3; ++k) {
            /* Constraints.cpp:114-119: beq = (1 - 0.01) v_node + scripted */
            fd[3 * i + k] = 3 * fixed_nodes[i] + k;
            fv[3 * i + k] = (1 - 0.01) * v[3 * fixed_nodes[i] + k] +
                            (fixed_scripted_vel ? fixed_scripted_vel[3 * i + k] : 0.0);
        }
    double t2 = now_s();
    double *vn = (double *)malloc(sizeof(double) * n);
    int src = orc_velocity_solve(&MDK, rhs, nf, fd, fv, v, tol, maxit, precond, vn, st, nt);
    double t3 = now_s();
    for (int i = 0; i < n; ++i) { v[i] = vn[i]; x[i] = x[i] + h * v[i]; }
    double t4 = now_s();
    if (ph) { ph[0] = t1 - t0; ph[1] = t2 - t1; ph[2] = t3 - t2; ph[3] = t4 - t3; ph[4] = t4 - t0; }
    orc_csr_free(&M); orc_csr_free(&MDK);
    free(f); free(rhs); free(fd); free(fv); free(vn);
    return rc ? rc : src;
}
