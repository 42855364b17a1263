/*
 * orc_assembly.c -- Forces::fill restated (oracle, FP64).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Follows Forces.cpp:912-930:
 *   faceBasedF (Forces.cpp:331-520, non-EOL branch 498-518): per face, P and Q
 *     (357-372), membrane + inertial elements, f[v] += fm_v + fi_v, and the
 *     triplets of fillxMI / fillxxMI (103-125): M blocks Mi and MDK blocks
 *     Mi + dampingB*h*h*Km for (a,a),(b,b),(c,c),(a,b),(a,c),(b,c); every
 *     off-diagonal block is pushed together with its transpose.
 *   edgeBasedF (Forces.cpp:685-910, non-EOL branch 885-908): per interior edge,
 *     the bending element and the triplets of fillxB / fillxxB (522-539) with
 *     dampingB*h*h*Kb for the 4 diagonal and 6 off-diagonal blocks.  As in the
 *     reference the bending force is NOT added to f unless
 *     ORC_INCLUDE_BENDING_FORCE is set (the non-EOL branch never adds fb).
 *   setFromTriplets (Forces.cpp:928-929): duplicates are summed per (row, col)
 *     in triplet insertion order (Eigen's collapseDuplicates is a left fold
 *     over the insertion order), and columns end sorted within each row.
 *
 * Element evaluation runs in parallel; triplet insertion is serial in the
 * reference order so the summation order does not depend on thread count.
 */
#include "oracle.h"
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void orc_csr_free(orc_csr *A)
{
    free(A->rowptr); free(A->col); free(A->val);
    A->rowptr = NULL; A->col = NULL; A->val = NULL; A->nnz = 0;
}

typedef struct {
    int n;
    int64_t *cnt;     /* per row triplet count, then start offsets */
    int64_t *fillp;   /* per row fill cursor */
    int *col;
    double *val;
    int64_t total;
} trip_builder;

static void tb_init(trip_builder *tb, int n)
{
    tb->n = n;
    tb->cnt = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
    tb->fillp = NULL; tb->col = NULL; tb->val = NULL; tb->total = 0;
}
static void tb_alloc(trip_builder *tb)
{
    int64_t s = 0;
    for (int r = 0; r < tb->n; ++r) { int64_t c = tb->cnt[r]; tb->cnt[r] = s; s += c; }
    tb->cnt[tb->n] = s;
    tb->total = s;
    tb->fillp = (int64_t *)malloc(sizeof(int64_t) * ((size_t)tb->n + 1));
    memcpy(tb->fillp, tb->cnt, sizeof(int64_t) * ((size_t)tb->n + 1));
    tb->col = (int *)malloc(sizeof(int) * (size_t)(s ? s : 1));
    tb->val = (double *)malloc(sizeof(double) * (size_t)(s ? s : 1));
}
/* pass 0 counts, pass 1 inserts */
static inline void tb_push(trip_builder *tb, int pass, int r, int c, double v)
{
    if (pass == 0) { tb->cnt[r]++; return; }
    int64_t p = tb->fillp[r]++;
    tb->col[p] = c;
    tb->val[p] = v;
}

static int cmp_int_idx(const void *a, const void *b)
{
    const int *x = (const int *)a, *y = (const int *)b;
    return (*x > *y) - (*x < *y);
}

/* collapse duplicates per row (left fold in insertion order), then sort columns */
static void tb_finish(trip_builder *tb, orc_csr *A, int nthreads)
{
    int n = tb->n;
    int64_t *rownnz = (int64_t *)calloc((size_t)n + 1, sizeof(int64_t));
#pragma omp parallel num_threads(nthreads)
    {
        int *wi = (int *)malloc(sizeof(int) * (size_t)n);
        for (int i = 0; i < n; ++i) wi[i] = -1;
#pragma omp for schedule(dynamic, 256)
        for (int r = 0; r < n; ++r) {
            int64_t b = tb->cnt[r], e = tb->cnt[r + 1], o = b;
            for (int64_t k = b; k < e; ++k) {
                int c = tb->col[k];
                if (wi[c] >= 0) {
                    tb->val[wi[c]] = tb->val[wi[c]] + tb->val[k];
                } else {
                    wi[c] = (int)o;
                    tb->col[o] = c;
                    tb->val[o] = tb->val[k];
                    ++o;
                }
            }
            for (int64_t k = b; k < o; ++k) wi[tb->col[k]] = -1;
            rownnz[r] = o - b;
            /* sort (col, val) pairs by column: insertion sort on short rows */
            for (int64_t k = b + 1; k < o; ++k) {
                int c = tb->col[k]; double v = tb->val[k];
                int64_t q = k - 1;
                while (q >= b && tb->col[q] > c) { tb->col[q + 1] = tb->col[q]; tb->val[q + 1] = tb->val[q]; --q; }
                tb->col[q + 1] = c; tb->val[q + 1] = v;
            }
        }
        free(wi);
    }
    A->n = n;
    A->rowptr = (int64_t *)malloc(sizeof(int64_t) * ((size_t)n + 1));
    int64_t s = 0;
    for (int r = 0; r < n; ++r) { A->rowptr[r] = s; s += rownnz[r]; }
    A->rowptr[n] = s;
    A->nnz = s;
    A->col = (int *)malloc(sizeof(int) * (size_t)(s ? s : 1));
    A->val = (double *)malloc(sizeof(double) * (size_t)(s ? s : 1));
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int r = 0; r < n; ++r) {
        memcpy(A->col + A->rowptr[r], tb->col + tb->cnt[r], sizeof(int) * (size_t)rownnz[r]);
        memcpy(A->val + A->rowptr[r], tb->val + tb->cnt[r], sizeof(double) * (size_t)rownnz[r]);
    }
    free(rownnz);
    free(tb->cnt); free(tb->fillp); free(tb->col); free(tb->val);
    (void)cmp_int_idx;
}

/* ---- EOL extension (Forces.cpp:177-329, 399-497, 580-683, 746-884) ----------
 * An EOL node e (EoL_index k) has two more DOFs at 3N + 2k.  Each face or hinge
 * with an EOL vertex keeps its x-x triplets and force, and adds, through the
 * deformation gradient F (deform_grad, UtilEOL.cpp:13-28: Dx DX^-1 of the face;
 * (F1 + F2) / 2 of the two faces of a hinge), the blocks of fillEOLInertia /
 * fillEOLMembrane / fillEOLBending:
 *   f_X(v) = -F^T f_v,  K(X_v, X_w) = F^T K_vw F,  K(X_v, x_w) = -F^T K_vw,
 *   K(x_u, X_v) = -K_uv F,
 * pushed by fillXMI / fillXXMI / fillXxMI / fillxXMI (Forces.cpp:126-175) and their
 * bending twins (540-578) in the order faceBasedF / edgeBasedF call them: X
 * diagonal blocks, X-X pairs, then per EOL vertex v its (X_v, x_w) blocks for
 * w >= v and the (x_u, X_v) blocks for u < v.  An EOL hinge adds its bending force
 * to f (edgeBasedF's EOL branch does, the non-EOL branch does not). */
typedef struct {
    const int *eidx;   /* per node: EoL_index or -1; NULL: no EOL */
    int E;
} eol_info;

/* deform_grad of the face (a, b, c): Dx DX^-1 (3x2, row-major F[r][c]) */
static void deform_grad(const double *x, const double *X, int a, int b, int c, double Fo[3][2])
{
    double Dx[3][2], DX[2][2];
    for (int r = 0; r < 3; ++r) {
        Dx[r][0] = x[3 * b + r] - x[3 * a + r];
        Dx[r][1] = x[3 * c + r] - x[3 * a + r];
    }
    DX[0][0] = X[2 * b] - X[2 * a];
    DX[0][1] = X[2 * c] - X[2 * a];
    DX[1][0] = X[2 * b + 1] - X[2 * a + 1];
    DX[1][1] = X[2 * c + 1] - X[2 * a + 1];
    const double det = DX[0][0] * DX[1][1] - DX[0][1] * DX[1][0];
    const double inv[2][2] = { { DX[1][1] / det, -DX[0][1] / det }, { -DX[1][0] / det, DX[0][0] / det } };
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 2; ++k) Fo[r][k] = Dx[r][0] * inv[0][k] + Dx[r][1] * inv[1][k];
}

/* element block (v, w) entry (j, k) of a column-major nv*3 square matrix */
static inline double eblk(const double *K, int nv, int v, int w, int j, int k)
{
    return K[(size_t)(3 * w + k) * (3 * nv) + 3 * v + j];
}

/* Pushes the EOL triplets of one element.  Mm: mass matrix or NULL (hinges);
 * K: stiffness; sK: factor on K (dampingB h^2); eo[v]: X offset or -1. */
static void push_eol(trip_builder *tm, trip_builder *tk, int pass, int nv, const int *xo, const int *eo,
                     double Fm_[3][2], const double *Mm, const double *K, double sK)
{
    /* value of a 2x2 X-X block entry and of the 2x3 / 3x2 X-x entries */
#define KV(v, w, j, k) eblk(K, nv, v, w, j, k)
#define MV(v, w, j, k) (Mm ? eblk(Mm, nv, v, w, j, k) : 0.0)
    double (*Fv)[2] = Fm_;
    /* X diagonal blocks: F^T K_vv F */
    for (int v = 0; v < nv; ++v) {
        if (eo[v] < 0) continue;
        for (int j = 0; j < 2; ++j)
            for (int k = 0; k < 2; ++k) {
                double kk = 0.0, mm = 0.0;
                for (int a = 0; a < 3; ++a)
                    for (int b = 0; b < 3; ++b) {
                        kk += Fv[a][j] * KV(v, v, a, b) * Fv[b][k];
                        mm += Fv[a][j] * MV(v, v, a, b) * Fv[b][k];
                    }
                if (Mm) tb_push(tm, pass, eo[v] + j, eo[v] + k, mm);
                tb_push(tk, pass, eo[v] + j, eo[v] + k, Mm ? mm + sK * kk : sK * kk);
            }
    }
    /* X-X pairs v < w */
    for (int v = 0; v < nv; ++v)
        for (int w = v + 1; w < nv; ++w) {
            if (eo[v] < 0 || eo[w] < 0) continue;
            for (int j = 0; j < 2; ++j)
                for (int k = 0; k < 2; ++k) {
                    double kk = 0.0, mm = 0.0;
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) {
                            kk += Fv[a][j] * KV(v, w, a, b) * Fv[b][k];
                            mm += Fv[a][j] * MV(v, w, a, b) * Fv[b][k];
                        }
                    const double md = Mm ? mm + sK * kk : sK * kk;
                    if (Mm) {
                        tb_push(tm, pass, eo[v] + j, eo[w] + k, mm);
                        tb_push(tm, pass, eo[w] + k, eo[v] + j, mm);
                    }
                    tb_push(tk, pass, eo[v] + j, eo[w] + k, md);
                    tb_push(tk, pass, eo[w] + k, eo[v] + j, md);
                }
        }
    /* X-x: per EOL v, (X_v, x_w) for w >= v, then (x_u, X_v) for u < v */
    for (int v = 0; v < nv; ++v) {
        if (eo[v] < 0) continue;
        for (int w = v; w < nv; ++w)
            for (int j = 0; j < 2; ++j)
                for (int k = 0; k < 3; ++k) {
                    double kk = 0.0, mm = 0.0;
                    for (int a = 0; a < 3; ++a) {
                        kk -= Fv[a][j] * KV(v, w, a, k);
                        mm -= Fv[a][j] * MV(v, w, a, k);
                    }
                    const double md = Mm ? mm + sK * kk : sK * kk;
                    if (Mm) {
                        tb_push(tm, pass, eo[v] + j, xo[w] + k, mm);
                        tb_push(tm, pass, xo[w] + k, eo[v] + j, mm);
                    }
                    tb_push(tk, pass, eo[v] + j, xo[w] + k, md);
                    tb_push(tk, pass, xo[w] + k, eo[v] + j, md);
                }
        for (int u = 0; u < v; ++u)
            for (int j = 0; j < 3; ++j)
                for (int k = 0; k < 2; ++k) {
                    double kk = 0.0, mm = 0.0;
                    for (int a = 0; a < 3; ++a) {
                        kk -= KV(u, v, j, a) * Fv[a][k];
                        mm -= MV(u, v, j, a) * Fv[a][k];
                    }
                    const double md = Mm ? mm + sK * kk : sK * kk;
                    if (Mm) {
                        tb_push(tm, pass, xo[u] + j, eo[v] + k, mm);
                        tb_push(tm, pass, eo[v] + k, xo[u] + j, mm);
                    }
                    tb_push(tk, pass, xo[u] + j, eo[v] + k, md);
                    tb_push(tk, pass, eo[v] + k, xo[u] + j, md);
                }
    }
#undef KV
#undef MV
}

static int fill_impl(int N, const double *x, const double *X,
                     int F, const int *tri, int H, const int *hinge,
                     const orc_material *mat, const double *grav, double h, int flags,
                     double *f, orc_csr *M, orc_csr *MDK, int nthreads, eol_info eol)
{
    int n = 3 * N + 2 * eol.E;
    if (nthreads < 1) nthreads = 1;
    /* ---- element evaluation (parallel; outputs per element) ---- */
    double *Fm = (double *)malloc(sizeof(double) * 9 * (size_t)F);   /* fm + fi */
    double *Fmem = (double *)malloc(sizeof(double) * 9 * (size_t)(F ? F : 1));   /* fm */
    double *Fin = (double *)malloc(sizeof(double) * 9 * (size_t)(F ? F : 1));    /* fi */
    double *Km = (double *)malloc(sizeof(double) * 81 * (size_t)F);
    double *Mi = (double *)malloc(sizeof(double) * 81 * (size_t)F);
    double *Fb = (double *)malloc(sizeof(double) * 12 * (size_t)(H ? H : 1));
    double *Kb = (double *)malloc(sizeof(double) * 144 * (size_t)(H ? H : 1));
    int bad = 0;
#pragma omp parallel for num_threads(nthreads) schedule(static) reduction(| : bad)
    for (int i = 0; i < F; ++i) {
        const int *t = tri + 3 * i;
        const double *xa = x + 3 * t[0], *xb = x + 3 * t[1], *xc = x + 3 * t[2];
        const double *Xa = X + 2 * t[0], *Xb = X + 2 * t[1], *Xc = X + 2 * t[2];
        double P[6], Q[4], Wm, Wi, fm[9], fi[9];
        orc_face_PQ(xa, xb, xc, Xa, Xb, Xc, P, Q);
        orc_membrane(xa, xb, xc, Xa, Xb, Xc, mat->e, mat->nu, P, Q, &Wm, fm, Km + 81 * (size_t)i);
        orc_inertial(xa, xb, xc, Xa, Xb, Xc, grav, mat->density, &Wi, fi, Mi + 81 * (size_t)i);
        for (int k = 0; k < 9; ++k) {
            Fm[9 * (size_t)i + k] = fm[k] + fi[k];
            Fmem[9 * (size_t)i + k] = fm[k];
            Fin[9 * (size_t)i + k] = fi[k];
        }
        if ((Xb[0] - Xa[0]) * (Xc[1] - Xa[1]) - (Xb[1] - Xa[1]) * (Xc[0] - Xa[0]) <= 0.0) bad = 1;
    }
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int e = 0; e < H; ++e) {
        const int *q = hinge + 4 * e;
        double Wb;
        orc_bending(x + 3 * q[0], x + 3 * q[1], x + 3 * q[2], x + 3 * q[3],
                    X + 2 * q[0], X + 2 * q[1], X + 2 * q[2], X + 2 * q[3],
                    mat->beta, &Wb, Fb + 12 * (size_t)e, Kb + 144 * (size_t)e);
    }
    /* ---- f (face order, then the bending of EOL hinges, or of all hinges when
     * asked for); the X rows of an EOL vertex get -F^T f_v per element, with the
     * membrane and inertial parts mapped separately and then added, as
     * faceBasedF does ---- */
    memset(f, 0, sizeof(double) * (size_t)n);
    for (int i = 0; i < F; ++i) {
        const int *t = tri + 3 * i;
        for (int v = 0; v < 3; ++v)
            for (int k = 0; k < 3; ++k) f[3 * t[v] + k] += Fm[9 * (size_t)i + 3 * v + k];
        if (eol.eidx && (eol.eidx[t[0]] >= 0 || eol.eidx[t[1]] >= 0 || eol.eidx[t[2]] >= 0)) {
            double Fd[3][2];
            deform_grad(x, X, t[0], t[1], t[2], Fd);
            for (int v = 0; v < 3; ++v) {
                if (eol.eidx[t[v]] < 0) continue;
                const int o = 3 * N + 2 * eol.eidx[t[v]];
                for (int j = 0; j < 2; ++j) {
                    double a = 0.0, b = 0.0;
                    for (int r = 0; r < 3; ++r) {
                        a -= Fd[r][j] * Fmem[9 * (size_t)i + 3 * v + r];
                        b -= Fd[r][j] * Fin[9 * (size_t)i + 3 * v + r];
                    }
                    f[o + j] += a + b;
                }
            }
        }
    }
    for (int e = 0; e < H; ++e) {
        const int *q = hinge + 4 * e;
        const int is_eol = eol.eidx && (eol.eidx[q[0]] >= 0 || eol.eidx[q[1]] >= 0 || eol.eidx[q[2]] >= 0 ||
                                        eol.eidx[q[3]] >= 0);
        if (!is_eol && !(flags & ORC_INCLUDE_BENDING_FORCE)) continue;
        for (int v = 0; v < 4; ++v)
            for (int k = 0; k < 3; ++k) f[3 * q[v] + k] += Fb[12 * (size_t)e + 3 * v + k];
        if (!is_eol) continue;
        double F1[3][2], F2[3][2], Fh[3][2];
        deform_grad(x, X, q[0], q[1], q[2], F1);
        deform_grad(x, X, q[1], q[0], q[3], F2);
        for (int r = 0; r < 3; ++r)
            for (int k = 0; k < 2; ++k) Fh[r][k] = (F1[r][k] + F2[r][k]) / 2;
        for (int v = 0; v < 4; ++v) {
            if (eol.eidx[q[v]] < 0) continue;
            const int o = 3 * N + 2 * eol.eidx[q[v]];
            for (int j = 0; j < 2; ++j) {
                double a = 0.0;
                for (int r = 0; r < 3; ++r) a -= Fh[r][j] * Fb[12 * (size_t)e + 3 * v + r];
                f[o + j] += a;
            }
        }
    }
    /* ---- triplets in reference order ---- */
    trip_builder tm, tk;
    tb_init(&tm, n);
    tb_init(&tk, n);
    const double dB = mat->dampingB;
    static const int fpairs[6][2] = { { 0, 0 }, { 1, 1 }, { 2, 2 }, { 0, 1 }, { 0, 2 }, { 1, 2 } };
    static const int hpairs[10][2] = { { 0, 0 }, { 1, 1 }, { 2, 2 }, { 3, 3 }, { 0, 1 },
                                       { 0, 2 }, { 0, 3 }, { 1, 2 }, { 1, 3 }, { 2, 3 } };
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) { tb_alloc(&tm); tb_alloc(&tk); }
        for (int i = 0; i < F; ++i) {
            const double *K = Km + 81 * (size_t)i, *Mm = Mi + 81 * (size_t)i;
            for (int pb = 0; pb < 6; ++pb) {
                int v = fpairs[pb][0], w = fpairs[pb][1];
                int i0 = 3 * tri[3 * i + v], i1 = 3 * tri[3 * i + w];
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k) {
                        /* Kme = Map<MatrixXd>(Km) is column-major: Kme(r,c) = Km[c*9+r] */
                        double kk = K[(3 * w + k) * 9 + 3 * v + j];
                        double mm = Mm[(3 * w + k) * 9 + 3 * v + j];
                        double md = mm + dB * h * h * kk;
                        if (v == w) {
                            tb_push(&tm, pass, i0 + j, i0 + k, mm);
                            tb_push(&tk, pass, i0 + j, i0 + k, md);
                        } else {
                            tb_push(&tm, pass, i0 + j, i1 + k, mm);
                            tb_push(&tm, pass, i1 + k, i0 + j, mm);
                            tb_push(&tk, pass, i0 + j, i1 + k, md);
                            tb_push(&tk, pass, i1 + k, i0 + j, md);
                        }
                    }
            }
            if (eol.eidx) {
                const int *t = tri + 3 * i;
                int xo[3], eo[3], any = 0;
                for (int v = 0; v < 3; ++v) {
                    xo[v] = 3 * t[v];
                    eo[v] = eol.eidx[t[v]] >= 0 ? 3 * N + 2 * eol.eidx[t[v]] : -1;
                    any |= eo[v] >= 0;
                }
                if (any) {
                    double Fd[3][2];
                    deform_grad(x, X, t[0], t[1], t[2], Fd);
                    push_eol(&tm, &tk, pass, 3, xo, eo, Fd, Mm, K, dB * h * h);
                }
            }
        }
        for (int e = 0; e < H; ++e) {
            const double *K = Kb + 144 * (size_t)e;
            for (int pb = 0; pb < 10; ++pb) {
                int v = hpairs[pb][0], w = hpairs[pb][1];
                int i0 = 3 * hinge[4 * e + v], i1 = 3 * hinge[4 * e + w];
                for (int j = 0; j < 3; ++j)
                    for (int k = 0; k < 3; ++k) {
                        double kk = dB * h * h * K[(3 * w + k) * 12 + 3 * v + j];
                        if (v == w) {
                            tb_push(&tk, pass, i0 + j, i0 + k, kk);
                        } else {
                            tb_push(&tk, pass, i0 + j, i1 + k, kk);
                            tb_push(&tk, pass, i1 + k, i0 + j, kk);
                        }
                    }
            }
            if (eol.eidx) {
                const int *q = hinge + 4 * e;
                int xo[4], eo[4], any = 0;
                for (int v = 0; v < 4; ++v) {
                    xo[v] = 3 * q[v];
                    eo[v] = eol.eidx[q[v]] >= 0 ? 3 * N + 2 * eol.eidx[q[v]] : -1;
                    any |= eo[v] >= 0;
                }
                if (any) {
                    double F1[3][2], F2[3][2], Fh[3][2];
                    deform_grad(x, X, q[0], q[1], q[2], F1);
                    deform_grad(x, X, q[1], q[0], q[3], F2);
                    for (int r = 0; r < 3; ++r)
                        for (int k = 0; k < 2; ++k) Fh[r][k] = (F1[r][k] + F2[r][k]) / 2;
                    push_eol(&tm, &tk, pass, 4, xo, eo, Fh, NULL, K, dB * h * h);
                }
            }
        }
    }
    free(Fm); free(Fmem); free(Fin); free(Km); free(Mi); free(Fb); free(Kb);
    tb_finish(&tm, M, nthreads);
    tb_finish(&tk, MDK, nthreads);
    return bad ? 3 : 0;
}

int orc_forces_fill(int N, const double *x, const double *X,
                    int F, const int *tri, int H, const int *hinge,
                    const orc_material *mat, const double *grav, double h, int flags,
                    double *f, orc_csr *M, orc_csr *MDK, int nthreads)
{
    eol_info none = { NULL, 0 };
    return fill_impl(N, x, X, F, tri, H, hinge, mat, grav, h, flags, f, M, MDK, nthreads, none);
}

int orc_forces_fill_eol(int N, const double *x, const double *X,
                        int F, const int *tri, int H, const int *hinge,
                        const orc_material *mat, const double *grav, double h, int flags,
                        int n_eol, const int *eol_nodes, double *f, orc_csr *M, orc_csr *MDK, int nthreads)
{
    int *eidx = (int *)malloc(sizeof(int) * (size_t)(N ? N : 1));
    for (int i = 0; i < N; ++i) eidx[i] = -1;
    for (int k = 0; k < n_eol; ++k) {
        if (eol_nodes[k] < 0 || eol_nodes[k] >= N || eidx[eol_nodes[k]] >= 0) { free(eidx); return 1; }
        eidx[eol_nodes[k]] = k;
    }
    eol_info info = { eidx, n_eol };
    const int rc = fill_impl(N, x, X, F, tri, H, hinge, mat, grav, h, flags, f, M, MDK, nthreads, info);
    free(eidx);
    return rc;
}

void orc_csr_matvec(const orc_csr *A, const double *x, double *y, int nthreads)
{
    if (nthreads < 1) nthreads = 1;
#pragma omp parallel for num_threads(nthreads) schedule(static)
    for (int r = 0; r < A->n; ++r) {
        double s = 0.0;
        for (int64_t k = A->rowptr[r]; k < A->rowptr[r + 1]; ++k) s += A->val[k] * x[A->col[k]];
        y[r] = s;
    }
}
