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

int orc_forces_fill(int N, const double *x, const double *X,
                    int F, const int *tri, int H, const int *hinge,
                    const orc_material *mat, const double *grav, double h, int flags,
                    double *f, orc_csr *M, orc_csr *MDK, int nthreads)
{
    int n = 3 * N;
    if (nthreads < 1) nthreads = 1;
    /* ---- element evaluation (parallel; outputs per element) ---- */
    double *Fm = (double *)malloc(sizeof(double) * 9 * (size_t)F);   /* fm + fi */
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
        for (int k = 0; k < 9; ++k) Fm[9 * (size_t)i + k] = fm[k] + fi[k];
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
    /* ---- f (face order, then optional bending) ---- */
    memset(f, 0, sizeof(double) * (size_t)n);
    for (int i = 0; i < F; ++i)
        for (int v = 0; v < 3; ++v)
            for (int k = 0; k < 3; ++k) f[3 * tri[3 * i + v] + k] += Fm[9 * (size_t)i + 3 * v + k];
    if (flags & ORC_INCLUDE_BENDING_FORCE)
        for (int e = 0; e < H; ++e)
            for (int v = 0; v < 4; ++v)
                for (int k = 0; k < 3; ++k) f[3 * hinge[4 * e + v] + k] += Fb[12 * (size_t)e + 3 * v + k];
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
        }
    }
    free(Fm); free(Km); free(Mi); free(Fb); free(Kb);
    tb_finish(&tm, M, nthreads);
    tb_finish(&tk, MDK, nthreads);
    return bad ? 3 : 0;
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
