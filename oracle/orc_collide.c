/*
 * orc_collide.c -- point/box-vs-cloth proximity queries (oracle, FP64).
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 * Restates btc::pointTriCollision (boxTriCollision.cpp:1067-1224) and
 * btc::boxTriCollision (boxTriCollision.cpp:602-1063) with their helpers:
 * createEdges (141-231), createFaceNormals (233-247), createVertNormals
 * (249-283), box tables (289-392), createBox (400-420), AABBs and check_AABB
 * (424-486), barycentric (488-505), lineline (507-525), linepoint (527-536),
 * intersect_square (550-600) and intersect_triangle3_inc (raytri.cpp:259-316).
 *
 * The reference algorithm is brute force, so this restatement is brute force
 * too.  Vector arithmetic is written out in the evaluation order Eigen uses for
 * 3-vectors (dot = (a0 b0 + a1 b1) + a2 b2, cross component-wise,
 * normalized() = v / sqrt(squaredNorm) when squaredNorm > 0) and this file is
 * compiled with -ffp-contract=off, so results are reproducible bit for bit.
 *
 * Quirks kept on purpose (they change which pairs are reported):
 *  - jitter is regenerated from seed 1 on every call (649-659, 1079-1090);
 *  - the 1e-3 AABB pad ignores the threshold (473);
 *  - V2-T1 reads the cloth normal of FACE i2 for VERTEX i2 (731);
 *  - createEdges' hash kmin + (n+1) kmax is a 32-bit int and wraps (170);
 *  - intersect_triangle3_inc's "-ZERO" expands to +1e-6 (raytri.cpp:258-300),
 *    so the inclusive test is in fact a 1e-6-exclusive one;
 *  - cloth edge normals (and their angle) come from the un-jittered positions
 *    because createEdges runs before the copy is perturbed (612-614, 629).
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define CD_PI 3.14159265358979323846

static inline void v_sub(const double *a, const double *b, double *o) { o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2]; }
static inline void v_add(const double *a, const double *b, double *o) { o[0] = a[0] + b[0]; o[1] = a[1] + b[1]; o[2] = a[2] + b[2]; }
static inline double v_dot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
static inline void v_cross(const double *a, const double *b, double *o)
{
    double c0 = a[1] * b[2] - a[2] * b[1], c1 = a[2] * b[0] - a[0] * b[2], c2 = a[0] * b[1] - a[1] * b[0];
    o[0] = c0; o[1] = c1; o[2] = c2;
}
static inline double v_norm(const double *a) { return sqrt(v_dot(a, a)); }
static inline void v_normalized(const double *a, double *o)
{
    double z = v_dot(a, a);
    if (z > 0.0) { double s = sqrt(z); o[0] = a[0] / s; o[1] = a[1] / s; o[2] = a[2] / s; }
    else { o[0] = a[0]; o[1] = a[1]; o[2] = a[2]; }
}
static inline void v_copy(const double *a, double *o) { o[0] = a[0]; o[1] = a[1]; o[2] = a[2]; }
static inline void v_neg(double *a) { a[0] = -a[0]; a[1] = -a[1]; a[2] = -a[2]; }

/* ---- box tables (boxTriCollision.cpp:289-392) ---------------------------- */
static const double box_verts_data[14][3] = {
    { -1, -1, -1 }, { -1, -1, 1 }, { -1, 1, -1 }, { -1, 1, 1 },
    { 1, -1, -1 },  { 1, -1, 1 },  { 1, 1, -1 },  { 1, 1, 1 },
    { -1, 0, 0 },   { 1, 0, 0 },   { 0, -1, 0 },  { 0, 1, 0 },
    { 0, 0, -1 },   { 0, 0, 1 } };
static const int box_faces[24][3] = {
    { 0, 8, 2 },   { 1, 8, 0 },   { 3, 8, 1 },   { 2, 8, 3 },
    { 4, 10, 0 },  { 5, 10, 4 },  { 1, 10, 5 },  { 0, 10, 1 },
    { 6, 9, 4 },   { 7, 9, 6 },   { 5, 9, 7 },   { 4, 9, 5 },
    { 2, 11, 6 },  { 3, 11, 2 },  { 7, 11, 3 },  { 6, 11, 7 },
    { 1, 13, 3 },  { 5, 13, 1 },  { 7, 13, 5 },  { 3, 13, 7 },
    { 0, 12, 4 },  { 2, 12, 0 },  { 6, 12, 2 },  { 4, 12, 6 } };
static const int box_edge_verts[12][4] = {
    { 0, 1, 8, 10 }, { 2, 0, 8, 12 }, { 1, 3, 8, 13 }, { 3, 2, 8, 11 },
    { 0, 4, 10, 12 }, { 5, 1, 10, 13 }, { 4, 5, 10, 9 }, { 6, 2, 11, 12 },
    { 4, 6, 9, 12 }, { 3, 7, 11, 13 }, { 7, 5, 9, 13 }, { 6, 7, 9, 11 } };
static const int box_vert_edges[8][3] = {
    { 0, 1, 4 }, { 0, 2, 5 }, { 1, 3, 7 }, { 2, 3, 9 },
    { 4, 6, 8 }, { 5, 6, 10 }, { 7, 8, 11 }, { 9, 10, 11 } };
static const int box_edge_faces[12][2] = {
    { 1, 7 }, { 0, 21 }, { 2, 16 }, { 3, 13 }, { 4, 20 }, { 6, 17 },
    { 5, 11 }, { 12, 22 }, { 8, 23 }, { 14, 19 }, { 10, 18 }, { 9, 15 } };

typedef struct {
    double v[14][3];
    double fn[24][3];
    double vn[14][3];
    double en[12][2][3];
    double eangle[12];
    double faabb[24][6];
    double aabb[6];
} box_geom;

/* createFaceNormals (233-247): dba.cross(-dac).normalized() */
static void face_normal(const double *xa, const double *xb, const double *xc, double *n)
{
    double dba[3], dac[3], c[3];
    v_sub(xb, xa, dba);
    v_sub(xa, xc, dac);
    v_neg(dac);
    v_cross(dba, dac, c);
    v_normalized(c, n);
}

/* createBox (400-420) with createVertNormals (249-283) */
static void make_box(box_geom *B, const double *whd, const double *E1 /* col-major 4x4 */)
{
    double S[16] = { 0 }, E[16];
    S[0] = 0.5 * whd[0]; S[5] = 0.5 * whd[1]; S[10] = 0.5 * whd[2]; S[15] = 1.0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = E1[0 * 4 + i] * S[j * 4 + 0];
            for (int k = 1; k < 4; ++k) s = s + E1[k * 4 + i] * S[j * 4 + k];
            E[j * 4 + i] = s;
        }
    for (int c = 0; c < 14; ++c) {
        double vv[4] = { box_verts_data[c][0], box_verts_data[c][1], box_verts_data[c][2], 1.0 };
        for (int i = 0; i < 3; ++i) {
            double s = E[0 * 4 + i] * vv[0];
            for (int k = 1; k < 4; ++k) s = s + E[k * 4 + i] * vv[k];
            B->v[c][i] = s;
        }
    }
    for (int f = 0; f < 24; ++f)
        face_normal(B->v[box_faces[f][0]], B->v[box_faces[f][1]], B->v[box_faces[f][2]], B->fn[f]);
    double acc[14][3], ang[14];
    memset(acc, 0, sizeof(acc));
    memset(ang, 0, sizeof(ang));
    for (int f = 0; f < 24; ++f) {
        const double *xa = B->v[box_faces[f][0]], *xb = B->v[box_faces[f][1]], *xc = B->v[box_faces[f][2]];
        double dba[3], dcb[3], dac[3], nd[3], nor[3], t[3];
        v_sub(xb, xa, dba); v_sub(xc, xb, dcb); v_sub(xa, xc, dac);
        v_copy(dac, nd); v_neg(nd);
        v_cross(dba, nd, nor);
        v_normalized(nor, nor); v_normalized(dba, dba); v_normalized(dcb, dcb); v_normalized(dac, dac);
        v_copy(dac, t); v_neg(t);
        double a1 = acos(v_dot(dba, t));
        v_copy(dba, t); v_neg(t);
        double a2 = acos(v_dot(dcb, t));
        v_copy(dcb, t); v_neg(t);
        double a3 = acos(v_dot(dac, t));
        int i0 = box_faces[f][0], i1 = box_faces[f][1], i2 = box_faces[f][2];
        for (int k = 0; k < 3; ++k) {
            acc[i0][k] = acc[i0][k] + a1 * nor[k];
            acc[i1][k] = acc[i1][k] + a2 * nor[k];
            acc[i2][k] = acc[i2][k] + a3 * nor[k];
        }
        ang[i0] += a1; ang[i1] += a2; ang[i2] += a3;
    }
    for (int c = 0; c < 14; ++c) {
        double nn[3] = { acc[c][0] / ang[c], acc[c][1] / ang[c], acc[c][2] / ang[c] };
        v_normalized(nn, B->vn[c]);
    }
    for (int k = 0; k < 12; ++k) {
        v_copy(B->fn[box_edge_faces[k][0]], B->en[k][0]);
        v_copy(B->fn[box_edge_faces[k][1]], B->en[k][1]);
        B->eangle[k] = acos(v_dot(B->en[k][0], B->en[k][1]));
    }
    for (int i = 0; i < 3; ++i) {
        double lo = B->v[0][i], hi = B->v[0][i];
        for (int c = 1; c < 14; ++c) { if (B->v[c][i] < lo) lo = B->v[c][i]; if (B->v[c][i] > hi) hi = B->v[c][i]; }
        B->aabb[i] = lo; B->aabb[i + 3] = hi;
    }
    for (int f = 0; f < 24; ++f) {
        const double *xa = B->v[box_faces[f][0]], *xb = B->v[box_faces[f][1]], *xc = B->v[box_faces[f][2]];
        for (int i = 0; i < 3; ++i) {
            double lo = xa[i], hi = xa[i];
            lo = (lo < xb[i]) ? lo : xb[i];  /* std::min(xb, lo) */
            lo = (lo < xc[i]) ? lo : xc[i];
            hi = (xb[i] < hi) ? hi : xb[i];  /* std::max(xb, hi) */
            hi = (xc[i] < hi) ? hi : xc[i];
            B->faabb[f][i] = lo; B->faabb[f][i + 3] = hi;
        }
    }
}

/* check_AABB (462-486): both boxes padded by a fixed 1e-3 */
static int check_aabb(const double *a1, const double *a2)
{
    const double t = 1e-3;
    double mn1[3], mx1[3], mn2[3], mx2[3];
    for (int i = 0; i < 3; ++i) { mn1[i] = a1[i] - t; mx1[i] = a1[i + 3] + t; mn2[i] = a2[i] - t; mx2[i] = a2[i + 3] + t; }
    return mx1[0] >= mn2[0] && mn1[0] <= mx2[0] && mx1[1] >= mn2[1] && mn1[1] <= mx2[1] &&
           mx1[2] >= mn2[2] && mn1[2] <= mx2[2];
}

static void barycentric(double *alpha, double *beta, const double *a, const double *b, const double *c, const double *p)
{
    double v0[3], v1[3], v2[3];
    v_sub(b, a, v0); v_sub(c, a, v1); v_sub(p, a, v2);
    double d00 = v_dot(v0, v0), d01 = v_dot(v0, v1), d11 = v_dot(v1, v1);
    double d20 = v_dot(v2, v0), d21 = v_dot(v2, v1);
    double denom = d00 * d11 - d01 * d01;
    double bt = (d11 * d20 - d01 * d21) / denom;
    double gm = (d00 * d21 - d01 * d20) / denom;
    *beta = bt;
    *alpha = 1.0 - bt - gm;
}

static void lineline(double *a, double *b, const double *A1, const double *A2, const double *B1, const double *B2)
{
    double B2B1[3], A1B1[3], A2A1[3], X[3], t[3];
    v_sub(B2, B1, B2B1); v_sub(A1, B1, A1B1); v_sub(A2, A1, A2A1);
    v_cross(A2A1, B2B1, X);
    v_cross(B2B1, A1B1, t); double nA = v_dot(t, X);
    v_cross(A2A1, A1B1, t); double nB = v_dot(t, X);
    v_cross(A2A1, B2B1, t); double d = v_dot(t, X);
    *a = nA / d;
    *b = nB / d;
}

static double linepoint(const double *A, const double *B, const double *P)
{
    double AP[3], AB[3];
    v_sub(P, A, AP); v_sub(B, A, AB);
    double ab2 = v_dot(AB, AB), apab = v_dot(AP, AB);
    return apab / ab2;
}

/* raytri.cpp:259-316 with EPSILON 1e-6 and ZERO = -EPSILON as the macros expand */
static int tri3_inc(const double *orig, const double *dir, const double *v0, const double *v1, const double *v2, double *t)
{
    const double EPS = 1e-6, ZERO = -1e-6;
    double e1[3], e2[3], tv[3], pv[3], qv[3];
    v_sub(v1, v0, e1); v_sub(v2, v0, e2);
    v_cross(dir, e2, pv);
    double det = v_dot(e1, pv);
    v_sub(orig, v0, tv);
    double inv_det = 1.0 / det;
    v_cross(tv, e1, qv);
    double u, v;
    if (det > EPS) {
        u = v_dot(tv, pv);
        if (u < -ZERO || u > det + ZERO) return 0;
        v = v_dot(dir, qv);
        if (v < -ZERO || u + v > det + ZERO) return 0;
    } else if (det < -EPS) {
        u = v_dot(tv, pv);
        if (u > ZERO || u < det - ZERO) return 0;
        v = v_dot(dir, qv);
        if (v > ZERO || u + v < det - ZERO) return 0;
    } else return 0;
    *t = v_dot(e2, qv) * inv_det;
    return 1;
}

/* the oracle's ray/triangle test exported for its pin against the compiled raytri.cpp */
int orc_tri3_inc(const double *orig, const double *dir, const double *v0, const double *v1, const double *v2, double *t)
{
    return tri3_inc(orig, dir, v0, v1, v2, t);
}

static int intersect_square(const double *x0, const double *dx, const double *xa, const double *xb, const double *xc, double *t)
{
    double xd[3], xe[3], d[3];
    for (int i = 0; i < 3; ++i) { d[i] = xc[i] - xa[i]; xd[i] = xa[i] + 2.0 * d[i]; }
    for (int i = 0; i < 3; ++i) { d[i] = xc[i] - xb[i]; xe[i] = xb[i] + 2.0 * d[i]; }
    if (tri3_inc(x0, dx, xa, xb, xc, t)) return 1;
    if (tri3_inc(x0, dx, xb, xd, xc, t)) return 1;
    if (tri3_inc(x0, dx, xd, xe, xc, t)) return 1;
    if (tri3_inc(x0, dx, xe, xa, xc, t)) return 1;
    *t = -1.0;
    return 0;
}

/* ---- createEdges (141-231) ------------------------------------------------ */
typedef struct { int lv[3]; int face; int hash; int ord; } face_edge;
static int cmp_fe(const void *a, const void *b)
{
    const face_edge *x = (const face_edge *)a, *y = (const face_edge *)b;
    if (x->hash != y->hash) return x->hash < y->hash ? -1 : 1;
    return (x->ord > y->ord) - (x->ord < y->ord);  /* stable */
}

int orc_create_edges(int N, int F, const int *tri, const double *verts, orc_edge *edges)
{
    (void)N;
    int n = 3 * F;
    face_edge *tmp = (face_edge *)malloc(sizeof(face_edge) * (size_t)(n ? n : 1));
    for (int k = 0; k < F; ++k)
        for (int i = 0; i < 3; ++i) {
            face_edge *fe = &tmp[3 * k + i];
            fe->lv[0] = i % 3; fe->lv[1] = (i + 1) % 3; fe->lv[2] = (i + 2) % 3;
            fe->face = k;
            int a = tri[3 * k + fe->lv[0]], b = tri[3 * k + fe->lv[1]];
            int kmin = (a < b ? a : b) + 1, kmax = (a < b ? b : a) + 1;
            /* int arithmetic in the reference: wraps modulo 2^32 */
            fe->hash = (int)((uint32_t)kmin + (uint32_t)(n + 1) * (uint32_t)kmax);
            fe->ord = 3 * k + i;
        }
    qsort(tmp, (size_t)n, sizeof(face_edge), cmp_fe);
    int ne = 0, k = 0;
    while (k < n) {
        int f0 = tmp[k].face;
        orc_edge *E = &edges[ne++];
        E->verts[0] = tri[3 * f0 + tmp[k].lv[0]];
        E->verts[1] = tri[3 * f0 + tmp[k].lv[1]];
        E->verts[2] = tri[3 * f0 + tmp[k].lv[2]];
        double n0[3], d1[3], d2[3], c[3];
        const double *xa = verts + 3 * tri[3 * f0], *xb = verts + 3 * tri[3 * f0 + 1], *xc = verts + 3 * tri[3 * f0 + 2];
        v_sub(xb, xa, d1); v_sub(xc, xa, d2); v_cross(d1, d2, c); v_normalized(c, n0);
        v_copy(n0, E->normals);
        if (k < n - 1 && tmp[k].hash == tmp[k + 1].hash) {
            int f1 = tmp[k + 1].face;
            double n1[3];
            const double *ya = verts + 3 * tri[3 * f1], *yb = verts + 3 * tri[3 * f1 + 1], *yc = verts + 3 * tri[3 * f1 + 2];
            v_sub(yb, ya, d1); v_sub(yc, ya, d2); v_cross(d1, d2, c); v_normalized(c, n1);
            E->verts[3] = tri[3 * f1 + tmp[k + 1].lv[2]];
            E->faces[0] = f0; E->faces[1] = f1;
            E->internal = 1;
            E->angle = acos(v_dot(n0, n1));
            v_copy(n1, E->normals + 3);
            k += 2;
        } else {
            E->verts[3] = -1;
            E->faces[0] = f0; E->faces[1] = -1;
            E->internal = 0;
            E->angle = 1e9;
            E->normals[3] = E->normals[4] = E->normals[5] = 0.0;
            k += 1;
        }
    }
    free(tmp);
    return ne;
}

/* ---- shared pieces ---------------------------------------------------------- */
static double *jittered(int N, const double *verts, double thr)
{
    double *r = (double *)malloc(sizeof(double) * 3 * (size_t)(N ? N : 1));
    double *o = (double *)malloc(sizeof(double) * 3 * (size_t)(N ? N : 1));
    orc_mt19937_uniform(1u, 3 * N, -1.0, 1.0, r);
    for (int i = 0; i < 3 * N; ++i) o[i] = verts[i] + r[i] * thr * 1e-3;
    free(r);
    return o;
}

static double *face_normals(int F, const int *tri, const double *v)
{
    double *fn = (double *)malloc(sizeof(double) * 3 * (size_t)(F ? F : 1));
    for (int k = 0; k < F; ++k)
        face_normal(v + 3 * tri[3 * k], v + 3 * tri[3 * k + 1], v + 3 * tri[3 * k + 2], fn + 3 * k);
    return fn;
}

static void body_aabb(int N, const double *v, double *aabb)
{
    for (int i = 0; i < 3; ++i) {
        double lo = v[i], hi = v[i];
        for (int k = 1; k < N; ++k) {
            double x = v[3 * k + i];
            if (x < lo) lo = x;
            if (x > hi) hi = x;
        }
        aabb[i] = lo; aabb[i + 3] = hi;
    }
}

static void rec_init(orc_collision *c)
{
    memset(c, 0, sizeof(*c));
    c->tri1 = -1; c->tri2 = -1; c->edge2 = -1; c->nedge1 = 0;
}

/* vertex (x1, normal nor1) against all cloth faces: 770-845 and 1141-1213 */
static int vert_vs_faces(const double *x1, const double *nor1, double thr, int F, const int *tri,
                         const double *v2, const double *fn2, orc_collision *best)
{
    int found = 0;
    for (int j2 = 0; j2 < F; ++j2) {
        const int *f2 = tri + 3 * j2;
        const double *x2a = v2 + 3 * f2[0], *x2b = v2 + 3 * f2[1], *x2c = v2 + 3 * f2[2];
        double nor2[3];
        v_copy(fn2 + 3 * j2, nor2);
        if (v_dot(nor1, nor2) < 0.0) v_neg(nor2);
        double dx[3];
        v_sub(x1, x2a, dx);
        double proj = v_dot(dx, nor2);
        if (proj < 0.0) continue;
        double x2[3];
        for (int i = 0; i < 3; ++i) x2[i] = x1[i] - proj * nor2[i];
        v_sub(x2, x1, dx);
        double dist = v_norm(dx);
        if (dist > 5.0 * thr) continue;
        double u, v;
        barycentric(&u, &v, x2a, x2b, x2c, x1);
        double w = 1.0 - u - v;
        if (u < 0.0 || 1.0 < u || v < 0.0 || 1.0 < v || w < 0.0 || 1.0 < w) continue;
        if (!found || dist < best->dist) {
            rec_init(best);
            best->dist = dist;
            v_copy(nor1, best->nor1); v_copy(nor2, best->nor2);
            v_copy(x1, best->pos1); v_copy(x2, best->pos2);
            best->count1 = 1; best->count2 = 3;
            best->verts2[0] = f2[0]; best->verts2[1] = f2[1]; best->verts2[2] = f2[2];
            best->weights1[0] = 1.0;
            best->weights2[0] = u; best->weights2[1] = v; best->weights2[2] = w;
            best->tri2 = j2;
            found = 1;
        }
    }
    return found;
}

static void push(orc_collision *out, int cap, int *cnt, const orc_collision *c)
{
    if (*cnt < cap) out[*cnt] = *c;
    ++*cnt;
}

static void snap_pos1(orc_collision *out, int from, int to, int cap, double thr)
{
    double snap = 0.1 * thr;
    for (int k = from; k < to && k < cap; ++k)
        for (int i = 0; i < 3; ++i) out[k].pos1_[i] = out[k].pos1[i] - snap * out[k].nor1[i];
}

int orc_point_tri(double thr, int P, const double *pts, const double *nrm,
                  int N, const double *verts, int F, const int *tri, int eol,
                  orc_collision *out, int cap)
{
    int cnt = 0;
    double *v2 = jittered(N, verts, thr);
    double *fn2 = face_normals(F, tri, v2);
    double aabbB2[6];
    body_aabb(N, v2, aabbB2);
    if (eol) {
        for (int i2 = 0; i2 < N; ++i2) {
            const double *x2 = v2 + 3 * i2;
            orc_collision best;
            int found = 0;
            for (int i1 = 0; i1 < P; ++i1) {
                const double *x1 = pts + 3 * i1;
                double dx[3];
                v_sub(x2, x1, dx);
                double dist = v_norm(dx);
                if (dist < thr) {
                    if (!found || dist < best.dist) {
                        rec_init(&best);
                        best.dist = dist;
                        v_copy(nrm + 3 * i1, best.nor1); v_copy(nrm + 3 * i1, best.nor2);
                        v_copy(x1, best.pos1); v_copy(x2, best.pos2);
                        best.count1 = 3; best.count2 = 1;
                        best.verts1[0] = i1; best.verts1[1] = -1; best.verts1[2] = -1;
                        best.verts2[0] = i2; best.verts2[1] = -1; best.verts2[2] = -1;
                        best.weights1[0] = 1.0; best.weights2[0] = 1.0;
                        found = 1;
                    }
                }
            }
            if (found) push(out, cap, &cnt, &best);
        }
    }
    for (int i1 = 0; i1 < P; ++i1) {
        const double *x1 = pts + 3 * i1;
        double aabbV1[6] = { x1[0], x1[1], x1[2], x1[0], x1[1], x1[2] };
        if (!check_aabb(aabbV1, aabbB2)) continue;
        orc_collision best;
        if (vert_vs_faces(x1, nrm + 3 * i1, thr, F, tri, v2, fn2, &best)) {
            best.verts1[0] = i1; best.verts1[1] = -1; best.verts1[2] = -1;
            push(out, cap, &cnt, &best);
        }
    }
    snap_pos1(out, 0, cnt, cap, thr);
    free(v2); free(fn2);
    return cnt;
}

int orc_box_tri(double thr, const double *whd, const double *E1,
                int N, const double *verts, int F, const int *tri, const int *isEOL, int eol,
                orc_collision *out, int cap)
{
    (void)isEOL;
    int cnt = 0;
    /* edges from the un-jittered positions (8-argument overload, 602-615) */
    orc_edge *edges2 = (orc_edge *)malloc(sizeof(orc_edge) * (size_t)(3 * F + 1));
    int ne2 = orc_create_edges(N, F, tri, verts, edges2);
    box_geom B;
    make_box(&B, whd, E1);
    double *v2 = jittered(N, verts, thr);
    double *fn2 = face_normals(F, tri, v2);
    double aabbB2[6];
    body_aabb(N, v2, aabbB2);
    int status = 0;
    if (!eol) {
        /* Vertex2-Triangle1 (675-764) */
        for (int i2 = 0; i2 < N; ++i2) {
            const double *x2 = v2 + 3 * i2;
            double aabbV2[6] = { x2[0], x2[1], x2[2], x2[0], x2[1], x2[2] };
            if (!check_aabb(aabbV2, B.aabb)) continue;
            int inside = 1;
            for (int j1 = 0; j1 < 24; ++j1) {
                double dx[3];
                v_sub(x2, B.v[box_faces[j1][0]], dx);
                if (v_dot(B.fn[j1], dx) > 0.0) { inside = 0; break; }
            }
            if (!inside) continue;
            orc_collision best;
            int found = 0;
            for (int j1 = 0; j1 < 24; ++j1) {
                const double *x1a = B.v[box_faces[j1][0]], *x1b = B.v[box_faces[j1][1]], *x1c = B.v[box_faces[j1][2]];
                const double *nor1 = B.fn[j1];
                double dx[3], x1[3];
                v_sub(x2, x1a, dx);
                double proj = v_dot(nor1, dx);
                if (proj > 0.0) continue;
                for (int i = 0; i < 3; ++i) x1[i] = x2[i] - proj * nor1[i];
                v_sub(x2, x1, dx);
                double dist = v_norm(dx);
                if (dist > 5.0 * thr) continue;
                double u, v;
                barycentric(&u, &v, x1a, x1b, x1c, x1);
                double w = 1.0 - u - v;
                if (u < 0.0 || 1.0 < u || v < 0.0 || 1.0 < v || w < 0.0 || 1.0 < w) continue;
                if (i2 >= F) { status = -2; continue; }  /* faceNors2.col(i2) would read past the end */
                double nor2[3];
                v_copy(fn2 + 3 * i2, nor2);
                if (v_dot(nor2, nor1) < 0.0) v_neg(nor2);
                if (!found || dist < best.dist) {
                    rec_init(&best);
                    best.dist = dist;
                    v_copy(nor1, best.nor1); v_copy(nor2, best.nor2);
                    v_copy(x1, best.pos1); v_copy(x2, best.pos2);
                    best.count1 = 3; best.count2 = 1;
                    for (int i = 0; i < 3; ++i) best.verts1[i] = box_faces[j1][i];
                    best.verts2[0] = i2; best.verts2[1] = -1; best.verts2[2] = -1;
                    best.weights1[0] = u; best.weights1[1] = v; best.weights1[2] = w;
                    best.weights2[0] = 1.0;
                    best.tri1 = j1;
                    found = 1;
                }
            }
            if (found) push(out, cap, &cnt, &best);
        }
    }
    /* Vertex1-Triangle2 (770-845): the 8 corners */
    for (int i1 = 0; i1 < 8; ++i1) {
        const double *x1 = B.v[i1];
        double aabbV1[6] = { x1[0], x1[1], x1[2], x1[0], x1[1], x1[2] };
        if (!check_aabb(aabbV1, aabbB2)) continue;
        orc_collision best;
        if (vert_vs_faces(x1, B.vn[i1], thr, F, tri, v2, fn2, &best)) {
            best.verts1[0] = i1; best.verts1[1] = -1; best.verts1[2] = -1;
            best.nedge1 = 3;
            for (int i = 0; i < 3; ++i) best.edge1[i] = box_vert_edges[i1][i];
            push(out, cap, &cnt, &best);
        }
    }
    /* Edge2-Edge1 (848-1017) */
    const double threshAng = 2.0 * CD_PI / 180.0;
    for (int k2 = 0; k2 < ne2; ++k2) {
        const orc_edge *e2 = &edges2[k2];
        const double *x2a = v2 + 3 * e2->verts[0], *x2b = v2 + 3 * e2->verts[1];
        double dx2[3], nor2[3], t[3];
        v_sub(x2b, x2a, dx2);
        double len2 = v_norm(dx2);
        v_add(e2->normals, e2->normals + 3, t);
        v_normalized(t, nor2);
        double aabbE[6];
        for (int i = 0; i < 3; ++i) {
            double lo = x2a[i], hi = x2a[i];
            lo = (lo < x2b[i]) ? lo : x2b[i];  /* std::min(x1, aabb) */
            hi = (x2b[i] < hi) ? hi : x2b[i];  /* std::max(x1, aabb) */
            aabbE[i] = lo; aabbE[i + 3] = hi;
        }
        for (int k1 = 0; k1 < 12; ++k1) {
            if (B.eangle[k1] < CD_PI / 6.0) continue;
            if (!check_aabb(B.faabb[box_edge_faces[k1][0]], aabbE) &&
                !check_aabb(B.faabb[box_edge_faces[k1][1]], aabbE)) continue;
            const double *x1a = B.v[box_edge_verts[k1][0]], *x1b = B.v[box_edge_verts[k1][1]];
            double dx1[3], tan1[3];
            v_sub(x1b, x1a, dx1);
            double len1 = v_norm(dx1);
            for (int i = 0; i < 3; ++i) tan1[i] = dx1[i] / len1;
            double angle = acos(v_dot(tan1, nor2));
            if (fabs(angle) < threshAng || fabs(CD_PI - angle) < threshAng) continue;
            angle = acos(v_dot(tan1, dx2) / len2);
            if (fabs(angle) < threshAng || fabs(CD_PI - angle) < threshAng) continue;
            double nor[3], c[3];
            v_cross(dx1, dx2, c);
            v_normalized(c, nor);
            const double *x1c = B.v[box_edge_verts[k1][2]], *x1d = B.v[box_edge_verts[k1][3]];
            const double *n1c = B.en[k1][0], *n1d = B.en[k1][1];
            double nor1[3];
            v_add(n1c, n1d, nor1);
            v_normalized(nor1, nor1);
            if (v_dot(nor, nor1) < 0.0) v_neg(nor);
            double angleCD = acos(v_dot(n1c, n1d));
            double angleCN = acos(v_dot(n1c, nor));
            if (angleCD < 0.0) { angleCD = -angleCD; angleCN = -angleCN; }
            if (angleCN < -threshAng || angleCN - angleCD > threshAng) continue;
            double u2c = 0, u2d = 0;
            int i2c = intersect_square(x2a, dx2, x1a, x1b, x1c, &u2c);
            int i2d = intersect_square(x2a, dx2, x1b, x1a, x1d, &u2d);
            i2c = i2c && (0.0 <= u2c && u2c <= 1.0);
            i2d = i2d && (0.0 <= u2d && u2d <= 1.0);
            double u1, u2;
            lineline(&u1, &u2, x1a, x1b, x2a, x2b);
            double thresh1 = 1.0 * thr / len1, thresh2 = 1.0 * thr / len2;
            if (u1 < -thresh1 || u1 > 1.0 + thresh1 || u2 < -thresh2 || u2 > 1.0 + thresh2) continue;
            double x1[3], x2[3], dx[3];
            if (i2c && i2d) {
                for (int i = 0; i < 3; ++i) {
                    x1[i] = (1.0 - u1) * x1a[i] + u1 * x1b[i];
                    x2[i] = (1.0 - u2) * x2a[i] + u2 * x2b[i];
                }
            } else if (i2c && !i2d) {
                v_sub(x2a, x1a, dx);
                if (v_dot(dx, n1c) < 0.0) { double m = (u2 < u2c) ? u2 : u2c; u2 = (0.0 < m) ? m : 0.0; }
                else { double m = (u2 < 1.0) ? u2 : 1.0; u2 = (u2c < m) ? m : u2c; }
                for (int i = 0; i < 3; ++i) x2[i] = (1.0 - u2) * x2a[i] + u2 * x2b[i];
                u1 = linepoint(x1a, x1b, x2);
                for (int i = 0; i < 3; ++i) x1[i] = (1.0 - u1) * x1a[i] + u1 * x1b[i];
            } else if (!i2c && i2d) {
                v_sub(x2a, x1b, dx);
                if (v_dot(dx, n1d) < 0.0) { double m = (u2 < u2d) ? u2 : u2d; u2 = (0.0 < m) ? m : 0.0; }
                else { double m = (u2 < 1.0) ? u2 : 1.0; u2 = (u2d < m) ? m : u2d; }
                for (int i = 0; i < 3; ++i) x2[i] = (1.0 - u2) * x2a[i] + u2 * x2b[i];
                u1 = linepoint(x1a, x1b, x2);
                for (int i = 0; i < 3; ++i) x1[i] = (1.0 - u1) * x1a[i] + u1 * x1b[i];
            } else continue;
            if (u1 < -thresh1 || u1 > 1.0 + thresh1 || u2 < -thresh2 || u2 > 1.0 + thresh2) continue;
            v_sub(x2, x1, dx);
            double th = 2.0 * thr;
            if (v_dot(dx, dx) > th * th) continue;
            orc_collision cc;
            rec_init(&cc);
            cc.dist = v_norm(dx);
            v_copy(nor1, cc.nor1); v_copy(nor, cc.nor2);
            v_copy(x1, cc.pos1); v_copy(x2, cc.pos2);
            cc.count1 = 2; cc.count2 = 2;
            cc.verts1[0] = box_edge_verts[k1][0]; cc.verts1[1] = box_edge_verts[k1][1]; cc.verts1[2] = -1;
            cc.verts2[0] = e2->verts[0]; cc.verts2[1] = e2->verts[1]; cc.verts2[2] = -1;
            cc.weights1[0] = 1.0 - u1; cc.weights1[1] = u1;
            cc.weights2[0] = 1.0 - u2; cc.weights2[1] = u2;
            cc.nedge1 = 1; cc.edge1[0] = k1;
            cc.edge2 = k2;
            v_copy(tan1, cc.edgeDir);
            push(out, cap, &cnt, &cc);
        }
    }
    /* corner de-duplication (1020-1052): keep the closest record per corner */
    if (cnt <= cap) {
        for (int i1 = 0; i1 < 8; ++i1) {
            int kmin = -1;
            double dmin = 1e9;
            for (int k = 0; k < cnt; ++k)
                if (out[k].count1 == 1 && out[k].verts1[0] == i1) {
                    double dx[3];
                    v_sub(out[k].pos2, B.v[i1], dx);
                    double d = v_dot(dx, dx);
                    if (d < dmin) { kmin = k; dmin = d; }
                }
            if (kmin != -1) {
                int *dl = (int *)malloc(sizeof(int) * (size_t)(cnt + 1));
                int nd = 0;
                for (int k = 0; k < cnt; ++k)
                    if (out[k].count1 == 1 && out[k].verts1[0] == i1 && k != kmin) dl[nd++] = k;
                for (int q = 0; q < nd; ++q) { out[dl[q]] = out[cnt - 1]; --cnt; }
                free(dl);
            }
        }
    }
    snap_pos1(out, 0, cnt, cap, thr);
    free(edges2); free(v2); free(fn2);
    return status ? status : cnt;
}
