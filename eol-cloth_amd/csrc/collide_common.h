// Shared host/device pieces of the proximity queries.  Compiled with
// -ffp-contract=off on both sides; expressions follow the evaluation order of the
// reference's Eigen 3-vector code so host and device results agree bit for bit.
//
// References: boxTriCollision.cpp -- box tables 289-392, createFaceNormals 233-247,
// check_AABB 462-486, barycentric 488-505, lineline 507-525, linepoint 527-536,
// intersect_square 550-600, E2-E1 loop 848-1017; raytri.cpp:259-316
// (intersect_triangle3_inc, whose "-ZERO" expands to +1e-6).
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>

#include "../../include/eolc.h"

namespace eolc_cd {

#define CDQ __host__ __device__ __forceinline__

constexpr double kPi = 3.14159265358979323846;

// ---- box tables (boxTriCollision.cpp:289-392) --------------------------------
__host__ __device__ constexpr int kBoxFaces[24][3] = {
    {0, 8, 2},  {1, 8, 0},  {3, 8, 1},  {2, 8, 3},  {4, 10, 0}, {5, 10, 4}, {1, 10, 5}, {0, 10, 1},
    {6, 9, 4},  {7, 9, 6},  {5, 9, 7},  {4, 9, 5},  {2, 11, 6}, {3, 11, 2}, {7, 11, 3}, {6, 11, 7},
    {1, 13, 3}, {5, 13, 1}, {7, 13, 5}, {3, 13, 7}, {0, 12, 4}, {2, 12, 0}, {6, 12, 2}, {4, 12, 6}};
__host__ __device__ constexpr int kBoxEdgeVerts[12][4] = {
    {0, 1, 8, 10}, {2, 0, 8, 12}, {1, 3, 8, 13}, {3, 2, 8, 11}, {0, 4, 10, 12}, {5, 1, 10, 13},
    {4, 5, 10, 9}, {6, 2, 11, 12}, {4, 6, 9, 12}, {3, 7, 11, 13}, {7, 5, 9, 13}, {6, 7, 9, 11}};
__host__ __device__ constexpr int kBoxVertEdges[8][3] = {{0, 1, 4},  {0, 2, 5},  {1, 3, 7},  {2, 3, 9},
                                                         {4, 6, 8}, {5, 6, 10}, {7, 8, 11}, {9, 10, 11}};
__host__ __device__ constexpr int kBoxEdgeFaces[12][2] = {{1, 7},   {0, 21}, {2, 16},  {3, 13}, {4, 20},  {6, 17},
                                                          {5, 11}, {12, 22}, {8, 23}, {14, 19}, {10, 18}, {9, 15}};
__host__ __device__ constexpr double kBoxUnitVerts[14][3] = {
    {-1, -1, -1}, {-1, -1, 1}, {-1, 1, -1}, {-1, 1, 1}, {1, -1, -1}, {1, -1, 1}, {1, 1, -1},
    {1, 1, 1},    {-1, 0, 0},  {1, 0, 0},   {0, -1, 0}, {0, 1, 0},   {0, 0, -1}, {0, 0, 1}};

struct BoxGeom {
    double v[14][3];
    double fn[24][3];
    double vn[14][3];
    double en[12][2][3];
    double eangle[12];
    double faabb[24][6];
    double aabb[6];
};

struct CdGrid {
    double org[3];
    double inv_cell;
    int dim[3];
    unsigned mask;
    double margin;
    double cond_max;
};

// ---- 3-vector arithmetic in Eigen's order -------------------------------------
CDQ void v_sub(const double *a, const double *b, double *o) { o[0] = a[0] - b[0]; o[1] = a[1] - b[1]; o[2] = a[2] - b[2]; }
CDQ void v_add(const double *a, const double *b, double *o) { o[0] = a[0] + b[0]; o[1] = a[1] + b[1]; o[2] = a[2] + b[2]; }
CDQ double v_dot(const double *a, const double *b) { return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]; }
CDQ void v_cross(const double *a, const double *b, double *o)
{
    double c0 = a[1] * b[2] - a[2] * b[1], c1 = a[2] * b[0] - a[0] * b[2], c2 = a[0] * b[1] - a[1] * b[0];
    o[0] = c0; o[1] = c1; o[2] = c2;
}
CDQ double v_norm(const double *a) { return sqrt(v_dot(a, a)); }
CDQ void v_normalized(const double *a, double *o)
{
    double z = v_dot(a, a);
    if (z > 0.0) {
        double s = sqrt(z);
        o[0] = a[0] / s; o[1] = a[1] / s; o[2] = a[2] / s;
    } else {
        o[0] = a[0]; o[1] = a[1]; o[2] = a[2];
    }
}
CDQ void v_copy(const double *a, double *o) { o[0] = a[0]; o[1] = a[1]; o[2] = a[2]; }
CDQ void v_neg(double *a) { a[0] = -a[0]; a[1] = -a[1]; a[2] = -a[2]; }

// createFaceNormals: dba.cross(-dac).normalized()
CDQ void face_normal(const double *xa, const double *xb, const double *xc, double *n)
{
    double dba[3], dac[3], c[3];
    v_sub(xb, xa, dba);
    v_sub(xa, xc, dac);
    v_neg(dac);
    v_cross(dba, dac, c);
    v_normalized(c, n);
}

// check_AABB: both boxes grown by a fixed 1e-3
CDQ bool check_aabb(const double *a1, const double *a2)
{
    const double t = 1e-3;
    double mn1[3], mx1[3], mn2[3], mx2[3];
    for (int i = 0; i < 3; ++i) { mn1[i] = a1[i] - t; mx1[i] = a1[i + 3] + t; mn2[i] = a2[i] - t; mx2[i] = a2[i + 3] + t; }
    return mx1[0] >= mn2[0] && mn1[0] <= mx2[0] && mx1[1] >= mn2[1] && mn1[1] <= mx2[1] && mx1[2] >= mn2[2] &&
           mn1[2] <= mx2[2];
}

CDQ void barycentric(double *alpha, double *beta, const double *a, const double *b, const double *c, const double *p)
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

CDQ void lineline(double *a, double *b, const double *A1, const double *A2, const double *B1, const double *B2)
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

CDQ double linepoint(const double *A, const double *B, const double *P)
{
    double AP[3], AB[3];
    v_sub(P, A, AP); v_sub(B, A, AB);
    return v_dot(AP, AB) / v_dot(AB, AB);
}

CDQ int tri3_inc(const double *orig, const double *dir, const double *v0, const double *v1, const double *v2, double *t)
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
    } else {
        return 0;
    }
    *t = v_dot(e2, qv) * inv_det;
    return 1;
}

CDQ int intersect_square(const double *x0, const double *dx, const double *xa, const double *xb, const double *xc,
                         double *t)
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

CDQ void rec_init(eolc_collision *c)
{
    c->dist = 0.0;
    for (int i = 0; i < 3; ++i) {
        c->nor1[i] = c->nor2[i] = c->pos1[i] = c->pos2[i] = c->pos1_[i] = 0.0;
        c->verts1[i] = c->verts2[i] = 0;
        c->weights1[i] = c->weights2[i] = 0.0;
        c->edge1[i] = 0;
        c->edgeDir[i] = 0.0;
    }
    c->count1 = c->count2 = 0;
    c->tri1 = c->tri2 = -1;
    c->nedge1 = 0;
    c->edge2 = -1;
}

// E2-E1 test of one cloth edge against the 12 box edges (boxTriCollision.cpp:848-1017).
// cedges: (v0, v1, face0, face1 or -1) per edge in createEdges order; x the
// un-jittered positions (edge normals), xj the jittered ones.  Writes the records
// to out (when not null), returns their count; bit k1 of *ambig is set when an
// angle comparison for box edge k1 fell within `band` of its threshold.
CDQ int ee_edge(int k2, double thr, const BoxGeom &box, const int *cedges, const double *x, const double *xj,
                const int *tri, eolc_collision *out, int *ambig, double band)
{
    const int *e2 = cedges + 4 * k2;
    const double *x2a = xj + 3 * e2[0], *x2b = xj + 3 * e2[1];
    double dx2[3], nor2[3], t[3], n0[3], n1[3] = {0.0, 0.0, 0.0};
    v_sub(x2b, x2a, dx2);
    const double len2 = v_norm(dx2);
    {
        const int f0 = e2[2];
        double d1[3], d2[3], c[3];
        v_sub(x + 3 * tri[3 * f0 + 1], x + 3 * tri[3 * f0], d1);
        v_sub(x + 3 * tri[3 * f0 + 2], x + 3 * tri[3 * f0], d2);
        v_cross(d1, d2, c);
        v_normalized(c, n0);
        const int f1 = e2[3];
        if (f1 >= 0) {
            v_sub(x + 3 * tri[3 * f1 + 1], x + 3 * tri[3 * f1], d1);
            v_sub(x + 3 * tri[3 * f1 + 2], x + 3 * tri[3 * f1], d2);
            v_cross(d1, d2, c);
            v_normalized(c, n1);
        }
    }
    v_add(n0, n1, t);
    v_normalized(t, nor2);
    double aabbE[6];
    for (int i = 0; i < 3; ++i) {
        double lo = x2a[i], hi = x2a[i];
        lo = (lo < x2b[i]) ? lo : x2b[i];
        hi = (x2b[i] < hi) ? hi : x2b[i];
        aabbE[i] = lo; aabbE[i + 3] = hi;
    }
    const double threshAng = 2.0 * kPi / 180.0;
    int cnt = 0;
    for (int k1 = 0; k1 < 12; ++k1) {
        if (box.eangle[k1] < kPi / 6.0) continue;
        if (!check_aabb(box.faabb[kBoxEdgeFaces[k1][0]], aabbE) && !check_aabb(box.faabb[kBoxEdgeFaces[k1][1]], aabbE))
            continue;
        const double *x1a = box.v[kBoxEdgeVerts[k1][0]], *x1b = box.v[kBoxEdgeVerts[k1][1]];
        double dx1[3], tan1[3];
        v_sub(x1b, x1a, dx1);
        const double len1 = v_norm(dx1);
        for (int i = 0; i < 3; ++i) tan1[i] = dx1[i] / len1;
        double angle = acos(v_dot(tan1, nor2));
        if (fabs(fabs(angle) - threshAng) < band || fabs(fabs(kPi - angle) - threshAng) < band) *ambig |= 1 << k1;
        if (fabs(angle) < threshAng || fabs(kPi - angle) < threshAng) continue;
        angle = acos(v_dot(tan1, dx2) / len2);
        if (fabs(fabs(angle) - threshAng) < band || fabs(fabs(kPi - angle) - threshAng) < band) *ambig |= 1 << k1;
        if (fabs(angle) < threshAng || fabs(kPi - angle) < threshAng) continue;
        double nor[3], c[3];
        v_cross(dx1, dx2, c);
        v_normalized(c, nor);
        const double *x1c = box.v[kBoxEdgeVerts[k1][2]], *x1d = box.v[kBoxEdgeVerts[k1][3]];
        const double *n1c = box.en[k1][0], *n1d = box.en[k1][1];
        double nor1[3];
        v_add(n1c, n1d, nor1);
        v_normalized(nor1, nor1);
        if (v_dot(nor, nor1) < 0.0) v_neg(nor);
        double angleCD = acos(v_dot(n1c, n1d));
        double angleCN = acos(v_dot(n1c, nor));
        if (angleCD < 0.0) { angleCD = -angleCD; angleCN = -angleCN; }
        if (fabs(angleCN + threshAng) < band || fabs(angleCN - angleCD - threshAng) < band) *ambig |= 1 << k1;
        if (angleCN < -threshAng || angleCN - angleCD > threshAng) continue;
        double u2c = 0, u2d = 0;
        int i2c = intersect_square(x2a, dx2, x1a, x1b, x1c, &u2c);
        int i2d = intersect_square(x2a, dx2, x1b, x1a, x1d, &u2d);
        i2c = i2c && (0.0 <= u2c && u2c <= 1.0);
        i2d = i2d && (0.0 <= u2d && u2d <= 1.0);
        double u1, u2;
        lineline(&u1, &u2, x1a, x1b, x2a, x2b);
        const double thresh1 = 1.0 * thr / len1, thresh2 = 1.0 * thr / len2;
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
        } else {
            continue;
        }
        if (u1 < -thresh1 || u1 > 1.0 + thresh1 || u2 < -thresh2 || u2 > 1.0 + thresh2) continue;
        v_sub(x2, x1, dx);
        const double th = 2.0 * thr;
        if (v_dot(dx, dx) > th * th) continue;
        if (out) {
            eolc_collision cc;
            rec_init(&cc);
            cc.dist = v_norm(dx);
            v_copy(nor1, cc.nor1); v_copy(nor, cc.nor2);
            v_copy(x1, cc.pos1); v_copy(x2, cc.pos2);
            cc.count1 = 2; cc.count2 = 2;
            cc.verts1[0] = kBoxEdgeVerts[k1][0]; cc.verts1[1] = kBoxEdgeVerts[k1][1]; cc.verts1[2] = -1;
            cc.verts2[0] = e2[0]; cc.verts2[1] = e2[1]; cc.verts2[2] = -1;
            cc.weights1[0] = 1.0 - u1; cc.weights1[1] = u1;
            cc.weights2[0] = 1.0 - u2; cc.weights2[1] = u2;
            cc.nedge1 = 1; cc.edge1[0] = k1;
            cc.edge2 = k2;
            v_copy(tan1, cc.edgeDir);
            out[cnt] = cc;
        }
        ++cnt;
    }
    return cnt;
}

#undef CDQ
}  // namespace eolc_cd
