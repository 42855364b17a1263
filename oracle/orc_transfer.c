/* Velocity transfer after a remesh, non-EOL nodes: CPU restatement (test
 * infrastructure only: the checker of the GPU transfer kernels).
 *
 * Follows Cloth::velocityTransfer (Cloth.cpp:195-230 and 265-281) with brute-force
 * loops exactly as written there:
 *   - a new vertex takes the velocity of the FIRST old vertex (index order) whose
 *     material-space distance, unsigned_vv_distance (geometry.cpp:228-230), is below
 *     1e-6 (the loop breaks at the first hit);
 *   - otherwise it takes the barycentric combination of the velocities of the FIRST
 *     old face (index order) for which is_inside holds (get_enclosing_face,
 *     geometry.cpp:284-290; is_inside, geometry.cpp:276-281, thresholds -10 EPSILON,
 *     -10 EPSILON, -100 EPSILON with EPSILON = 1e-7f, util.hpp:38, evaluated in float
 *     as the C expression is), barycentric coordinates from get_barycentric_coords
 *     (geometry.cpp:253-271), v = b0 vA + b1 vB + b2 vC;
 *   - a vertex with neither reports face -1 and velocity 0 (the reference would
 *     dereference a null face).
 * pow(d, 2) of unsigned_vv_distance is written d * d. */
#include <math.h>

#include "oracle.h"

static void bary_coords(const double *p, const double *a, const double *b, const double *c, double out[3])
{
    const double v0x = a[0] - c[0], v0y = a[1] - c[1];
    const double v1x = b[0] - c[0], v1y = b[1] - c[1];
    const double v2x = p[0] - c[0], v2y = p[1] - c[1];
    const double dot00 = v0x * v0x + v0y * v0y;
    const double dot01 = v0x * v1x + v0y * v1y;
    const double dot02 = v0x * v2x + v0y * v2y;
    const double dot11 = v1x * v1x + v1y * v1y;
    const double dot12 = v1x * v2x + v1y * v2y;
    const double invDenom = 1.0 / (dot00 * dot11 - dot01 * dot01);
    const double u = (dot11 * dot02 - dot01 * dot12) * invDenom;
    const double v = (dot00 * dot12 - dot01 * dot02) * invDenom;
    out[0] = u;
    out[1] = v;
    out[2] = 1 - u - v;
}

int orc_velocity_transfer(int n_old, const double *u_old, const double *v_old, int f_old, const int *tri_old,
                          int n_new, const double *u_new, double *v_new, int *src, int *face, double *bary)
{
    const float eps = 1e-7f;
    const double t01 = (double)(-10 * eps), t2 = (double)(-100 * eps);
    int missing = 0;
    for (int n = 0; n < n_new; ++n) {
        const double *p = u_new + 2 * n;
        int found = -1;
        for (int j = 0; j < n_old; ++j) {
            const double dx = u_old[2 * j] - p[0], dy = u_old[2 * j + 1] - p[1];
            if (sqrt(dx * dx + dy * dy) < 1e-6) {
                found = j;
                break;
            }
        }
        src[n] = found;
        face[n] = -1;
        bary[3 * n] = bary[3 * n + 1] = bary[3 * n + 2] = 0.0;
        if (found >= 0) {
            for (int c = 0; c < 3; ++c) v_new[3 * n + c] = v_old[3 * found + c];
            continue;
        }
        int fi = -1;
        double b[3] = {0, 0, 0};
        for (int f = 0; f < f_old; ++f) {
            const int *t = tri_old + 3 * f;
            bary_coords(p, u_old + 2 * t[0], u_old + 2 * t[1], u_old + 2 * t[2], b);
            if (b[0] >= t01 && b[1] >= t01 && b[2] >= t2) {
                fi = f;
                break;
            }
        }
        face[n] = fi;
        if (fi < 0) {
            ++missing;
            for (int c = 0; c < 3; ++c) v_new[3 * n + c] = 0.0;
            continue;
        }
        const int *t = tri_old + 3 * fi;
        for (int c = 0; c < 3; ++c) {
            bary[3 * n + c] = b[c];
            v_new[3 * n + c] = b[0] * v_old[3 * t[0] + c] + b[1] * v_old[3 * t[1] + c] + b[2] * v_old[3 * t[2] + c];
        }
    }
    return missing;
}

/* The EOL correction of a new LAG vertex (Cloth.cpp:254-262): every vertex j of its
 * old face that was an EOL node (eidx[j] >= 0) contributes vw = v_j + (-F) V_j,
 * F = deform_grad(old face) (UtilEOL.cpp:13-28: Dx DX^-1 at the old world positions,
 * DX^-1 by its adjugate over the determinant), and v = b0 vwA + b1 vwB + b2 vwC.
 * Vertices matched to an old vertex keep that vertex's v. */
int orc_velocity_transfer_eol(int n_old, const double *u_old, const double *x_old, const double *v_old,
                              const int *eidx, const double *V_old, int f_old, const int *tri_old,
                              int n_new, const double *u_new, double *v_new, int *src, int *face, double *bary)
{
    const int missing = orc_velocity_transfer(n_old, u_old, v_old, f_old, tri_old, n_new, u_new, v_new, src, face, bary);
    for (int n = 0; n < n_new; ++n) {
        if (face[n] < 0) continue;
        const int *t = tri_old + 3 * face[n];
        if (eidx[t[0]] < 0 && eidx[t[1]] < 0 && eidx[t[2]] < 0) continue;
        const int a = t[0], b = t[1], c = t[2];
        const double d00 = u_old[2 * b] - u_old[2 * a], d01 = u_old[2 * c] - u_old[2 * a];
        const double d10 = u_old[2 * b + 1] - u_old[2 * a + 1], d11 = u_old[2 * c + 1] - u_old[2 * a + 1];
        const double det = d00 * d11 - d01 * d10;
        const double i00 = d11 / det, i01 = -d01 / det, i10 = -d10 / det, i11 = d00 / det;
        double F[3][2], vw[3][3];
        for (int r = 0; r < 3; ++r) {
            const double e0 = x_old[3 * b + r] - x_old[3 * a + r], e1 = x_old[3 * c + r] - x_old[3 * a + r];
            F[r][0] = e0 * i00 + e1 * i10;
            F[r][1] = e0 * i01 + e1 * i11;
        }
        for (int k = 0; k < 3; ++k) {
            const int j = t[k];
            for (int r = 0; r < 3; ++r) vw[k][r] = v_old[3 * j + r];
            if (eidx[j] >= 0) {
                const double V0 = V_old[2 * eidx[j]], V1 = V_old[2 * eidx[j] + 1];
                for (int r = 0; r < 3; ++r) vw[k][r] = vw[k][r] + ((-F[r][0]) * V0 + (-F[r][1]) * V1);
            }
        }
        const double b0 = bary[3 * n], b1 = bary[3 * n + 1], b2 = bary[3 * n + 2];
        for (int r = 0; r < 3; ++r) v_new[3 * n + r] = b0 * vw[0][r] + b1 * vw[1][r] + b2 * vw[2][r];
    }
    return missing;
}

