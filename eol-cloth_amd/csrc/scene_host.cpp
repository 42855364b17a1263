// Obstacle motion of the scene step (host-only; no device work): the rigid-body
// update of Box::step (Box.cpp:103-114) on Rigid::integrate (Rigid.cpp:79-112).
// Frames are 4x4 homogeneous transforms stored column-major, as Eigen stores
// Matrix4d.  A handful of boxes per step: host arithmetic, not worth a launch.
#include <cmath>

#include "eolc_internal.h"

namespace {

inline double &at(double *E, int r, int c) { return E[4 * c + r]; }
inline double at(const double *E, int r, int c) { return E[4 * c + r]; }

// E1 = E0 * exp(h [w]^, h v; 0, 0), phi = (w, v) in the body frame.  For |w| above
// 1e-10 the exponential is the closed form for a unit axis: rotation by
// theta = |w| h about w/|w| (Rodrigues) and translation (I - R)(u x s) + (u.s) theta u
// with u = w/|w|, s = v/|w|.  Otherwise a pure translation h v.
void integrate(const double *E0, const double *phi, double h, double *E1)
{
    double T[16] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1};
    double w[3] = {phi[0], phi[1], phi[2]}, v[3] = {phi[3], phi[4], phi[5]};
    for (int k = 0; k < 3; ++k) at(T, k, 3) = h * v[k];
    const double wlen = std::sqrt(w[0] * w[0] + w[1] * w[1] + w[2] * w[2]);
    if (wlen > 1e-10) {
        for (int k = 0; k < 3; ++k) {
            w[k] /= wlen;
            v[k] /= wlen;
        }
        const double th = wlen * h, c = std::cos(th), s = std::sin(th), c1 = 1.0 - c;
        const double R[3][3] = {{c + w[0] * w[0] * c1, -w[2] * s + w[0] * w[1] * c1, w[1] * s + w[0] * w[2] * c1},
                                {w[2] * s + w[0] * w[1] * c1, c + w[1] * w[1] * c1, -w[0] * s + w[1] * w[2] * c1},
                                {-w[1] * s + w[0] * w[2] * c1, w[0] * s + w[1] * w[2] * c1, c + w[2] * w[2] * c1}};
        const double cr[3] = {w[1] * v[2] - w[2] * v[1], w[2] * v[0] - w[0] * v[2], w[0] * v[1] - w[1] * v[0]};
        const double wv = w[0] * v[0] + w[1] * v[1] + w[2] * v[2];
        // (I - R) is formed before it multiplies u x s: near the small-angle branch
        // u x s is large and (I - R)(u x s) would cancel catastrophically if taken
        // as (u x s) - R (u x s)
        for (int r = 0; r < 3; ++r) {
            double d = 0.0;
            for (int k = 0; k < 3; ++k) d += ((r == k ? 1.0 : 0.0) - R[r][k]) * cr[k];
            for (int k = 0; k < 3; ++k) at(T, r, k) = R[r][k];
            at(T, r, 3) = wv * th * w[r] + d;
        }
    }
    double out[16];
    for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) {
            double acc = 0.0;
            for (int k = 0; k < 4; ++k) acc += at(E0, r, k) * at(T, k, c);
            at(out, r, c) = acc;
        }
    for (int k = 0; k < 16; ++k) E1[k] = out[k];
}

bool finite(const double *a, int n)
{
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(a[i])) return false;
    return true;
}

}  // namespace

extern "C" int eolc_rigid_integrate(const double *E0, const double *phi, double h, double *E1)
{
    if (!E0 || !phi || !E1 || !finite(E0, 16) || !finite(phi, 6) || !std::isfinite(h)) return EOLC_ERR_INVALID;
    integrate(E0, phi, h, E1);
    return EOLC_OK;
}

extern "C" int eolc_box_step(const double *E, const double *vel, double h, double *E_new)
{
    if (!E || !vel || !E_new || !finite(E, 16) || !finite(vel, 6) || !std::isfinite(h)) return EOLC_ERR_INVALID;
    // body-frame twist: the adjoint blockdiag(R^T, R^T) applied to (w, v)
    double phi[6];
    for (int r = 0; r < 3; ++r) {
        phi[r] = phi[3 + r] = 0.0;
        for (int k = 0; k < 3; ++k) {
            phi[r] += at(E, k, r) * vel[k];
            phi[3 + r] += at(E, k, r) * vel[3 + k];
        }
    }
    integrate(E, phi, h, E_new);
    return EOLC_OK;
}
