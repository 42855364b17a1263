// C entry points onto the reference's element kernels (ComputeMembrane.h:5-19,
// ComputeInertial.h:5-16, ComputeBending.h:5-18) and ray/triangle test
// (raytri.cpp:259-262), so ctypes can call them.  Test infrastructure only.
#include "ComputeMembrane.h"
#include "ComputeInertial.h"
#include "ComputeBending.h"

int intersect_triangle3_inc(const double *orig, const double *dir,
                            const double *vert0, const double *vert1, const double *vert2,
                            double *t, double *u, double *v);

extern "C" {
void ref_membrane(const double *xa, const double *xb, const double *xc, const double *Xa,
                  const double *Xb, const double *Xc, double e, double nu, const double *P,
                  const double *Q, double *W, double *f, double *K)
{
    ComputeMembrane(xa, xb, xc, Xa, Xb, Xc, e, nu, P, Q, W, f, K);
}
void ref_inertial(const double *xa, const double *xb, const double *xc, const double *Xa,
                  const double *Xb, const double *Xc, const double *g, double rho, double *W,
                  double *f, double *M)
{
    ComputeInertial(xa, xb, xc, Xa, Xb, Xc, g, rho, W, f, M);
}
void ref_bending(const double *x0, const double *x1, const double *x2, const double *x3,
                 const double *X0, const double *X1, const double *X2, const double *X3,
                 double beta, double *W, double *f, double *K)
{
    ComputeBending(x0, x1, x2, x3, X0, X1, X2, X3, beta, W, f, K);
}
int ref_tri3_inc(const double *orig, const double *dir, const double *v0, const double *v1,
                 const double *v2, double *t, double *u, double *v)
{
    return intersect_triangle3_inc(orig, dir, v0, v1, v2, t, u, v);
}
}
