// EOL rows of Forces::fill (gfx950, FP64): the two material-space DOFs of an EOL
// node (Forces.cpp:177-329 fillEOLInertia / fillEOLMembrane, 580-683
// fillEOLBending, assembled by faceBasedF 399-497 and edgeBasedF 746-884).
//
// For an element with an EOL vertex, F is the deformation gradient of the face
// (deform_grad, UtilEOL.cpp:13-28: Dx DX^-1) or, for a hinge, the mean over its two
// faces.  With B_vw the element's 3x3 block (faces: m_vw I + dampingB h^2 K_vw from
// the staging buffers of the element kernels; hinges: dampingB h^2 K_vw) the EOL
// blocks are
//   (X_v, X_w) = F^T B_vw F,   (X_v, x_u) = -F^T B_vu  (u >= v),  -(B_uv F)^T (u < v),
//   f_X(v) = -F^T f_v,
// the values fillXMI ... fillxXB push (with their transposes).  An EOL hinge also
// adds its bending force to the x rows, which the non-EOL branch never does.
//   k_eol_elem   one thread per EOL element: F and the staged values
//   k_eol_gather one thread per X-row entry: contributions summed in element order
//   k_eol_fx     one thread per X DOF: the X forces, in element order
//   k_eol_xforce one thread per vertex of an EOL hinge: its bending forces added to
//                f after the face forces, in hinge order
#include "eolc_internal.h"

namespace {

__device__ __forceinline__ void deform_grad(const double *__restrict__ x, const double *__restrict__ X, int a, int b,
                                            int c, double F[3][2])
{
    double Dx[3][2];
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        Dx[r][0] = x[3 * b + r] - x[3 * a + r];
        Dx[r][1] = x[3 * c + r] - x[3 * a + r];
    }
    const double d00 = X[2 * b] - X[2 * a], d01 = X[2 * c] - X[2 * a];
    const double d10 = X[2 * b + 1] - X[2 * a + 1], d11 = X[2 * c + 1] - X[2 * a + 1];
    const double det = d00 * d11 - d01 * d10;
    const double i00 = d11 / det, i01 = -d01 / det, i10 = -d10 / det, i11 = d00 / det;
#pragma unroll
    for (int r = 0; r < 3; ++r) {
        F[r][0] = Dx[r][0] * i00 + Dx[r][1] * i10;
        F[r][1] = Dx[r][0] * i01 + Dx[r][1] * i11;
    }
}

// local pair (v <= w) -> slot offset within the element (fillxMI/fillxxMI and
// fillxB/fillxxB order)
__device__ __forceinline__ int pair_index(int nv, int v, int w)
{
    if (v == w) return v;
    if (nv == 3) return v == 0 ? 2 + w : 5;   // (0,1) 3, (0,2) 4, (1,2) 5
    // hinges: (0,1) 4, (0,2) 5, (0,3) 6, (1,2) 7, (1,3) 8, (2,3) 9
    return v == 0 ? 3 + w : (v == 1 ? 5 + w : 9);
}

// block (v, w) of the element, entry (a, b); mass scalar added on the diagonal of
// face blocks
__device__ __forceinline__ double blk(const double *__restrict__ K, const double *__restrict__ m, int nv, int v, int w,
                                      int a, int b)
{
    const int p = v <= w ? pair_index(nv, v, w) : pair_index(nv, w, v);
    const double k = v <= w ? K[9 * p + 3 * a + b] : K[9 * p + 3 * b + a];
    return (m && a == b) ? m[p] + k : k;
}
__device__ __forceinline__ double mblk(const double *__restrict__ m, int nv, int v, int w, int a, int b)
{
    if (!m || a != b) return 0.0;
    return m[v <= w ? pair_index(nv, v, w) : pair_index(nv, w, v)];
}

__global__ void __launch_bounds__(256)
k_eol_elem(int ne, int F, const int *__restrict__ elem, const int *__restrict__ tri, const int *__restrict__ hinge,
           const int *__restrict__ eidx, const double *__restrict__ x, const double *__restrict__ X,
           const double *__restrict__ stageK, const double *__restrict__ stageM, const double *__restrict__ stageF,
           const int64_t *__restrict__ soff, const int64_t *__restrict__ foff, double *__restrict__ stK,
           double *__restrict__ stM, double *__restrict__ stF)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const int id = elem[e];
    const bool face = id < F;
    const int nv = face ? 3 : 4;
    int vt[4];
    const double *K, *m = nullptr, *fv;
    double Fg[3][2];
    if (face) {
        for (int k = 0; k < 3; ++k) vt[k] = tri[3 * id + k];
        vt[3] = -1;
        K = stageK + (size_t)9 * 6 * id;
        m = stageM + (size_t)6 * id;
        fv = stageF + (size_t)9 * id;
        deform_grad(x, X, vt[0], vt[1], vt[2], Fg);
    } else {
        const int h = id - F;
        for (int k = 0; k < 4; ++k) vt[k] = hinge[4 * h + k];
        K = stageK + (size_t)9 * (6 * (size_t)F + 10 * (size_t)h);
        fv = stageF + (size_t)9 * F + (size_t)12 * h;
        double F1[3][2], F2[3][2];
        deform_grad(x, X, vt[0], vt[1], vt[2], F1);
        deform_grad(x, X, vt[1], vt[0], vt[3], F2);
        for (int r = 0; r < 3; ++r)
            for (int k = 0; k < 2; ++k) Fg[r][k] = (F1[r][k] + F2[r][k]) / 2;
    }
    int loc[4], nE = 0;
    for (int v = 0; v < nv; ++v)
        if (eidx[vt[v]] >= 0) loc[nE++] = v;
    const int rec = 4 * nE + 6 * nv;
    double *sk = stK + soff[e], *sm = stM + soff[e], *sf = stF + foff[e];
    for (int ev = 0; ev < nE; ++ev) {
        const int v = loc[ev];
        double *rk = sk + (size_t)ev * rec, *rm = sm + (size_t)ev * rec;
        for (int ew = 0; ew < nE; ++ew) {
            const int w = loc[ew];
            // (X_v, X_w): F^T B_vw F for v <= w; the transpose of the (w, v) value otherwise
            const int p = v <= w ? v : w, q = v <= w ? w : v;
            for (int j = 0; j < 2; ++j)
                for (int k = 0; k < 2; ++k) {
                    const int jj = v <= w ? j : k, kk = v <= w ? k : j;
                    double gk = 0.0, gm = 0.0;
                    for (int a = 0; a < 3; ++a)
                        for (int b = 0; b < 3; ++b) {
                            gk += Fg[a][jj] * blk(K, nullptr, nv, p, q, a, b) * Fg[b][kk];
                            gm += Fg[a][jj] * mblk(m, nv, p, q, a, b) * Fg[b][kk];
                        }
                    rk[4 * ew + 2 * j + k] = gm + gk;
                    rm[4 * ew + 2 * j + k] = gm;
                }
        }
        for (int u = 0; u < nv; ++u)
            for (int j = 0; j < 2; ++j)
                for (int k = 0; k < 3; ++k) {
                    double gk = 0.0, gm = 0.0;
                    if (u >= v) {
                        for (int a = 0; a < 3; ++a) {
                            gk -= Fg[a][j] * blk(K, nullptr, nv, v, u, a, k);
                            gm -= Fg[a][j] * mblk(m, nv, v, u, a, k);
                        }
                    } else {
                        for (int a = 0; a < 3; ++a) {
                            gk -= blk(K, nullptr, nv, u, v, k, a) * Fg[a][j];
                            gm -= mblk(m, nv, u, v, k, a) * Fg[a][j];
                        }
                    }
                    rk[4 * nE + 6 * u + 3 * j + k] = gm + gk;
                    rm[4 * nE + 6 * u + 3 * j + k] = gm;
                }
        for (int j = 0; j < 2; ++j) {
            double s = 0.0;
            for (int a = 0; a < 3; ++a) s -= Fg[a][j] * fv[3 * v + a];
            sf[2 * ev + j] = s;
        }
    }
}

__global__ void __launch_bounds__(256)
k_eol_gather(int64_t n, const int64_t *__restrict__ ptr, const int *__restrict__ list, const double *__restrict__ stK,
             const double *__restrict__ stM, double *__restrict__ A, double *__restrict__ M)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double a = 0.0, m = 0.0;
    for (int64_t k = ptr[i]; k < ptr[i + 1]; ++k) {
        a += stK[list[k]];
        m += stM[list[k]];
    }
    A[i] = a;
    M[i] = m;
}

__global__ void __launch_bounds__(256)
k_eol_fx(int n, const int *__restrict__ ptr, const int *__restrict__ list, const double *__restrict__ stF,
         double *__restrict__ fX)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double s = 0.0;
    for (int k = ptr[i]; k < ptr[i + 1]; ++k) s += stF[list[k]];
    fX[i] = s;
}

__global__ void __launch_bounds__(256)
k_eol_xforce(int n, const int *__restrict__ verts, const int *__restrict__ ptr, const int *__restrict__ slots,
             const double *__restrict__ stageF, double *__restrict__ f)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int v = verts[i];
    double f0 = f[3 * v], f1 = f[3 * v + 1], f2 = f[3 * v + 2];
    for (int k = ptr[i]; k < ptr[i + 1]; ++k) {
        const double *s = stageF + (size_t)3 * slots[k];
        f0 += s[0]; f1 += s[1]; f2 += s[2];
    }
    f[3 * v] = f0; f[3 * v + 1] = f1; f[3 * v + 2] = f2;
}

inline int blocks_for(int64_t n, int t = 256) { return (int)((n + t - 1) / t); }

}  // namespace

namespace eolc {

hipError_t eol_fill(eolc_ctx *c)
{
    const EolState &S = c->eol;
    if (S.E == 0) return hipSuccess;
    if (S.ne > 0)
        hipLaunchKernelGGL(k_eol_elem, dim3(blocks_for(S.ne)), dim3(256), 0, c->stream, S.ne, c->topo.F, S.d_elem,
                           c->tri, c->hinge, S.d_eidx, c->x, c->X, c->stageK, c->stageM, c->stageF, S.d_soff,
                           S.d_foff, S.d_stK, S.d_stM, S.d_stF);
    const int64_t nnz = S.rowptr.empty() ? 0 : S.rowptr.back();
    if (nnz > 0)
        hipLaunchKernelGGL(k_eol_gather, dim3(blocks_for(nnz)), dim3(256), 0, c->stream, nnz, S.d_cptr, S.d_clist,
                           S.d_stK, S.d_stM, S.d_A, S.d_M);
    hipLaunchKernelGGL(k_eol_fx, dim3(blocks_for(2 * S.E)), dim3(256), 0, c->stream, 2 * S.E, S.d_fptr, S.d_flist,
                       S.d_stF, S.d_fX);
    if (S.nxv > 0 && !(c->flags & EOLC_OPT_INCLUDE_BENDING_FORCE))
        hipLaunchKernelGGL(k_eol_xforce, dim3(blocks_for(S.nxv)), dim3(256), 0, c->stream, S.nxv, S.d_xv, S.d_xvptr,
                           S.d_xvslot, c->stageF, c->f);
    return hipGetLastError();
}

}  // namespace eolc
