// Batched cloths (gfx950): many independent cloth instances, concatenated into one
// context, solved in ONE persistent launch.
//
// Every instance's vertex range is padded to a multiple of 64, so each 64-row
// SELL slice of the concatenated matrix belongs to exactly one instance; padding
// vertices are isolated and fixed.  Workgroup w of the persistent solve owns
// instance order[w] (largest first, so consecutive large instances land on
// different XCDs, which receive workgroups round-robin) and runs that instance's
// whole mixed-precision PCG -- the velocity solve of GeneralizedSolver::velocitySolve
// for solver "none" (GeneralizedSolver.cpp:107-172) -- to its own tolerance:
//   FP32 difference-form SpMV (row mass plus S'_ij (p_j - p_i), the diagonal
//   entry skipped) and block-Jacobi, FP64 dot products, FP64 true
//   residual at every reliable update (delta), iterate accumulated in FP64.
// No grid-wide synchronisation exists: the instances are independent, a workgroup
// only touches its own rows, and an instance that has converged simply leaves
// (its workgroup exits) -- the per-instance convergence mask.
//
// Reductions: each wave reduces its slice's 64 rows with a fixed shuffle tree and
// writes one partial per slice; after a barrier wave 0 sums the instance's slice
// partials in slice order.  The result does not depend on the workgroup size, so
// the autotuned launch shape never changes the answer.
#include <hip/hip_runtime.h>

#include <algorithm>

#include "eolc_internal.h"

// Multiply-adds are written out (fma where fused); the compiler contracts nothing,
// so every workgroup-size instantiation rounds identically.
#pragma clang fp contract(off)

namespace {

constexpr int SL = eolc::kSlice;

inline int blocks_for(int64_t n, int t = 256) { return (int)((n + t - 1) / t); }

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Sum over the instance's slices of K per-row values; the result is returned to
// every thread.  part: K x nslices scratch, red: K doubles of LDS.
template <int K>
__device__ __forceinline__ void inst_sum(double (&v)[K], const eolc::BatchInst &I, const double *__restrict__ part,
                                         int nslices, double *red)
{
    __syncthreads();   // the per-slice partials of every wave are written
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (wave == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double s = 0.0;
            for (int sl = I.s0 + lane; sl < I.s1; sl += 64) s += part[k * nslices + sl];
            s = wave_sum(s);
            if (lane == 0) red[k] = s;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = red[k];
}

// Per-slice partial of K values held one per lane (all 64 lanes of the wave).
template <int K>
__device__ __forceinline__ void slice_partial(const double (&v)[K], int sl, double *__restrict__ part, int nslices)
{
#pragma unroll
    for (int k = 0; k < K; ++k) {
        const double s = wave_sum(v[k]);
        if ((threadIdx.x & 63) == 0) part[k * nslices + sl] = s;
    }
}

// FP64 row product y = (A x)_row over the SELL entries of one row.
__device__ __forceinline__ void row_spmv64(const double *__restrict__ A, int64_t nbs, const int *__restrict__ cl,
                                           int base, int len, const double *__restrict__ x, double &y0, double &y1,
                                           double &y2)
{
    y0 = y1 = y2 = 0.0;
    for (int kk = 0; kk < len; ++kk) {
        const int e = base + kk * SL;
        const int j = cl[e];
        const double u0 = x[3 * j], u1 = x[3 * j + 1], u2 = x[3 * j + 2];
        y0 += fma(A[2 * nbs + e], u2, fma(A[1 * nbs + e], u1, A[0 * nbs + e] * u0));
        y1 += fma(A[5 * nbs + e], u2, fma(A[4 * nbs + e], u1, A[3 * nbs + e] * u0));
        y2 += fma(A[8 * nbs + e], u2, fma(A[7 * nbs + e], u1, A[6 * nbs + e] * u0));
    }
}

// out = b - A x (masked), returns ||out||^2 over the instance.  b may be null.
template <int WGS>
__device__ double inst_residual64(const eolc::BatchInst &I, const int *__restrict__ slice_off,
                                  const int *__restrict__ srow, const int *__restrict__ slen,
                                  const int *__restrict__ scol, const double *__restrict__ A, int64_t nbs,
                                  const unsigned char *__restrict__ fixmask, const double *__restrict__ x,
                                  const double *__restrict__ b, double *__restrict__ out, double *part, int nslices,
                                  double *red)
{
    constexpr int NW = WGS / 64;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int sl = I.s0 + wave; sl < I.s1; sl += NW) {
        const int g = sl * SL + lane;
        const int row = srow[g];
        double y0, y1, y2;
        row_spmv64(A, nbs, scol, slice_off[sl] + lane, slen[g], x, y0, y1, y2);
        double o0 = (b ? b[3 * row] : 0.0) - y0, o1 = (b ? b[3 * row + 1] : 0.0) - y1,
               o2 = (b ? b[3 * row + 2] : 0.0) - y2;
        const unsigned fm = fixmask[row];
        if (fm & 1) o0 = 0.0;
        if (fm & 2) o1 = 0.0;
        if (fm & 4) o2 = 0.0;
        out[3 * row] = o0; out[3 * row + 1] = o1; out[3 * row + 2] = o2;
        const double v[1] = {(o0 * o0 + o1 * o1) + o2 * o2};
        slice_partial<1>(v, sl, part, nslices);
    }
    double s[1];
    inst_sum<1>(s, I, part, nslices, red);
    return s[0];
}

// r32 = (float) r64, z = D r32; returns (r.z, r.r) over the instance.
template <int WGS>
__device__ void inst_precond(const eolc::BatchInst &I, const int *__restrict__ srow, int N,
                             const float *__restrict__ D, const double *__restrict__ r64, float *__restrict__ r32,
                             float *__restrict__ z32, double *part, int nslices, double *red, double &rz, double &rr)
{
    constexpr int NW = WGS / 64;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int sl = I.s0 + wave; sl < I.s1; sl += NW) {
        const int i = srow[sl * SL + lane];
        const float r0 = (float)r64[3 * i], r1 = (float)r64[3 * i + 1], r2 = (float)r64[3 * i + 2];
        r32[3 * i] = r0; r32[3 * i + 1] = r1; r32[3 * i + 2] = r2;
        const float *d = D + i;
        const double s0 = r0, s1 = r1, s2 = r2;
        const double e00 = d[0], e01 = d[N], e02 = d[2 * N], e11 = d[3 * N], e12 = d[4 * N], e22 = d[5 * N];
        const float z0 = (float)fma(e02, s2, fma(e01, s1, e00 * s0));
        const float z1 = (float)fma(e12, s2, fma(e11, s1, e01 * s0));
        const float z2 = (float)fma(e22, s2, fma(e12, s1, e02 * s0));
        z32[3 * i] = z0; z32[3 * i + 1] = z1; z32[3 * i + 2] = z2;
        const double v[2] = {(s0 * (double)z0 + s1 * (double)z1) + s2 * (double)z2, (s0 * s0 + s1 * s1) + s2 * s2};
        slice_partial<2>(v, sl, part, nslices);
    }
    double s[2];
    inst_sum<2>(s, I, part, nslices, red);
    rz = s[0];
    rr = s[1];
}

// Multilevel coarse levels inside the persistent solve: the hierarchy of the
// concatenated topology never spans two instances, and its hierarchical numbering
// gives every instance a contiguous aggregate range per level (lo), so a workgroup
// applies its instance's levels alone, with workgroup barriers between levels.
struct BatchML {
    eolc::MLLevels lv;
    const int *hord, *ptr1, *agg1, *par, *cptr;
    const int *lo;   // (B + 1) x (kMLMaxLevels + 1): first aggregate of instance b at level l
    const double *Dinv;
    double *R, *C;
};

__device__ __forceinline__ void ml_inv_apply(const double *__restrict__ d, double r0, double r1, double r2, double &c0,
                                             double &c1, double &c2)
{
    c0 = fma(d[2], r2, fma(d[1], r1, d[0] * r0));
    c1 = fma(d[4], r2, fma(d[3], r1, d[1] * r0));
    c2 = fma(d[5], r2, fma(d[4], r1, d[2] * r0));
}

// z = D r + sum_l Z_l D_l^-1 Z_l^T r over instance b (r32 of its rows complete);
// returns r.z over the instance.  Sums in member / child order, FP64.
template <int WGS>
__device__ double inst_ml_apply(const eolc::BatchInst &I, int b, const BatchML &m, const int *__restrict__ srow,
                                int N, const float *__restrict__ D, const float *__restrict__ r32,
                                float *__restrict__ z32, const unsigned char *__restrict__ fixmask, double *part,
                                int nslices, double *red)
{
    constexpr int NW = WGS / 64;
    constexpr int S = eolc::kMLMaxLevels + 1;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int L = m.lv.L;
    __syncthreads();   // every row's r32 is written
    {
        const int lo = m.lo[b * S + 1], hi = m.lo[(b + 1) * S + 1];
        for (int a = lo + (int)threadIdx.x; a < hi; a += WGS) {
            double t0 = 0.0, t1 = 0.0, t2 = 0.0;
            for (int k = m.ptr1[a]; k < m.ptr1[a + 1]; ++k) {
                const int i = m.hord[k];
                t0 += (double)r32[3 * i];
                t1 += (double)r32[3 * i + 1];
                t2 += (double)r32[3 * i + 2];
            }
            m.R[3 * a] = t0; m.R[3 * a + 1] = t1; m.R[3 * a + 2] = t2;
        }
    }
    __syncthreads();
    for (int l = 2; l <= L; ++l) {
        const int lo = m.lo[b * S + l], hi = m.lo[(b + 1) * S + l];
        for (int A = lo + (int)threadIdx.x; A < hi; A += WGS) {
            const int c0 = m.cptr[m.lv.coff[l] + A], c1 = m.cptr[m.lv.coff[l] + A + 1];
            double t0 = 0.0, t1 = 0.0, t2 = 0.0;
            for (int a = c0; a < c1; ++a) {
                const double *q = m.R + 3 * (int64_t)(m.lv.off[l - 1] + a);
                t0 += q[0]; t1 += q[1]; t2 += q[2];
            }
            double *o = m.R + 3 * (int64_t)(m.lv.off[l] + A);
            o[0] = t0; o[1] = t1; o[2] = t2;
        }
        __syncthreads();
    }
    for (int l = L; l >= 1; --l) {
        const int lo = m.lo[b * S + l], hi = m.lo[(b + 1) * S + l];
        for (int a = lo + (int)threadIdx.x; a < hi; a += WGS) {
            const int64_t g = m.lv.off[l] + a;
            double c0, c1, c2;
            ml_inv_apply(m.Dinv + 6 * g, m.R[3 * g], m.R[3 * g + 1], m.R[3 * g + 2], c0, c1, c2);
            if (l < L) {
                const double *up = m.C + 3 * (int64_t)(m.lv.off[l + 1] + m.par[g]);
                c0 += up[0]; c1 += up[1]; c2 += up[2];
            }
            m.C[3 * g] = c0; m.C[3 * g + 1] = c1; m.C[3 * g + 2] = c2;
        }
        __syncthreads();
    }
    for (int sl = I.s0 + wave; sl < I.s1; sl += NW) {
        const int i = srow[sl * SL + lane];
        const float r0 = r32[3 * i], r1 = r32[3 * i + 1], r2 = r32[3 * i + 2];
        float f0, f1, f2;
        eolc::apply_sym3<float>(D + i, N, r0, r1, r2, f0, f1, f2);
        const double *cc = m.C + 3 * (int64_t)m.agg1[i];
        const unsigned fm = fixmask[i];
        const float z0 = (fm & 1u) ? 0.0f : (float)((double)f0 + cc[0]);
        const float z1 = (fm & 2u) ? 0.0f : (float)((double)f1 + cc[1]);
        const float z2 = (fm & 4u) ? 0.0f : (float)((double)f2 + cc[2]);
        z32[3 * i] = z0; z32[3 * i + 1] = z1; z32[3 * i + 2] = z2;
        const double v[1] = {((double)r0 * (double)z0 + (double)r1 * (double)z1) + (double)r2 * (double)z2};
        slice_partial<1>(v, sl, part, nslices);
    }
    double s[1];
    inst_sum<1>(s, I, part, nslices, red);
    return s[0];
}

// r32 = (float) r64 and r.r over the instance (the multilevel path forms z after).
template <int WGS>
__device__ double inst_convert(const eolc::BatchInst &I, const int *__restrict__ srow, const double *__restrict__ r64,
                               float *__restrict__ r32, double *part, int nslices, double *red)
{
    constexpr int NW = WGS / 64;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    for (int sl = I.s0 + wave; sl < I.s1; sl += NW) {
        const int i = srow[sl * SL + lane];
        const float r0 = (float)r64[3 * i], r1 = (float)r64[3 * i + 1], r2 = (float)r64[3 * i + 2];
        r32[3 * i] = r0; r32[3 * i + 1] = r1; r32[3 * i + 2] = r2;
        const double s0 = r0, s1 = r1, s2 = r2;
        const double v[1] = {(s0 * s0 + s1 * s1) + s2 * s2};
        slice_partial<1>(v, sl, part, nslices);
    }
    double s[1];
    inst_sum<1>(s, I, part, nslices, red);
    return s[0];
}

struct BatchArgs {
    const eolc::BatchInst *inst;
    const int *order;
    const int *slice_off, *srow, *slen, *scol;
    const double *A64;
    const float *S32, *M32, *D32;
    int64_t nbs;
    int N, nslices;
    const unsigned char *fixmask;
    const double *vfix, *rhs;
    double *v, *xs64, *r64, *b64;
    float *r32, *z32, *q32, *xl32, *p32a, *p32b;
    double *part;
    double tol2, delta2;
    int maxit;
    int *stats;
    double *dstats;
    BatchML ml;
};

// Persistent mixed-precision PCG, one workgroup per instance.  ML: the preconditioner
// adds the instance's multilevel coarse levels to the block-Jacobi blocks.
template <int WGS, bool ML>
__global__ void __launch_bounds__(WGS) k_batch_pcg(BatchArgs a)
{
    constexpr int NW = WGS / 64;
    __shared__ double red[4];
    const int b = a.order[blockIdx.x];
    const eolc::BatchInst I = a.inst[b];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int N = a.N, ns = a.nslices;
    const int64_t nbs = a.nbs;

    // b_red = rhs - A vfix (masked); x0 = masked v_old; r = b_red - A x0
    const double bnorm2 = inst_residual64<WGS>(I, a.slice_off, a.srow, a.slen, a.scol, a.A64, nbs, a.fixmask, a.vfix,
                                               a.rhs, a.b64, a.part, ns, red);
    for (int sl = I.s0 + wave; sl < I.s1; sl += NW) {
        const int i = a.srow[sl * SL + lane];
        const unsigned fm = a.fixmask[i];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            a.xs64[3 * i + c] = ((fm >> c) & 1) ? 0.0 : a.v[3 * i + c];
            a.xl32[3 * i + c] = 0.0f;
            a.p32a[3 * i + c] = 0.0f;
        }
    }
    __syncthreads();
    double rmax2 = inst_residual64<WGS>(I, a.slice_off, a.srow, a.slen, a.scol, a.A64, nbs, a.fixmask, a.xs64, a.b64,
                                        a.r64, a.part, ns, red);
    double rz, rr;
    if constexpr (ML) {
        rr = inst_convert<WGS>(I, a.srow, a.r64, a.r32, a.part, ns, red);
        rz = inst_ml_apply<WGS>(I, b, a.ml, a.srow, N, a.D32, a.r32, a.z32, a.fixmask, a.part, ns, red);
    } else {
        inst_precond<WGS>(I, a.srow, N, a.D32, a.r64, a.r32, a.z32, a.part, ns, red, rz, rr);
    }
    double rz_prev = 1.0;
    int k = 0, first = 0, refinements = 0, status = EOLC_ERR_MAXIT;
    float *pold = a.p32a, *pnew = a.p32b;
    const double thr_final = a.tol2 * bnorm2;
    for (;;) {
        const double thr = fmax(thr_final, a.delta2 * rmax2);
        if (!isfinite(rr) || !isfinite(rz)) {
            status = EOLC_ERR_NONFINITE;
            break;
        }
        if (rr <= thr) {
            // reliable update: fold the FP32 correction in, FP64 true residual
            for (int sl = I.s0 + wave; sl < I.s1; sl += NW) {
                const int i = a.srow[sl * SL + lane];
#pragma unroll
                for (int c = 0; c < 3; ++c) {
                    a.xs64[3 * i + c] += (double)a.xl32[3 * i + c];
                    a.xl32[3 * i + c] = 0.0f;
                }
            }
            __syncthreads();
            rmax2 = inst_residual64<WGS>(I, a.slice_off, a.srow, a.slen, a.scol, a.A64, nbs, a.fixmask, a.xs64, a.b64,
                                         a.r64, a.part, ns, red);
            ++refinements;
            if (!isfinite(rmax2)) {
                status = EOLC_ERR_NONFINITE;
                break;
            }
            if (rmax2 <= thr_final) {
                status = EOLC_OK;
                break;
            }
            if (k >= a.maxit || refinements > eolc::kMaxRefinements) break;
            if constexpr (ML) {
                rr = inst_convert<WGS>(I, a.srow, a.r64, a.r32, a.part, ns, red);
                rz = inst_ml_apply<WGS>(I, b, a.ml, a.srow, N, a.D32, a.r32, a.z32, a.fixmask, a.part, ns, red);
            } else {
                inst_precond<WGS>(I, a.srow, N, a.D32, a.r64, a.r32, a.z32, a.part, ns, red, rz, rr);
            }
            continue;
        }
        if (k >= a.maxit) break;
        const double beta = (k == first) ? 0.0 : rz / rz_prev;
        const float betaT = (float)beta;
        // p_new = z + beta p_old (own rows; neighbours formed on the fly), q = A p_new
        for (int sl = I.s0 + wave; sl < I.s1; sl += NW) {
            const int g = sl * SL + lane;
            const int row = a.srow[g], len = a.slen[g] - 1;   // the diagonal (last entry) is skipped
            const int base = a.slice_off[sl] + lane;
            const float ui0 = fmaf(betaT, pold[3 * row], a.z32[3 * row]);
            const float ui1 = fmaf(betaT, pold[3 * row + 1], a.z32[3 * row + 1]);
            const float ui2 = fmaf(betaT, pold[3 * row + 2], a.z32[3 * row + 2]);
            const float mb = a.M32[row];
            double y0 = (double)(mb * ui0), y1 = (double)(mb * ui1), y2 = (double)(mb * ui2);
            for (int kk = 0; kk < len; ++kk) {
                const int e = base + kk * SL;
                const int j = a.scol[e];
                const float u0 = fmaf(betaT, pold[3 * j], a.z32[3 * j]);
                const float u1 = fmaf(betaT, pold[3 * j + 1], a.z32[3 * j + 1]);
                const float u2 = fmaf(betaT, pold[3 * j + 2], a.z32[3 * j + 2]);
                const float d0 = u0 - ui0, d1 = u1 - ui1, d2 = u2 - ui2;
                const float *s = a.S32 + e;
                y0 += (double)fmaf(s[2 * nbs], d2, fmaf(s[1 * nbs], d1, s[0 * nbs] * d0));
                y1 += (double)fmaf(s[5 * nbs], d2, fmaf(s[4 * nbs], d1, s[3 * nbs] * d0));
                y2 += (double)fmaf(s[8 * nbs], d2, fmaf(s[7 * nbs], d1, s[6 * nbs] * d0));
            }
            const unsigned fm = a.fixmask[row];
            if (fm & 1) y0 = 0.0;
            if (fm & 2) y1 = 0.0;
            if (fm & 4) y2 = 0.0;
            pnew[3 * row] = ui0; pnew[3 * row + 1] = ui1; pnew[3 * row + 2] = ui2;
            const float q0 = (float)y0, q1 = (float)y1, q2 = (float)y2;
            a.q32[3 * row] = q0; a.q32[3 * row + 1] = q1; a.q32[3 * row + 2] = q2;
            const double v[1] = {((double)ui0 * (double)q0 + (double)ui1 * (double)q1) + (double)ui2 * (double)q2};
            slice_partial<1>(v, sl, a.part, ns);
        }
        double pq[1];
        inst_sum<1>(pq, I, a.part, ns, red);
        if (!(pq[0] > 0.0)) {
            status = isfinite(pq[0]) ? EOLC_ERR_NOT_SPD : EOLC_ERR_NONFINITE;
            break;
        }
        const double alpha = rz / pq[0];
        const float alphaT = (float)alpha;
        // x += alpha p, r -= alpha q, z = D r
        for (int sl = I.s0 + wave; sl < I.s1; sl += NW) {
            const int i = a.srow[sl * SL + lane];
#pragma unroll
            for (int c = 0; c < 3; ++c) a.xl32[3 * i + c] = fmaf(alphaT, pnew[3 * i + c], a.xl32[3 * i + c]);
            const float r0 = fmaf(-alphaT, a.q32[3 * i], a.r32[3 * i]);
            const float r1 = fmaf(-alphaT, a.q32[3 * i + 1], a.r32[3 * i + 1]);
            const float r2 = fmaf(-alphaT, a.q32[3 * i + 2], a.r32[3 * i + 2]);
            a.r32[3 * i] = r0; a.r32[3 * i + 1] = r1; a.r32[3 * i + 2] = r2;
            if constexpr (ML) {
                const double v[1] = {((double)r0 * (double)r0 + (double)r1 * (double)r1) + (double)r2 * (double)r2};
                slice_partial<1>(v, sl, a.part, ns);
            } else {
                float z0, z1, z2;
                eolc::apply_sym3<float>(a.D32 + i, N, r0, r1, r2, z0, z1, z2);
                a.z32[3 * i] = z0; a.z32[3 * i + 1] = z1; a.z32[3 * i + 2] = z2;
                const double v[2] = {((double)r0 * (double)z0 + (double)r1 * (double)z1) + (double)r2 * (double)z2,
                                     ((double)r0 * (double)r0 + (double)r1 * (double)r1) + (double)r2 * (double)r2};
                slice_partial<2>(v, sl, a.part, ns);
            }
        }
        rz_prev = rz;
        if constexpr (ML) {
            double s1[1];
            inst_sum<1>(s1, I, a.part, ns, red);
            rr = s1[0];
            rz = inst_ml_apply<WGS>(I, b, a.ml, a.srow, N, a.D32, a.r32, a.z32, a.fixmask, a.part, ns, red);
        } else {
            double s2[2];
            inst_sum<2>(s2, I, a.part, ns, red);
            rz = s2[0];
            rr = s2[1];
        }
        float *t = pold;
        pold = pnew;
        pnew = t;
        ++k;
    }
    if (status != EOLC_OK && status != EOLC_ERR_NOT_SPD) {
        for (int sl = I.s0 + wave; sl < I.s1; sl += NW) {
            const int i = a.srow[sl * SL + lane];
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                a.xs64[3 * i + c] += (double)a.xl32[3 * i + c];
                a.xl32[3 * i + c] = 0.0f;
            }
        }
    }
    // v = fixed ? vfix : x
    for (int sl = I.s0 + wave; sl < I.s1; sl += NW) {
        const int i = a.srow[sl * SL + lane];
        const unsigned fm = a.fixmask[i];
#pragma unroll
        for (int c = 0; c < 3; ++c) a.v[3 * i + c] = ((fm >> c) & 1) ? a.vfix[3 * i + c] : a.xs64[3 * i + c];
    }
    __syncthreads();
    const double true_rr = inst_residual64<WGS>(I, a.slice_off, a.srow, a.slen, a.scol, a.A64, nbs, a.fixmask, a.v,
                                                a.rhs, a.r64, a.part, ns, red);
    if (threadIdx.x == 0) {
        a.stats[4 * b + 0] = k;
        a.stats[4 * b + 1] = refinements;
        a.stats[4 * b + 2] = status;
        a.dstats[4 * b + 0] = bnorm2;
        a.dstats[4 * b + 1] = true_rr;
        a.dstats[4 * b + 2] = bnorm2 > 0.0 ? sqrt(rr / bnorm2) : 0.0;
    }
}

// Fixed values of the caller's fixed nodes (all three DOFs), Constraints.cpp:114-119:
// v_fixed = (1 - 0.01) v_old + scripted velocity.
__global__ void k_batch_vfix(int n, const int *__restrict__ nodes, const double *__restrict__ scripted,
                             const double *__restrict__ v, double *__restrict__ vfix)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * n) return;
    const int i = nodes[t / 3], c = t % 3;
    vfix[3 * i + c] = (1 - 0.01) * v[3 * i + c] + scripted[t];
}

// All DOFs of the listed nodes fixed (set once per batch topology).
__global__ void k_batch_fixmask(int n, const int *__restrict__ nodes, unsigned char *__restrict__ fixmask)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) fixmask[nodes[t]] = 7;
}

template <int WGS>
void launch_batch(eolc_ctx *c, const BatchArgs &args, bool ml)
{
    if (ml)
        hipLaunchKernelGGL((k_batch_pcg<WGS, true>), dim3(c->batch.B), dim3(WGS), 0, c->stream, args);
    else
        hipLaunchKernelGGL((k_batch_pcg<WGS, false>), dim3(c->batch.B), dim3(WGS), 0, c->stream, args);
}

}  // namespace

namespace eolc {

hipError_t batch_set_fixmask(eolc_ctx *c, const int *d_nodes, int n)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_batch_fixmask, dim3(blocks_for(n)), dim3(256), 0, c->stream, n, d_nodes, c->fixmask);
    return hipGetLastError();
}

// Enqueue the solve part of one batched step: fixed values, block-Jacobi blocks and
// the persistent PCG (all kernels, no host synchronisation: capturable).
hipError_t batch_solve_async(eolc_ctx *c, double tol, int maxit, int variant)
{
    BatchState &S = c->batch;
    const Topology &t = c->topo;
    const int nu = (int)S.user_fixed.size();
    if (nu > 0)
        hipLaunchKernelGGL(k_batch_vfix, dim3(blocks_for(3 * nu)), dim3(256), 0, c->stream, nu, S.d_user_fixed, S.d_scripted,
                           c->v, c->vfix);
    const bool ml = S.ml && S.d_ml_lo && c->ml.lv.L > 0;
    c->ml_on = ml;   // pcg_setup then also forms the coarse levels' blocks (translations only)
    hipError_t e = pcg_setup(c, true, 0, false);
    c->ml_on = false;
    if (e != hipSuccess) return e;
    BatchArgs a;
    a.inst = S.d_inst; a.order = S.d_order;
    a.slice_off = c->slice_off; a.srow = c->srow; a.slen = c->slen; a.scol = c->scol;
    a.A64 = c->A64; a.S32 = c->S32; a.M32 = c->M32; a.D32 = c->D32;
    a.nbs = t.nbs; a.N = t.N; a.nslices = t.nslices;
    a.fixmask = c->fixmask; a.vfix = c->vfix; a.rhs = c->rhs;
    a.v = c->v; a.xs64 = c->xs64; a.r64 = c->r64; a.b64 = c->b64;
    a.r32 = c->r32; a.z32 = c->z32; a.q32 = c->q32; a.xl32 = c->xl32; a.p32a = c->p32[0]; a.p32b = c->p32[1];
    a.part = S.slice_part;
    a.tol2 = tol * tol; a.delta2 = c->batch_delta * c->batch_delta; a.maxit = maxit;
    a.stats = S.stats; a.dstats = S.dstats;
    const MLState &m = c->ml;
    a.ml = BatchML{m.lv, m.hord, m.ptr1, m.agg1, m.par, m.cptr, S.d_ml_lo, m.Dinv, m.R, m.C};
    switch (variant) {
    case 0: launch_batch<256>(c, a, ml); break;
    case 1: launch_batch<512>(c, a, ml); break;
    default: launch_batch<1024>(c, a, ml); break;
    }
    return hipGetLastError();
}

}  // namespace eolc
