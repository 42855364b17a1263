// Additive multilevel preconditioner over an aggregation hierarchy (gfx950).
//
//   z = D0 r + sum_{l = 1..L} Z_l D_l^-1 Z_l^T r
//
// D0 holds the fine 3x3 blocks (block-Jacobi inverse, or the learned blocks).  Level
// l groups the vertices into nested aggregates (ml_host.cpp); Z_l is the piecewise-
// constant prolongation that copies one 3-vector per aggregate to its member
// vertices (zero on fixed DOFs), and D_l is the 3x3 diagonal block of the Galerkin
// operator Z_l^T A Z_l: the sum of all blocks A_ij, fixed DOFs masked, whose two
// vertices lie in the same level-l aggregate.  Each term is symmetric positive
// (semi-)definite for an SPD A, so z = P r defines an SPD preconditioner: an additive
// multilevel method of the BPX family (Bramble, Pasciak and Xu 1990) with
// aggregation coarse spaces.  The coarse terms carry the smooth, slowly converging
// modes (sheet-wide translations and bending) that no per-vertex block reaches.
//
// Set-up per matrix (ml_setup): k_ml_diag1 sums, per level-1 aggregate, the blocks
// of its member rows bucketed by the first level whose aggregate holds both vertices
// (`lam`, precomputed per SELL entry); k_ml_diagup adds children upwards.  Every
// level's blocks are inverted in FP64.
//
// Per PCG iteration, after the vector update (ml_apply):
//   k_ml_restrict : one wavefront per level-2 aggregate, vertices staged through LDS:
//                   R_1(a) = sum of r over a's members, R_2 = sum of the children's R_1,
//                   and per level-1 aggregate a, C(a) = D_1^-1 R_1(a) + D_2^-1 R_2(parent)
//   k_ml_coarse   : one workgroup (levels >= 3 only): R_l for l >= 3, then
//                   C_l = D_l^-1 R_l + C_{l+1}(parent) from the top level down to 3
//   k_ml_prolong  : z_i = D0 r_i + C(a_i) + C_3(level-3 aggregate of i); r.z partials
// All coarse arithmetic is FP64 in a fixed order, so the result does not depend on
// the launch shape.
#include "eolc_internal.h"
#include "pcg_device.h"

using eolc::MLLevels;
using eolc::SolverScalars;

namespace {

constexpr int kML = eolc::kMLMaxLevels;

// Inverse of a symmetric 3x3 block (upper entries out).  A component whose diagonal
// is not positive (all of the aggregate's DOFs in it fixed) is left out: its row
// and column of the inverse are zero.  A block that is not positive definite on
// the remaining components gives a zero inverse (no correction from it).
__device__ __forceinline__ void inv_spd3(const double a[9], double *__restrict__ out)
{
    double s00 = a[0], s11 = a[4], s22 = a[8];
    double s01 = 0.5 * (a[1] + a[3]), s02 = 0.5 * (a[2] + a[6]), s12 = 0.5 * (a[5] + a[7]);
    const bool k0 = s00 > 0.0, k1 = s11 > 0.0, k2 = s22 > 0.0;
    if (!k0) { s00 = 1.0; s01 = 0.0; s02 = 0.0; }
    if (!k1) { s11 = 1.0; s01 = 0.0; s12 = 0.0; }
    if (!k2) { s22 = 1.0; s02 = 0.0; s12 = 0.0; }
    const double c00 = s11 * s22 - s12 * s12, c01 = s02 * s12 - s01 * s22, c02 = s01 * s12 - s02 * s11;
    const double m2 = s00 * s11 - s01 * s01;
    const double det = s00 * c00 + s01 * c01 + s02 * c02;
    double o[6] = {0, 0, 0, 0, 0, 0};
    if (s00 > 0.0 && m2 > 0.0 && det > 0.0 && isfinite(det)) {
        const double id = 1.0 / det;
        o[0] = c00 * id;
        o[1] = c01 * id;
        o[2] = c02 * id;
        o[3] = (s00 * s22 - s02 * s02) * id;
        o[4] = (s01 * s02 - s00 * s12) * id;
        o[5] = m2 * id;
        if (!k0) { o[0] = 0.0; o[1] = 0.0; o[2] = 0.0; }
        if (!k1) { o[1] = 0.0; o[3] = 0.0; o[4] = 0.0; }
        if (!k2) { o[2] = 0.0; o[4] = 0.0; o[5] = 0.0; }
    }
#pragma unroll
    for (int k = 0; k < 6; ++k) out[k] = o[k];
}

__device__ __forceinline__ void apply_inv(const double *__restrict__ d, double r0, double r1, double r2, double &c0,
                                          double &c1, double &c2)
{
    c0 = fma(d[2], r2, fma(d[1], r1, d[0] * r0));
    c1 = fma(d[4], r2, fma(d[3], r1, d[1] * r0));
    c2 = fma(d[5], r2, fma(d[4], r1, d[2] * r0));
}

// One thread per level-1 aggregate: V_1[m](a) = sum over a's member rows of the
// blocks (fixed DOFs masked) whose first common level is <= m + 1, m = 0..L-1;
// D_1^-1 from V_1[0].
__global__ void __launch_bounds__(256) k_ml_diag1(MLLevels lv, const int *__restrict__ hord,
                                                  const int *__restrict__ ptr1, const int *__restrict__ gpos,
                                                  const int *__restrict__ slice_off, const int *__restrict__ slen,
                                                  const int *__restrict__ scol, const unsigned char *__restrict__ lam,
                                                  const double *__restrict__ A, int64_t nbs,
                                                  const unsigned char *__restrict__ fixmask, double *__restrict__ V,
                                                  double *__restrict__ Dinv)
{
    const int a = blockIdx.x * blockDim.x + threadIdx.x;
    const int n1 = lv.n[1], L = lv.L;
    if (a >= n1) return;
    double acc[kML][9];
#pragma unroll
    for (int m = 0; m < kML; ++m)
#pragma unroll
        for (int c = 0; c < 9; ++c) acc[m][c] = 0.0;
    for (int k = ptr1[a]; k < ptr1[a + 1]; ++k) {
        const int i = hord[k];
        const int g = gpos[i];
        const int len = slen[g];
        const int base = slice_off[g / eolc::kSlice] + (g % eolc::kSlice);
        const unsigned fi = fixmask[i];
        for (int kk = 0; kk < len; ++kk) {
            const int e = base + kk * eolc::kSlice;
            const int lm = lam[e];
            if (lm > L) continue;
            const unsigned fj = fixmask[scol[e]];
            double b[9];
#pragma unroll
            for (int c = 0; c < 9; ++c) {
                const int r = c / 3, s = c % 3;
                b[c] = (((fi >> r) | (fj >> s)) & 1u) ? 0.0 : A[(int64_t)c * nbs + e];
            }
#pragma unroll
            for (int m = 0; m < kML; ++m)
                if (m < L && lm <= m + 1)
#pragma unroll
                    for (int c = 0; c < 9; ++c) acc[m][c] += b[c];
        }
    }
#pragma unroll
    for (int m = 0; m < kML; ++m)
        if (m < L)
#pragma unroll
            for (int c = 0; c < 9; ++c) V[(int64_t)(m * 9 + c) * n1 + a] = acc[m][c];
    inv_spd3(acc[0], Dinv + 6 * (int64_t)a);
}

// One thread per level-l aggregate (l >= 2): V_l[m] = sum of the children's V_{l-1}[m]
// for m >= l - 1; D_l^-1 from V_l[l - 1].
__global__ void __launch_bounds__(256) k_ml_diagup(int l, MLLevels lv, const int *__restrict__ cptr,
                                                   double *__restrict__ V, double *__restrict__ Dinv)
{
    const int A = blockIdx.x * blockDim.x + threadIdx.x;
    const int L = lv.L, nl = lv.n[l], nc = lv.n[l - 1];
    if (A >= nl) return;
    const double *__restrict__ Vc = V + (int64_t)9 * L * lv.off[l - 1];
    double *__restrict__ Vl = V + (int64_t)9 * L * lv.off[l];
    const int c0 = cptr[lv.coff[l] + A], c1 = cptr[lv.coff[l] + A + 1];
    double d[9];
    for (int m = l - 1; m < L; ++m)
        for (int c = 0; c < 9; ++c) {
            double s = 0.0;
            for (int a = c0; a < c1; ++a) s += Vc[(int64_t)(m * 9 + c) * nc + a];
            Vl[(int64_t)(m * 9 + c) * nl + A] = s;
            if (m == l - 1) d[c] = s;
        }
    inv_spd3(d, Dinv + 6 * (int64_t)(lv.off[l] + A));
}

// Ordering of LDS traffic between the lanes of one wavefront (no workgroup barrier:
// the wavefronts of a workgroup run different trip counts).
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Restriction of r to levels 1 and 2: one wavefront per level-2 aggregate.  Its
// vertices are a contiguous range of hord; the lanes stage 64 of them at a time in
// LDS (one vertex per lane), then each lane sums the members of one level-1 child
// that fall in the chunk, in member order.  R_2 is the wavefront sum of the
// children's R_1 (fixed shuffle tree).  With a single level, a wavefront takes 64
// level-1 aggregates.
// two wavefronts per workgroup: 13.9 us per restriction at 1024^2, against 14.8 (1),
// 14.2 (4) and 14.4 (8) in the same kernel trace
constexpr int kRestrictWG = 128;
template <class T>
__global__ void __launch_bounds__(kRestrictWG) k_ml_restrict(MLLevels lv, const int *__restrict__ hord,
                                                             const int *__restrict__ ptr1, const int *__restrict__ cptr,
                                                             const int *__restrict__ kptr2,
                                                             const T *__restrict__ r, double *__restrict__ R,
                                                             const double *__restrict__ Dinv, double *__restrict__ C,
                                                             const SolverScalars *scal, int check_done)
{
    __shared__ double stage[kRestrictWG / 64][3 * 64];
    const int lane = threadIdx.x % 64;
    double *__restrict__ buf = stage[threadIdx.x / 64];
    const int w = (blockIdx.x * blockDim.x + threadIdx.x) / 64;   // uniform in the wavefront
    const bool two = lv.L >= 2;
    int a0, a1, ka = 0, kz = 0;
    // the aggregate's child and member ranges are read together with the stop flag
    // (one latency round; a stopped launch only reads these words)
    if (two) {
        if (w >= lv.n[2]) return;
        a0 = cptr[lv.coff[2] + w];
        a1 = cptr[lv.coff[2] + w + 1];
        ka = kptr2[w];
        kz = kptr2[w + 1];
    } else {
        a0 = w * 64;
        if (a0 >= lv.n[1]) return;
        a1 = min(a0 + 64, lv.n[1]);
    }
    if (check_done && scal->done) return;
    // inverse blocks of the level-2 aggregate and of this lane's first child: they do
    // not depend on r, so they load with the member indices
    double dw[6] = {0, 0, 0, 0, 0, 0}, da[6] = {0, 0, 0, 0, 0, 0};
    if (two) {
        const int64_t gw = lv.off[2] + w;
#pragma unroll
        for (int k = 0; k < 6; ++k) dw[k] = Dinv[6 * gw + k];
        if (a0 + lane < a1)
#pragma unroll
            for (int k = 0; k < 6; ++k) da[k] = Dinv[6 * (int64_t)(a0 + lane) + k];
    }
    double s0 = 0.0, s1 = 0.0, s2 = 0.0;
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;   // R_1 of this lane's first child
    int k0 = two ? ka : ptr1[a0];
    for (int g0 = a0; g0 < a1; g0 += 64) {   // groups of up to 64 children
        const int g1 = min(g0 + 64, a1);
        const int a = g0 + lane;
        const bool mine = a < g1;
        const int m0 = mine ? ptr1[a] : 0, m1 = mine ? ptr1[a + 1] : 0;
        const int k1 = (two && g1 == a1) ? kz : ptr1[g1];
        double t0 = 0.0, t1 = 0.0, t2 = 0.0;
        for (int kb = k0; kb < k1; kb += 64) {
            const int k = kb + lane;
            if (k < k1) {
                const int i = hord[k];
                buf[3 * lane] = (double)r[3 * i];
                buf[3 * lane + 1] = (double)r[3 * i + 1];
                buf[3 * lane + 2] = (double)r[3 * i + 2];
            }
            wave_sync();
            const int e0 = max(m0, kb), e1 = min(m1, kb + 64);
            for (int m = e0; m < e1; ++m) {
                t0 += buf[3 * (m - kb)];
                t1 += buf[3 * (m - kb) + 1];
                t2 += buf[3 * (m - kb) + 2];
            }
            wave_sync();
        }
        if (mine) {
            R[3 * a] = t0; R[3 * a + 1] = t1; R[3 * a + 2] = t2;
            s0 += t0; s1 += t1; s2 += t2;
            if (g0 == a0) { f0 = t0; f1 = t1; f2 = t2; }
        }
        k0 = k1;   // the next group starts where this one ends
    }
    if (two) {
        // wave_sum leaves the total in lane 0; every lane needs it below
        s0 = __shfl(wave_sum(s0), 0, 64);
        s1 = __shfl(wave_sum(s1), 0, 64);
        s2 = __shfl(wave_sum(s2), 0, 64);
        if (lane == 0) {
            double *o = R + 3 * (lv.off[2] + (int64_t)w);
            o[0] = s0; o[1] = s1; o[2] = s2;
        }
        // levels 1 and 2 of each child, for the prolongation: C(a) = D_1^-1 R_1(a) +
        // D_2^-1 R_2(w)
        double e0, e1, e2, c0, c1, c2;
        apply_inv(dw, s0, s1, s2, e0, e1, e2);
        if (a0 + lane < a1) {
            const int64_t a = a0 + lane;
            apply_inv(da, f0, f1, f2, c0, c1, c2);
            C[3 * a] = c0 + e0; C[3 * a + 1] = c1 + e1; C[3 * a + 2] = c2 + e2;
        }
        // more than 64 children: each lane re-reads the R_1 it wrote above
        for (int g0 = a0 + 64; g0 < a1; g0 += 64) {
            const int a = g0 + lane;
            if (a < a1) {
                apply_inv(Dinv + 6 * (int64_t)a, R[3 * a], R[3 * a + 1], R[3 * a + 2], c0, c1, c2);
                C[3 * a] = c0 + e0; C[3 * a + 1] = c1 + e1; C[3 * a + 2] = c2 + e2;
            }
        }
    }
}

// Levels >= 3 in one workgroup: R_l from the children (l = 3..L), then from the top
// down C_l = D_l^-1 R_l + C_{l+1}(parent) (l = L..3).  These levels hold a few hundred
// aggregates for a million vertices: then each thread owns one, their R and C live
// in LDS and only C_3 is written out.  Larger hierarchies loop over global memory.
constexpr int kCoarseWG = 1024;
constexpr int kCoarseLds = 1024;
__global__ void __launch_bounds__(kCoarseWG) k_ml_coarse(MLLevels lv, const int *__restrict__ cptr,
                                                         const int *__restrict__ par, const double *__restrict__ Dinv,
                                                         double *__restrict__ R, double *__restrict__ C,
                                                         const SolverScalars *scal, int check_done)
{
    __shared__ double lds[6 * kCoarseLds];
    const int L = lv.L;
    const int base = lv.off[3];
    const int M = lv.off[L + 1] - base;   // aggregates of levels >= 3
    if (M <= kCoarseWG) {
        // one aggregate per thread: its inverse block, parent and child range are
        // loaded up front (with level 3's children sums), then every phase runs in LDS
        double *Rl = lds, *Cl = lds + 3 * kCoarseLds;
        const int j = threadIdx.x;
        int l = 0, pa = 0, c0 = 0, c1 = 0;
        double d[6] = {0, 0, 0, 0, 0, 0};
        if (j < M) {
            l = 3;
            while (l < L && j >= lv.off[l + 1] - base) ++l;
            const int a = j - (lv.off[l] - base);
            const int64_t g = base + j;
#pragma unroll
            for (int k = 0; k < 6; ++k) d[k] = Dinv[6 * g + k];
            if (l < L) pa = par[g];
            c0 = cptr[lv.coff[l] + a];
            c1 = cptr[lv.coff[l] + a + 1];
        }
        // the stop flag is read with the first loads (uniform: every thread leaves
        // before the first barrier, or none does)
        if (check_done && scal->done) return;
        if (j < M) {
            if (l == 3) {
                const double *src = R + 3 * (int64_t)lv.off[2];
                double s0 = 0.0, s1 = 0.0, s2 = 0.0;
                for (int b = c0; b < c1; b += 8) {   // children in order, 8 loads in flight
                    double v[8][3];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const bool ok = b + u < c1;
                        const double *q = src + 3 * (ok ? b + u : c0);
                        v[u][0] = ok ? q[0] : 0.0;
                        v[u][1] = ok ? q[1] : 0.0;
                        v[u][2] = ok ? q[2] : 0.0;
                    }
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        s0 += v[u][0]; s1 += v[u][1]; s2 += v[u][2];
                    }
                }
                Rl[3 * j] = s0; Rl[3 * j + 1] = s1; Rl[3 * j + 2] = s2;
            }
        }
        __syncthreads();
        for (int lev = 4; lev <= L; ++lev) {
            if (j < M && l == lev) {
                const int cb = lv.off[lev - 1] - base;
                double s0 = 0.0, s1 = 0.0, s2 = 0.0;
                for (int b = c0; b < c1; ++b) {
                    s0 += Rl[3 * (cb + b)]; s1 += Rl[3 * (cb + b) + 1]; s2 += Rl[3 * (cb + b) + 2];
                }
                Rl[3 * j] = s0; Rl[3 * j + 1] = s1; Rl[3 * j + 2] = s2;
            }
            __syncthreads();
        }
        for (int lev = L; lev >= 3; --lev) {
            if (j < M && l == lev) {
                double e0, e1, e2;
                apply_inv(d, Rl[3 * j], Rl[3 * j + 1], Rl[3 * j + 2], e0, e1, e2);
                if (lev < L) {
                    const int up = lv.off[lev + 1] - base + pa;
                    e0 += Cl[3 * up]; e1 += Cl[3 * up + 1]; e2 += Cl[3 * up + 2];
                }
                Cl[3 * j] = e0; Cl[3 * j + 1] = e1; Cl[3 * j + 2] = e2;
            }
            __syncthreads();
        }
        if (j < M && l == 3) {
            double *o = C + 3 * ((int64_t)base + j);
            o[0] = Cl[3 * j]; o[1] = Cl[3 * j + 1]; o[2] = Cl[3 * j + 2];
        }
        return;
    }
    if (check_done && scal->done) return;
    const bool in_lds = M <= kCoarseLds;
    double *Rb = in_lds ? lds : R + 3 * (int64_t)base;   // level l at off[l] - off[3]
    double *Cb = in_lds ? lds + 3 * kCoarseLds : C + 3 * (int64_t)base;
    for (int l = 3; l <= L; ++l) {
        for (int A = threadIdx.x; A < lv.n[l]; A += kCoarseWG) {
            const int c0 = cptr[lv.coff[l] + A], c1 = cptr[lv.coff[l] + A + 1];
            const double *src = (l == 3) ? R + 3 * (int64_t)lv.off[2] : Rb + 3 * (lv.off[l - 1] - base);
            double s0 = 0.0, s1 = 0.0, s2 = 0.0;
            // children in order; the loads of 8 children are issued before their sums
            // (a missing child adds an exact 0)
            for (int a = c0; a < c1; a += 8) {
                double v[8][3];
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    const bool ok = a + u < c1;
                    const double *q = src + 3 * (ok ? a + u : c0);
                    v[u][0] = ok ? q[0] : 0.0;
                    v[u][1] = ok ? q[1] : 0.0;
                    v[u][2] = ok ? q[2] : 0.0;
                }
#pragma unroll
                for (int u = 0; u < 8; ++u) {
                    s0 += v[u][0]; s1 += v[u][1]; s2 += v[u][2];
                }
            }
            double *o = Rb + 3 * (lv.off[l] - base + A);
            o[0] = s0; o[1] = s1; o[2] = s2;
        }
        __syncthreads();
    }
    for (int l = L; l >= 3; --l) {
        for (int a = threadIdx.x; a < lv.n[l]; a += kCoarseWG) {
            const int g = lv.off[l] + a;
            const double *q = Rb + 3 * (g - base);
            double c0, c1, c2;
            apply_inv(Dinv + 6 * (int64_t)g, q[0], q[1], q[2], c0, c1, c2);
            if (l < L) {
                const double *up = Cb + 3 * (lv.off[l + 1] - base + par[g]);
                c0 += up[0]; c1 += up[1]; c2 += up[2];
            }
            double *o = Cb + 3 * (g - base);
            o[0] = c0; o[1] = c1; o[2] = c2;
        }
        __syncthreads();
    }
    if (in_lds)
        for (int t = threadIdx.x; t < 3 * lv.n[3]; t += kCoarseWG) C[3 * (int64_t)base + t] = Cb[t];
}

// z = D0 r + coarse correction (zero on fixed DOFs); r.z partials -> PB[parity] slot 0.
// Levels 1 and 2 are applied here, per vertex, from the restricted sums; levels >= 3
// arrive accumulated in C_3.
// FINE = false: the coarse corrections alone (z_i = C(a_i) + C_3), for the learned
// blocks' training loss, whose fine term is the network's.
template <class T, bool FINE = true, class TD = T>
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_ml_prolong(int parity, int N, MLLevels lv, const int *__restrict__ agg1, const int *__restrict__ anc3,
             const double *__restrict__ Dinv, const double *__restrict__ R, const double *__restrict__ C,
             const TD *__restrict__ D, const T *__restrict__ r, T *__restrict__ z,
             const unsigned char *__restrict__ fixmask, const SolverScalars *scal, double *partials, int check_done)
{
    __shared__ double lds[2 * (WG / 64) + 2];
    int rb, re;
    wg_rows(N, &rb, &re);
    const int i0 = rb + (int)threadIdx.x;
    // the first vertex's aggregate indices are read together with the stop flag, so
    // both corrections below load in one round
    int a = 0, u = 0;
    if (i0 < re) {
        a = agg1[i0];
        if (lv.L >= 3) u = anc3[i0];
    }
    if (check_done && scal->done) return;
    double srz = 0.0;
    for (int i = i0; i < re; i += WG) {
        if (i != i0) {
            a = agg1[i];
            if (lv.L >= 3) u = anc3[i];
        }
        double c0, c1, c2;
        if (lv.L >= 2) {
            // levels 1 and 2, combined per level-1 aggregate by the restriction
            c0 = C[3 * a]; c1 = C[3 * a + 1]; c2 = C[3 * a + 2];
            if (lv.L >= 3) {
                const double *up = C + 3 * (int64_t)u;
                c0 += up[0]; c1 += up[1]; c2 += up[2];
            }
        } else {
            apply_inv(Dinv + 6 * (int64_t)a, R[3 * a], R[3 * a + 1], R[3 * a + 2], c0, c1, c2);
        }
        const T r0 = r[3 * i], r1 = r[3 * i + 1], r2 = r[3 * i + 2];
        double g0 = 0.0, g1 = 0.0, g2 = 0.0;
        if constexpr (FINE) {
            TD f0, f1, f2;
            eolc::apply_sym3<TD>(D + i, N, (TD)r0, (TD)r1, (TD)r2, f0, f1, f2);
            g0 = (double)f0; g1 = (double)f1; g2 = (double)f2;
        }
        const unsigned fm = fixmask[i];
        const T z0 = (fm & 1u) ? (T)0 : (T)(g0 + c0);
        const T z1 = (fm & 2u) ? (T)0 : (T)(g1 + c1);
        const T z2 = (fm & 4u) ? (T)0 : (T)(g2 + c2);
        z[3 * i] = z0; z[3 * i + 1] = z1; z[3 * i + 2] = z2;
        srz += ((double)r0 * (double)z0 + (double)r1 * (double)z1) + (double)r2 * (double)z2;
    }
    double v[1] = {srz};
    write_partials<1>(v, PB(partials, parity), lds);
}

inline int blocks_for(int64_t n, int t = 256) { return (int)((n + t - 1) / t); }

}  // namespace

namespace eolc {

hipError_t ml_setup(eolc_ctx *c)
{
    const MLState &m = c->ml;
    const MLLevels &lv = m.lv;
    if (lv.L <= 0) return hipSuccess;
    const Topology &t = c->topo;
    hipLaunchKernelGGL(k_ml_diag1, dim3(blocks_for(lv.n[1])), dim3(256), 0, c->stream, lv, m.hord, m.ptr1, m.gpos,
                       c->slice_off, c->slen, c->scol, m.lam, c->A64, t.nbs, c->fixmask, m.V, m.Dinv);
    for (int l = 2; l <= lv.L; ++l)
        hipLaunchKernelGGL(k_ml_diagup, dim3(blocks_for(lv.n[l])), dim3(256), 0, c->stream, l, lv, m.cptr, m.V,
                           m.Dinv);
    return hipGetLastError();
}

hipError_t ml_apply(eolc_ctx *c, bool mixed, int parity, int check_done)
{
    const MLState &m = c->ml;
    const MLLevels &lv = m.lv;
    const int N = c->topo.N;
    const int G = pcg_grid(N);
    const int waves = lv.L >= 2 ? lv.n[2] : (lv.n[1] + 63) / 64;
    const int rblocks = blocks_for((int64_t)64 * waves, kRestrictWG);
    if (mixed) {
        hipLaunchKernelGGL(k_ml_restrict<float>, dim3(rblocks), dim3(kRestrictWG), 0, c->stream, lv, m.hord, m.ptr1, m.cptr,
                           m.kptr2, c->r32, m.R, m.Dinv, m.C, c->scal, check_done);
    } else {
        hipLaunchKernelGGL(k_ml_restrict<double>, dim3(rblocks), dim3(kRestrictWG), 0, c->stream, lv, m.hord, m.ptr1, m.cptr,
                           m.kptr2, c->r64, m.R, m.Dinv, m.C, c->scal, check_done);
    }
    if (lv.L >= 3)
        hipLaunchKernelGGL(k_ml_coarse, dim3(1), dim3(kCoarseWG), 0, c->stream, lv, m.cptr, m.par, m.Dinv, m.R, m.C,
                           c->scal, check_done);
    if (mixed) {
        hipLaunchKernelGGL(k_ml_prolong<float>, dim3(G), dim3(WG), 0, c->stream, parity, N, lv, m.agg1, m.anc3, m.Dinv,
                           m.R, m.C, c->D32, c->r32, c->z32, c->fixmask, c->scal, c->partials, check_done);
    } else {
        hipLaunchKernelGGL(k_ml_prolong<double>, dim3(G), dim3(WG), 0, c->stream, parity, N, lv, m.agg1, m.anc3,
                           m.Dinv, m.R, m.C, c->D64, c->r64, c->z64, c->fixmask, c->scal, c->partials, check_done);
    }
    return hipGetLastError();
}

// z64 = the coarse corrections of r64 alone (no fine blocks), FP64, fixed DOFs zero.
hipError_t ml_coarse64(eolc_ctx *c)
{
    const MLState &m = c->ml;
    const MLLevels &lv = m.lv;
    const int N = c->topo.N;
    const int G = pcg_grid(N);
    if (lv.L <= 0) return hipMemsetAsync(c->z64, 0, sizeof(double) * 3 * (size_t)N, c->stream);
    const int waves = lv.L >= 2 ? lv.n[2] : (lv.n[1] + 63) / 64;
    const int rblocks = blocks_for((int64_t)64 * waves, kRestrictWG);
    hipLaunchKernelGGL(k_ml_restrict<double>, dim3(rblocks), dim3(kRestrictWG), 0, c->stream, lv, m.hord, m.ptr1, m.cptr,
                       m.kptr2, c->r64, m.R, m.Dinv, m.C, c->scal, 0);
    if (lv.L >= 3)
        hipLaunchKernelGGL(k_ml_coarse, dim3(1), dim3(kCoarseWG), 0, c->stream, lv, m.cptr, m.par, m.Dinv, m.R, m.C,
                           c->scal, 0);
    hipLaunchKernelGGL((k_ml_prolong<double, false>), dim3(G), dim3(WG), 0, c->stream, 0, N, lv, m.agg1, m.anc3, m.Dinv,
                       m.R, m.C, c->D64, c->r64, c->z64, c->fixmask, c->scal, c->partials, 0);
    return hipGetLastError();
}

}  // namespace eolc
