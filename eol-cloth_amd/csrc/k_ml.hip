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
//
// Linear coarse fields (single-cloth solves, levels >= kMLLinFrom = 3): there Z_l
// carries, per aggregate A and component, the basis phi = (1, (X - c_A) / s_A) of
// the material coordinates X (c_A the centroid, s_A the RMS radius of A's members),
// so a coarse correction is a field linear on each aggregate instead of a constant,
// and D_l is the 9 x 9 diagonal block of Z_l^T A Z_l, applied through the LDL^T
// factors of its unit-diagonal scaling, which leave out a basis function whose
// scaled pivot falls to 1e-10 (fact9_psd).  The smooth bending modes of a sheet are linear to
// first order over an aggregate; at 512^2 this cut the PCG iterations by a quarter
// (826 -> 617 at step 0, 890 -> 696 at step 6 of the benchmark trajectory).  The
// restriction forms the first moments sum (phi_1, phi_2) r of each level-2 aggregate,
// from which the coarse kernel gets every level >= 3 by the exact change of
// centroid and radius; the prolongation evaluates the level-3 field at each vertex.
// The batched solve keeps the translations on every level.
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

// Upper triangle of a symmetric 9 x 9 block, row by row: 45 entries.
__device__ __forceinline__ constexpr int pk9(int p, int q) { return p * 9 - p * (p - 1) / 2 + (q - p); }
// Strictly lower triangle of a 9 x 9 factor, row by row: 36 entries.
__device__ __forceinline__ constexpr int lo9(int i, int j) { return i * (i - 1) / 2 + j; }

constexpr double kPivotTol = 1e-10;
constexpr int kFact9 = 54;   // per aggregate: 36 factor entries, 9 inverse pivots, 9 scales

// Factorisation of a symmetric positive semi-definite 9 x 9 block G (packed) for
// applying its inverse: with the scales s_p = G_pp^-1/2 the scaled block s G s has
// a unit diagonal; LDL^T of it in index order leaves out a basis function whose
// pivot is at or below kPivotTol (its share of the block not spanned by the earlier
// ones), or whose diagonal is not positive, and inverts the rest.  Stored: the unit
// lower factor (its columns of left-out functions zero), 1/d (0 when left out) and
// s.  The scaling keeps a block whose stretch modes are 1e10 times stiffer than its
// translations (a collapsed element inside the aggregate) from losing the
// translations, and applying the factors keeps the operator positive where an
// explicit inverse of such a block would not be.
__device__ void fact9_psd(const double *__restrict__ G, double *__restrict__ F)
{
    double Lm[9][9], d[9], sc[9];
    bool kept[9];
    for (int p = 0; p < 9; ++p) {
        const double g = G[pk9(p, p)];
        sc[p] = (g > 0.0 && isfinite(g)) ? 1.0 / sqrt(g) : 0.0;
    }
    for (int k = 0; k < 9; ++k) {
        double dk = sc[k] * sc[k] * G[pk9(k, k)];
        for (int j = 0; j < k; ++j)
            if (kept[j]) dk -= Lm[k][j] * Lm[k][j] * d[j];
        kept[k] = sc[k] > 0.0 && dk > kPivotTol && isfinite(dk);
        for (int i = 0; i < 9; ++i) Lm[i][k] = 0.0;
        d[k] = 0.0;
        if (!kept[k]) continue;
        d[k] = dk;
        for (int i = k + 1; i < 9; ++i) {
            double v = sc[k] * sc[i] * G[pk9(k, i)];
            for (int j = 0; j < k; ++j)
                if (kept[j]) v -= Lm[i][j] * Lm[k][j] * d[j];
            Lm[i][k] = v / dk;
        }
    }
    for (int i = 1; i < 9; ++i)
        for (int j = 0; j < i; ++j) F[lo9(i, j)] = Lm[i][j];
    for (int p = 0; p < 9; ++p) {
        F[36 + p] = kept[p] ? 1.0 / d[p] : 0.0;
        F[45 + p] = sc[p];
    }
}

// out = G^+ r from fact9_psd's factors: s L^-T D^-1 L^-1 s r.
__device__ __forceinline__ void apply9(const double *__restrict__ F, const double *r, double *out)
{
    double y[9];
#pragma unroll
    for (int i = 0; i < 9; ++i) {
        double v = F[45 + i] * r[i];
#pragma unroll
        for (int j = 0; j < i; ++j) v = fma(-F[lo9(i, j)], y[j], v);
        y[i] = v;
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) y[i] *= F[36 + i];
#pragma unroll
    for (int i = 8; i >= 0; --i) {
        double v = y[i];
#pragma unroll
        for (int j = i + 1; j < 9; ++j) v = fma(-F[lo9(j, i)], y[j], v);
        y[i] = v;
    }
#pragma unroll
    for (int i = 0; i < 9; ++i) out[i] = F[45 + i] * y[i];
}

// One workgroup per set-up chunk (level l, aggregate g, hord positions [k0, k1)):
// sum over the chunk's rows i and their entries (i, j) inside g (first common level
// <= l) of phi(i) phi(j)^T (x) A_ij, fixed DOFs masked, in the aggregate's basis;
// the upper triangle (45 sums) in a fixed order per thread, then over the workgroup.
constexpr int kGalWG = 256;
__global__ void __launch_bounds__(kGalWG) k_ml_gal9_part(const int *__restrict__ gchunk, const int *__restrict__ hord,
                                                         const int *__restrict__ gpos,
                                                         const int *__restrict__ slice_off,
                                                         const int *__restrict__ slen, const int *__restrict__ scol,
                                                         const unsigned char *__restrict__ lam,
                                                         const double *__restrict__ A, int64_t nbs,
                                                         const unsigned char *__restrict__ fixmask,
                                                         const double *__restrict__ X, const double *__restrict__ geo,
                                                         double *__restrict__ gpart)
{
    __shared__ double red[kGalWG / 64][45];
    const int ch = blockIdx.x;
    const int l = gchunk[4 * ch], g = gchunk[4 * ch + 1], k0 = gchunk[4 * ch + 2], k1 = gchunk[4 * ch + 3];
    const double cu = geo[4 * (int64_t)g], cv = geo[4 * (int64_t)g + 1], is = geo[4 * (int64_t)g + 3];
    double acc[45];
#pragma unroll
    for (int p = 0; p < 45; ++p) acc[p] = 0.0;
    for (int k = k0 + (int)threadIdx.x; k < k1; k += kGalWG) {
        const int i = hord[k];
        const int gp = gpos[i];
        const int len = slen[gp];
        const int base = slice_off[gp / eolc::kSlice] + (gp % eolc::kSlice);
        const unsigned fi = fixmask[i];
        const double pi[3] = {1.0, (X[2 * (int64_t)i] - cu) * is, (X[2 * (int64_t)i + 1] - cv) * is};
        for (int kk = 0; kk < len; ++kk) {
            const int e = base + kk * eolc::kSlice;
            if ((int)lam[e] > l) continue;
            const int j = scol[e];
            const unsigned fj = fixmask[j];
            const double pj[3] = {1.0, (X[2 * (int64_t)j] - cu) * is, (X[2 * (int64_t)j + 1] - cv) * is};
            double b[9];
#pragma unroll
            for (int c = 0; c < 9; ++c) {
                const int r = c / 3, t = c % 3;
                b[c] = (((fi >> r) | (fj >> t)) & 1u) ? 0.0 : A[(int64_t)c * nbs + e];
            }
#pragma unroll
            for (int p = 0; p < 9; ++p)
#pragma unroll
                for (int q = p; q < 9; ++q) acc[pk9(p, q)] = fma(pi[p / 3] * pj[q / 3], b[(p % 3) * 3 + q % 3], acc[pk9(p, q)]);
        }
    }
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int p = 0; p < 45; ++p) {
        const double v = wave_sum(acc[p]);
        if (lane == 0) red[wid][p] = v;
    }
    __syncthreads();
    if (threadIdx.x < 45) {
        double v = 0.0;
#pragma unroll
        for (int w = 0; w < kGalWG / 64; ++w) v += red[w][threadIdx.x];
        gpart[45 * (int64_t)ch + threadIdx.x] = v;
    }
}

// One thread per aggregate of levels >= 3: its chunks' sums in order, then the
// block's factors.
__global__ void __launch_bounds__(64) k_ml_gal9_fin(int n_hi, const int *__restrict__ gcp,
                                                    const double *__restrict__ gpart, double *__restrict__ D9)
{
    const int h = blockIdx.x * 64 + threadIdx.x;
    if (h >= n_hi) return;
    double G[45];
    for (int p = 0; p < 45; ++p) G[p] = 0.0;
    for (int ch = gcp[h]; ch < gcp[h + 1]; ++ch)
        for (int p = 0; p < 45; ++p) G[p] += gpart[45 * (int64_t)ch + p];
    fact9_psd(G, D9 + kFact9 * (int64_t)h);
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

__device__ __forceinline__ void lift9(const double *__restrict__ q, double al, double bu, double bv, double *sm)
{
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        sm[a] += q[a];
        sm[3 + a] += fma(al, q[3 + a], bu * q[a]);
        sm[6 + a] += fma(al, q[6 + a], bv * q[a]);
    }
}
__device__ __forceinline__ void lower9(const double *__restrict__ P, double al, double bu, double bv, double *cf)
{
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        cf[a] += fma(bv, P[6 + a], fma(bu, P[3 + a], P[a]));
        cf[3 + a] = fma(al, P[3 + a], cf[3 + a]);
        cf[6 + a] = fma(al, P[6 + a], cf[6 + a]);
    }
}

// The restriction of one level-2 aggregate w by one wavefront (two levels or more),
// or of the 64 level-1 aggregates from w * 64 (one level).  Writes R_1 of its
// children, R_2(w) and C of its children; returns R_2(w) in every lane and, with
// feat2 (the members' level-2 coordinate features in hord order), the first moments
// sum phi_1 r and sum phi_2 r in lane 0.  Without the moments (LIN false) the
// inverse blocks load before the member loop, hiding their latency; with them they
// load after it, which keeps the loop's registers at six waves per SIMD.
template <class T, bool LIN>
__device__ __forceinline__ void restrict_wave(const MLLevels &lv, int w, const int *__restrict__ hord,
                                              const int *__restrict__ ptr1, const int *__restrict__ cptr,
                                              const int *__restrict__ kptr2, const T *__restrict__ r,
                                              double *__restrict__ R, const double *__restrict__ Dinv,
                                              double *__restrict__ C, const double *__restrict__ feat2,
                                              double *__restrict__ buf, double &s0, double &s1, double &s2,
                                              double (&mom)[6], const SolverScalars *scal, int check_done)
{
    const int lane = threadIdx.x % 64;
    const bool two = lv.L >= 2;
    int a0, a1, ka = 0, kz = 0;
    if (two) {
        a0 = cptr[lv.coff[2] + w];
        a1 = cptr[lv.coff[2] + w + 1];
        ka = kptr2[w];
        kz = kptr2[w + 1];
    } else {
        a0 = w * 64;
        a1 = min(a0 + 64, lv.n[1]);
    }
    // the stop flag is read together with the ranges (one latency round; a stopped
    // launch only reads these words)
    if (check_done && scal->done) return;
    // inverse blocks of the level-2 aggregate and of this lane's first child: they do
    // not depend on r, so they load with the member indices
    double dw[6] = {0, 0, 0, 0, 0, 0}, da[6] = {0, 0, 0, 0, 0, 0};
    auto load_inv = [&]() {
        const int64_t gw = lv.off[2] + w;
#pragma unroll
        for (int k = 0; k < 6; ++k) dw[k] = Dinv[6 * gw + k];
        if (a0 + lane < a1)
#pragma unroll
            for (int k = 0; k < 6; ++k) da[k] = Dinv[6 * (int64_t)(a0 + lane) + k];
    };
    if (two && !LIN) load_inv();
    s0 = s1 = s2 = 0.0;
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;   // R_1 of this lane's first child
    double mu0 = 0.0, mu1 = 0.0, mu2 = 0.0, mv0 = 0.0, mv1 = 0.0, mv2 = 0.0;
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
                double u = 0.0, v = 0.0;   // features first: they do not wait for hord
                if (LIN) {
                    u = feat2[2 * (int64_t)k];
                    v = feat2[2 * (int64_t)k + 1];
                }
                const int i = hord[k];
                // the compiler otherwise issues the feature load after the gather of r
                // below, a third dependent memory round per 64 members
                if (LIN) asm volatile("" ::: "memory");
                const double q0 = (double)r[3 * i], q1 = (double)r[3 * i + 1], q2 = (double)r[3 * i + 2];
                buf[3 * lane] = q0;
                buf[3 * lane + 1] = q1;
                buf[3 * lane + 2] = q2;
                if (LIN) {
                    mu0 = fma(u, q0, mu0); mu1 = fma(u, q1, mu1); mu2 = fma(u, q2, mu2);
                    mv0 = fma(v, q0, mv0); mv1 = fma(v, q1, mv1); mv2 = fma(v, q2, mv2);
                }
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
    if (!two) return;
    if (LIN) load_inv();
    // wave_sum leaves the total in lane 0; every lane needs it below
    s0 = __shfl(wave_sum(s0), 0, 64);
    s1 = __shfl(wave_sum(s1), 0, 64);
    s2 = __shfl(wave_sum(s2), 0, 64);
    if (lane == 0) {
        double *o = R + 3 * (lv.off[2] + (int64_t)w);
        o[0] = s0; o[1] = s1; o[2] = s2;
    }
    if (LIN) {
        mom[0] = wave_sum(mu0); mom[1] = wave_sum(mu1); mom[2] = wave_sum(mu2);
        mom[3] = wave_sum(mv0); mom[4] = wave_sum(mv1); mom[5] = wave_sum(mv2);
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

// Translations only: one wavefront per level-2 aggregate (or per 64 level-1 ones).
template <class T>
__global__ void __launch_bounds__(kRestrictWG) k_ml_restrict(MLLevels lv, const int *__restrict__ hord,
                                                             const int *__restrict__ ptr1, const int *__restrict__ cptr,
                                                             const int *__restrict__ kptr2,
                                                             const T *__restrict__ r, double *__restrict__ R,
                                                             const double *__restrict__ Dinv, double *__restrict__ C,
                                                             const SolverScalars *scal, int check_done)
{
    __shared__ double stage[kRestrictWG / 64][3 * 64];
    const int w = (blockIdx.x * blockDim.x + threadIdx.x) / 64;   // uniform in the wavefront
    if (lv.L >= 2 ? w >= lv.n[2] : w * 64 >= lv.n[1]) return;
    double s0, s1, s2, mom[6];
    restrict_wave<T, false>(lv, w, hord, ptr1, cptr, kptr2, r, R, Dinv, C, nullptr, stage[threadIdx.x / 64], s0, s1, s2,
                     mom, scal, check_done);
}

// Linear coarse fields: kRestrict3Parts workgroups per level-3 aggregate U, whose
// kRestrict3Parts x (kRestrict3WG / 64) wavefronts take U's level-2 children in turn
// (wavefront g of U takes children c0 + g, c0 + g + 16, ...).  Each child's R_2 and
// first moments are re-expressed on U's basis, (R_2, alpha M_1 + beta_u R_2, alpha
// M_2 + beta_v R_2) with alpha = s_w / s_U and beta = (c_w - c_U) / s_U, summed per
// wavefront in child order and over the workgroup's wavefronts in order: one partial
// R9 per workgroup, which the coarse kernel adds in order.  At 1024^2 the kernel takes
// 26 us against 14 us for the translations' restriction.  It ran at 31 us with the
// inverse blocks loaded before the member loop (101 VGPRs, four waves per SIMD) and
// at 28 us with the feature load issued after the gather of r; 512-thread workgroups
// (32 wavefronts per level-3 aggregate) took 36 us, and loading the next 64 members'
// indices and features during the gather gained nothing.
constexpr int kRestrict3WG = 256;
using eolc::kRestrict3Parts;
template <class T>
__global__ void __launch_bounds__(kRestrict3WG) k_ml_restrict3(MLLevels lv, const int *__restrict__ hord,
                                                               const int *__restrict__ ptr1,
                                                               const int *__restrict__ cptr,
                                                               const int *__restrict__ kptr2, const T *__restrict__ r,
                                                               double *__restrict__ R, const double *__restrict__ Dinv,
                                                               double *__restrict__ C,
                                                               const double *__restrict__ feat2,
                                                               const double *__restrict__ geo, double *__restrict__ R9p,
                                                               const SolverScalars *scal, int check_done)
{
    constexpr int NW = kRestrict3WG / 64;
    __shared__ double stage[NW][3 * 64];
    __shared__ double red[NW][9];
    const int U = blockIdx.x / kRestrict3Parts, part = blockIdx.x % kRestrict3Parts;
    const int lane = threadIdx.x % 64, wid = threadIdx.x / 64;
    const int c0 = cptr[lv.coff[3] + U], c1 = cptr[lv.coff[3] + U + 1];
    const int64_t gU = lv.off[3] + U;
    const double cu = geo[4 * gU], cv = geo[4 * gU + 1], is = geo[4 * gU + 3];
    if (check_done && scal->done) return;
    // lane 0 keeps the wavefront's running sum in LDS
    if (lane == 0)
#pragma unroll
        for (int q = 0; q < 9; ++q) red[wid][q] = 0.0;
    for (int w = c0 + part * NW + wid; w < c1; w += NW * kRestrict3Parts) {
        double s0, s1, s2, mom[6];
        restrict_wave<T, true>(lv, w, hord, ptr1, cptr, kptr2, r, R, Dinv, C, feat2, stage[wid], s0, s1, s2, mom, scal, 0);
        if (lane == 0) {
            const int64_t gw = lv.off[2] + w;
            const double al = geo[4 * gw + 2] * is, bu = (geo[4 * gw] - cu) * is, bv = (geo[4 * gw + 1] - cv) * is;
            double t[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
            const double q[9] = {s0, s1, s2, mom[0], mom[1], mom[2], mom[3], mom[4], mom[5]};
            lift9(q, al, bu, bv, t);
#pragma unroll
            for (int k = 0; k < 9; ++k) red[wid][k] += t[k];
        }
    }
    __syncthreads();
    if (threadIdx.x < 9) {
        double v = 0.0;
#pragma unroll
        for (int k = 0; k < NW; ++k) v += red[k][threadIdx.x];
        R9p[9 * (int64_t)blockIdx.x + threadIdx.x] = v;
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
        int l = 0, pa = 0, c0 = 0, c1 = 0, offm1 = 0, offp1 = 0;
        double d[6] = {0, 0, 0, 0, 0, 0};
        if (j < M) {
            // the level and its neighbours' offsets from constant indices into the
            // arguments (scalar loads; a variable index costs a vector load per use)
            l = 3;
            int offl = lv.off[3], coffl = lv.coff[3];
            offm1 = lv.off[2];
            offp1 = lv.off[4];
#pragma unroll
            for (int k = 4; k <= eolc::kMLMaxLevels; ++k)
                if (k <= L && j >= lv.off[k] - base) {   // offsets increase with the level
                    l = k;
                    offl = lv.off[k];
                    coffl = lv.coff[k];
                    offm1 = lv.off[k - 1];
                    offp1 = lv.off[k + 1];
                }
            const int a = j - (offl - base);
            const int64_t g = base + j;
#pragma unroll
            for (int k = 0; k < 6; ++k) d[k] = Dinv[6 * g + k];
            if (l < L) pa = par[g];
            c0 = cptr[coffl + a];
            c1 = cptr[coffl + a + 1];
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
                const int cb = offm1 - base;   // level lev - 1
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
                    const int up = offp1 - base + pa;   // the parent on level lev + 1
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

// Levels >= 3 with linear fields, in one workgroup, from the level-3 sums R9_3 of
// the restriction: R9_l (l > 3) = sum over the level-(l-1) children c of (R9_c[0],
// alpha R9_c[1] + beta_u R9_c[0], alpha R9_c[2] + beta_v R9_c[0]) with alpha = s_c /
// s_A, beta = (c_c - c_A) / s_A (the child's basis on the parent's; R9_c[m] the
// three components of feature m).  Then from the top down C9_l = D9_l^+ R9_l plus the
// parent's field re-expressed on the aggregate's basis (the same alpha, beta); the
// level-3 fields go to C9.
// One aggregate per thread (the linear fields apply when these levels hold at most
// kCoarse9WG aggregates, ml_prepare_linear): every load (the level-3 sums, geometry,
// parents, child ranges and the thread's own factors) is issued in one round, then
// the levels run in LDS and registers.
using eolc::kCoarse9WG;
__global__ void __launch_bounds__(kCoarse9WG) k_ml_coarse9(MLLevels lv, const int *__restrict__ cptr,
                                                           const int *__restrict__ par, const double *__restrict__ geo,
                                                           const double *__restrict__ D9, const double *__restrict__ R9p,
                                                           double *__restrict__ C9, const SolverScalars *scal,
                                                           int check_done)
{
    __shared__ double sR[9 * kCoarse9WG], sC[9 * kCoarse9WG], sg[4 * kCoarse9WG];
    __shared__ int sp[kCoarse9WG];
    const int L = lv.L, base = lv.off[3];
    const int M = lv.off[L + 1] - base, n3 = lv.n[3];
    const int h = threadIdx.x;   // this thread's aggregate (levels >= 3 concatenated)
    // Every load is issued before the first wait: the level-3 partial sums first (they
    // depend on h only), then the aggregate's factors, geometry, parent and child
    // range.  The level and the offsets of its neighbour levels come from the uniform
    // offsets with constant indices (scalar loads; a variable index reads the argument
    // block with a vector load, one more memory round per level), and the LDS stores
    // follow the loads.
    double r3[kRestrict3Parts][9];
    if (h < n3)
#pragma unroll
        for (int q = 0; q < kRestrict3Parts; ++q)
#pragma unroll
            for (int k = 0; k < 9; ++k) r3[q][k] = R9p[9 * ((int64_t)h * kRestrict3Parts + q) + k];
    int l = 0, A = 0, c0 = 0, c1 = 0, pp = 0, offm1 = 0, offp1 = 0;
    double F[kFact9], g4[4];
    if (h < M) {
        l = 3;
        int offl = lv.off[3], coffl = lv.coff[3];
        offm1 = lv.off[2];
        offp1 = lv.off[4];
#pragma unroll
        for (int k = 4; k <= eolc::kMLMaxLevels; ++k)
            if (k <= L && h >= lv.off[k] - base) {   // offsets increase with the level
                l = k;
                offl = lv.off[k];
                coffl = lv.coff[k];
                offm1 = lv.off[k - 1];
                offp1 = lv.off[k + 1];
            }
        A = h - (offl - base);
        if (l > 3) {
            c0 = cptr[coffl + A];
            c1 = cptr[coffl + A + 1];
        }
#pragma unroll
        for (int k = 0; k < kFact9; ++k) F[k] = D9[kFact9 * (int64_t)h + k];
#pragma unroll
        for (int k = 0; k < 4; ++k) g4[k] = geo[4 * (int64_t)(base + h) + k];
        pp = l < L ? par[base + h] : 0;
    }
    if (check_done && scal->done) return;
    if (h < M) {
#pragma unroll
        for (int k = 0; k < 4; ++k) sg[4 * h + k] = g4[k];
        sp[h] = pp;
    }
    // the level-3 sums: the restriction's partial sums of each aggregate, in order
    if (h < n3)
#pragma unroll
        for (int k = 0; k < 9; ++k) {
            double v = r3[0][k];
#pragma unroll
            for (int q = 1; q < kRestrict3Parts; ++q) v += r3[q][k];
            sR[9 * h + k] = v;
        }
    __syncthreads();
    for (int lev = 4; lev <= L; ++lev) {
        if (h < M && l == lev) {
            const double cu = sg[4 * h], cv = sg[4 * h + 1], is = sg[4 * h + 3];
            double sm[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) sm[q] = 0.0;
            const int cb = offm1 - base;   // level lev - 1
            for (int ch = c0; ch < c1; ++ch) {
                const int hc = cb + ch;
                lift9(sR + 9 * hc, sg[4 * hc + 2] * is, (sg[4 * hc] - cu) * is, (sg[4 * hc + 1] - cv) * is, sm);
            }
#pragma unroll
            for (int q = 0; q < 9; ++q) sR[9 * h + q] = sm[q];
        }
        __syncthreads();
    }
    for (int lev = L; lev >= 3; --lev) {
        if (h < M && l == lev) {
            double rv[9], cf[9];
#pragma unroll
            for (int q = 0; q < 9; ++q) rv[q] = sR[9 * h + q];
            apply9(F, rv, cf);
            if (lev < L) {
                const int hp = offp1 - base + sp[h];   // the parent on level lev + 1
                const double ip = sg[4 * hp + 3];
                lower9(sC + 9 * hp, sg[4 * h + 2] * ip, (sg[4 * h] - sg[4 * hp]) * ip, (sg[4 * h + 1] - sg[4 * hp + 1]) * ip,
                       cf);
            }
#pragma unroll
            for (int q = 0; q < 9; ++q) sC[9 * h + q] = cf[q];
        }
        __syncthreads();
    }
    for (int t = threadIdx.x; t < 9 * n3; t += kCoarse9WG) C9[t] = sC[t];
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
             const unsigned char *__restrict__ fixmask, const SolverScalars *scal, double *partials, int check_done,
             const double *__restrict__ feat3, const double *__restrict__ C9)
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
                if (C9) {   // the level-3 aggregate's linear field at this vertex
                    const double *P = C9 + 9 * (int64_t)(u - lv.off[3]);
                    const double gu = feat3[2 * (int64_t)i], gv = feat3[2 * (int64_t)i + 1];
                    c0 += fma(gv, P[6], fma(gu, P[3], P[0]));
                    c1 += fma(gv, P[7], fma(gu, P[4], P[1]));
                    c2 += fma(gv, P[8], fma(gu, P[5], P[2]));
                } else {
                    const double *up = C + 3 * (int64_t)u;
                    c0 += up[0]; c1 += up[1]; c2 += up[2];
                }
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

hipError_t ml_setup(eolc_ctx *c, bool linear)
{
    MLState &m = c->ml;
    const MLLevels &lv = m.lv;
    m.lin_on = false;
    if (lv.L <= 0) return hipSuccess;
    const Topology &t = c->topo;
    hipLaunchKernelGGL(k_ml_diag1, dim3(blocks_for(lv.n[1])), dim3(256), 0, c->stream, lv, m.hord, m.ptr1, m.gpos,
                       c->slice_off, c->slen, c->scol, m.lam, c->A64, t.nbs, c->fixmask, m.V, m.Dinv);
    for (int l = 2; l <= lv.L; ++l)
        hipLaunchKernelGGL(k_ml_diagup, dim3(blocks_for(lv.n[l])), dim3(256), 0, c->stream, l, lv, m.cptr, m.V,
                           m.Dinv);
    if (linear && m.lin_ok) {
        const int n_hi = lv.off[lv.L + 1] - lv.off[kMLLinFrom];
        hipLaunchKernelGGL(k_ml_gal9_part, dim3(m.nchunks), dim3(kGalWG), 0, c->stream, m.gchunk, m.hord, m.gpos,
                           c->slice_off, c->slen, c->scol, m.lam, c->A64, t.nbs, c->fixmask, c->X, m.geo, m.gpart);
        hipLaunchKernelGGL(k_ml_gal9_fin, dim3(blocks_for(n_hi, 64)), dim3(64), 0, c->stream, n_hi, m.gcp, m.gpart,
                           m.D9);
        m.lin_on = true;
    }
    return hipGetLastError();
}

// The restriction, the coarse levels and the prolongation's coarse term, with the
// linear fields when the last ml_setup formed them.
static void ml_coarse_chain(eolc_ctx *c, bool mixed, int check_done)
{
    const MLState &m = c->ml;
    const MLLevels &lv = m.lv;
    const bool lin = m.lin_on && lv.L >= kMLLinFrom;
    if (lin) {
        if (mixed)
            hipLaunchKernelGGL(k_ml_restrict3<float>, dim3(lv.n[3] * kRestrict3Parts), dim3(kRestrict3WG), 0,
                               c->stream, lv, m.hord, m.ptr1, m.cptr, m.kptr2, c->r32, m.R, m.Dinv, m.C, m.feat2,
                               m.geo, m.R9p, c->scal, check_done);
        else
            hipLaunchKernelGGL(k_ml_restrict3<double>, dim3(lv.n[3] * kRestrict3Parts), dim3(kRestrict3WG), 0,
                               c->stream, lv, m.hord, m.ptr1, m.cptr, m.kptr2, c->r64, m.R, m.Dinv, m.C, m.feat2,
                               m.geo, m.R9p, c->scal, check_done);
        hipLaunchKernelGGL(k_ml_coarse9, dim3(1), dim3(kCoarse9WG), 0, c->stream, lv, m.cptr, m.par, m.geo, m.D9,
                           m.R9p, m.C9, c->scal, check_done);
        return;
    }
    const int waves = lv.L >= 2 ? lv.n[2] : (lv.n[1] + 63) / 64;
    const int rblocks = blocks_for((int64_t)64 * waves, kRestrictWG);
    if (mixed)
        hipLaunchKernelGGL(k_ml_restrict<float>, dim3(rblocks), dim3(kRestrictWG), 0, c->stream, lv, m.hord, m.ptr1,
                           m.cptr, m.kptr2, c->r32, m.R, m.Dinv, m.C, c->scal, check_done);
    else
        hipLaunchKernelGGL(k_ml_restrict<double>, dim3(rblocks), dim3(kRestrictWG), 0, c->stream, lv, m.hord, m.ptr1,
                           m.cptr, m.kptr2, c->r64, m.R, m.Dinv, m.C, c->scal, check_done);
    if (lv.L >= 3)
        hipLaunchKernelGGL(k_ml_coarse, dim3(1), dim3(kCoarseWG), 0, c->stream, lv, m.cptr, m.par, m.Dinv, m.R, m.C,
                           c->scal, check_done);
}

hipError_t ml_apply(eolc_ctx *c, bool mixed, int parity, int check_done)
{
    const MLState &m = c->ml;
    const MLLevels &lv = m.lv;
    const int N = c->topo.N;
    const int G = pcg_grid(N);
    ml_coarse_chain(c, mixed, check_done);
    const double *c9 = (m.lin_on && lv.L >= kMLLinFrom) ? m.C9 : nullptr;
    if (mixed) {
        hipLaunchKernelGGL(k_ml_prolong<float>, dim3(G), dim3(WG), 0, c->stream, parity, N, lv, m.agg1, m.anc3, m.Dinv,
                           m.R, m.C, c->D32, c->r32, c->z32, c->fixmask, c->scal, c->partials, check_done, m.feat3,
                           c9);
    } else {
        hipLaunchKernelGGL(k_ml_prolong<double>, dim3(G), dim3(WG), 0, c->stream, parity, N, lv, m.agg1, m.anc3,
                           m.Dinv, m.R, m.C, c->D64, c->r64, c->z64, c->fixmask, c->scal, c->partials, check_done,
                           m.feat3, c9);
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
    ml_coarse_chain(c, false, 0);
    const double *c9 = (m.lin_on && lv.L >= kMLLinFrom) ? m.C9 : nullptr;
    hipLaunchKernelGGL((k_ml_prolong<double, false>), dim3(G), dim3(WG), 0, c->stream, 0, N, lv, m.agg1, m.anc3, m.Dinv,
                       m.R, m.C, c->D64, c->r64, c->z64, c->fixmask, c->scal, c->partials, 0, m.feat3, c9);
    return hipGetLastError();
}

}  // namespace eolc
