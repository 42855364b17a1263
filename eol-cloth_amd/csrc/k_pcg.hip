// Block-Jacobi PCG on the 3x3 BSR matrix MDK (gfx950).
//
// Solver "none" of GeneralizedSolver::velocitySolve (GeneralizedSolver.cpp:107-172)
// as the Dirichlet-reduced solve of MDK v = M v_old + h f: fixed DOFs are masked
// out of every matrix-vector product and carry their prescribed values.
//
// Iteration k = two kernels:
//   k_pcg_spmv<k>   : rz_k, rr_k <- sum of PB[k&1]; stop test; beta = rz_k / rz_{k-1};
//                     p_new = z + beta p_old (own rows and, on the fly, neighbours);
//                     q = A p_new; partial p_new.q -> PA.
//   k_pcg_update<k> : pq <- sum of PA; alpha = rz_k / pq; x += alpha p_new,
//                     r -= alpha q, z = D^-1 r; partials r.z, r.r -> PB[(k+1)&1].
// With the multilevel preconditioner the update forms x, r and r.r only, and the
// restriction, coarse-level and prolongation kernels of k_ml.hip form z and r.z.
// There is no grid-wide fence inside a kernel: every workgroup re-reduces the
// previous kernel's per-workgroup partials itself, in one fixed order, so all
// workgroups reach bitwise-identical scalars and identical stop decisions.  The
// launch boundary is the only synchronisation.  A `done` flag (set by workgroup 0
// of the kernel that decides to stop) turns later launches into no-ops, so the host
// enqueues chunks of iterations and synchronises only between chunks.
//
// Workgroup -> work: launch w is mapped to logical workgroup (w % 8) * (G / 8) + w / 8,
// so each of the 8 XCDs (which receive workgroups round-robin) streams one
// contiguous eighth of the matrix and keeps its neighbour gathers in its own L2.
//
// The matrix is in SELL-C-sigma form (C = 64, see eolc_internal.h): each wavefront
// of the SpMV owns one 64-row slice, one thread per block row; at column position
// k the 64 lanes read 64 consecutive entries of each of the 9 SoA value arrays
// (A[c * nbs + e]) and of the column array, fully coalesced, and because rows are
// sorted by length inside the slice the active lanes are always a prefix, so the
// padding is never fetched.  Products and row sums are accumulated in FP64 also
// when A, p, q are FP32.
#include <type_traits>

#include "eolc_internal.h"
#include "pcg_device.h"

using eolc::SolverScalars;

// Every multiply-add below is written out (fma where fused): no contraction is
// left to the compiler, so all launch variants of a kernel round identically.
#pragma clang fp contract(off)

namespace {

// ---------------------------------------------------------------------------
// DIFF = false (FP64 path): q = A p with the assembled blocks of MDK.
// DIFF = true  (FP32 path): q_i = Mbar_i p_i + sum_{j != i} S'_ij (p_j - p_i) with
// S'_ij = dampingB h^2 K_ij + m_ij I and the row mass Mbar_i = sum_j m_ij (exact
// algebra: sum_j m_ij p_j = Mbar_i p_i + sum_{j != i} m_ij (p_j - p_i)).  The
// stiffness then annihilates translations exactly and its FP32 rounding is relative
// to the local differences p_j - p_i, not to |p|; the smooth, slowly converging modes
// see the mass through Mbar_i.  An assembled FP32 copy of M + h^2 K instead loses the
// mass (~1e-8) under the rounding of the stiffness (~1e-2 x 6e-8).  Each row's
// diagonal block is its last SELL entry, so the FP32 loop stops one entry early.
//
// Launch variants (same arithmetic in the same order, so bitwise-identical results;
// chosen per matrix by eolc_autotune):
//   U   entries per loop trip (1 or 2): all loads of a trip are issued together
//   PF  the column index of the next trip is loaded one trip ahead, so the gathers of
//       p and z do not wait for a dependent load
//   NT  matrix values and columns are read with non-temporal loads (streamed past the
//       caches, which then keep the gathered vectors)
//   W   waves per SIMD the register allocation targets
template <class T, bool NT>
__device__ __forceinline__ T ldm(const T *p)
{
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

// One block of the product, accumulated into (y0, y1, y2):
//   DIFF:  y += S' (u - ui)   (FP32 terms, FP64 sums)
//   else:  y += A u          (FP64)
template <class T, bool DIFF>
__device__ __forceinline__ void block_apply(T a0, T a1, T a2, T a3, T a4, T a5, T a6, T a7, T a8, T u0, T u1, T u2,
                                            T ui0, T ui1, T ui2, double &y0, double &y1, double &y2)
{
    if constexpr (DIFF) {
        const T d0 = u0 - ui0, d1 = u1 - ui1, d2 = u2 - ui2;
        y0 += (double)fma(a2, d2, fma(a1, d1, a0 * d0));
        y1 += (double)fma(a5, d2, fma(a4, d1, a3 * d0));
        y2 += (double)fma(a8, d2, fma(a7, d1, a6 * d0));
    } else {
        y0 += fma((double)a2, (double)u2, fma((double)a1, (double)u1, (double)a0 * (double)u0));
        y1 += fma((double)a5, (double)u2, fma((double)a4, (double)u1, (double)a3 * (double)u0));
        y2 += fma((double)a8, (double)u2, fma((double)a7, (double)u1, (double)a6 * (double)u0));
    }
}

template <class T, bool DIFF>
struct SpmvEntry {
    T a[9];
    T u0, u1, u2;
    template <bool NT>
    __device__ __forceinline__ void load(const T *__restrict__ A, int64_t nbs, int e, int j, T betaT,
                                         const T *__restrict__ z, const T *__restrict__ pold)
    {
        u0 = fma(betaT, pold[3 * j], z[3 * j]);
        u1 = fma(betaT, pold[3 * j + 1], z[3 * j + 1]);
        u2 = fma(betaT, pold[3 * j + 2], z[3 * j + 2]);
#pragma unroll
        for (int c = 0; c < 9; ++c) a[c] = ldm<T, NT>(A + c * nbs + e);
    }
    __device__ __forceinline__ void accumulate(T ui0, T ui1, T ui2, double &y0, double &y1, double &y2) const
    {
        block_apply<T, DIFF>(a[0], a[1], a[2], a[3], a[4], a[5], a[6], a[7], a[8], u0, u1, u2, ui0, ui1, ui2, y0, y1,
                             y2);
    }
};

// Symmetric FP32 layout (SYM, difference form only): every off-diagonal block is
// stored once, as a 40-byte record {column, S'_ij} with its upper row; a lower entry
// (column < row) is a record {column, upper record of (j, i)} and applies that block
// transposed, read through the caches (row j was streamed shortly before).  Same
// entries, same order, same values (the assembled blocks satisfy K_ji = K_ij^T
// exactly) as the full layout, so the result is bitwise identical.
struct SymLayout {
    const int *supper, *uoff, *loff, *lrec;
    const float *S32u;
    const float *S32a;   // record layout: {column bits, S'_ij (9)} for every SELL entry, in SELL order
    const double *S64u;  // the symmetric records in FP64 (layout 4 of the FP64 product)
};

// Upper records of the symmetric layout: 10 values {column bits, S'_ij (9)}, read as
// five 2-vectors (40 B records in FP32, 80 B in FP64).
template <class T> struct SymRec;
template <> struct SymRec<float> {
    typedef float v2 __attribute__((ext_vector_type(2)));
    static __device__ __forceinline__ const float *base(const SymLayout &s) { return s.S32u; }
    static __device__ __forceinline__ int col(v2 r) { return __float_as_int(r.x); }
};
template <> struct SymRec<double> {
    typedef double v2 __attribute__((ext_vector_type(2)));
    static __device__ __forceinline__ const double *base(const SymLayout &s) { return s.S64u; }
    static __device__ __forceinline__ int col(v2 r) { return (int)__double_as_longlong(r.x); }
};

// Block layouts (LAY): 0 = nine SoA value arrays plus the column array (coalesced
// 4-byte loads); 1 = symmetric records (above); 2 = one 40-byte record {column,
// S'_ij} per SELL entry, read as five 8-byte loads per lane (five load instructions
// per block instead of ten); 3 = layout 0 with the next entry's loads issued before
// the current entry's products; 4 = symmetric records in either precision (the
// FP64 product then runs in the difference form too, on 80-byte records: 41% fewer
// algorithmic bytes than its assembled SoA form, different rounding).  1 and 2 are
// FP32 difference form only; 4 is layout 1 for FP32.
//
// STG (LDS-staged gathers): the workgroup's vertex window [wwin[2 lw], wwin[2 lw + 1])
// holds its rows and every column they reference (host-computed; at most
// kStageVerts vertices).  Its p_old and z are loaded (coalesced) before the partial
// sums are re-reduced, p_new = z + beta p_old is formed once per window vertex into
// LDS, and every neighbour gather reads LDS instead of global memory: the chain
// column index -> global gather -> product becomes column index -> LDS read.  Same
// values (the same fma per component), so bitwise-identical results.
template <class T, bool DIFF, int U, bool PF, bool NT, int W, int LAY, bool STG = false>
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(W, W)))
k_pcg_spmv(int k, int N, int nslices, const int *__restrict__ slice_off, const int *__restrict__ srow,
           const int *__restrict__ slen, const int *__restrict__ scol, const T *__restrict__ A,
           const T *__restrict__ Mm, int64_t nbs, const unsigned char *__restrict__ fixmask, const T *__restrict__ z,
           const T *__restrict__ pold, T *__restrict__ pnew, T *__restrict__ q, SolverScalars *scal, double *partials,
           SymLayout sym, const int *__restrict__ wwin)
{
    static_assert(U == 1 || U == 2, "entries per trip");
    static_assert(!(STG && U == 2), "the staged variants take one entry per trip");
    __shared__ double lds[3 * (WG / 64) + 3];
    __shared__ T su[STG ? 3 * eolc::kStageVerts : 1];
    if (scal->done) return;
    // the own row's indices and vectors do not depend on beta: their loads are
    // issued before the partial sums are re-reduced, so the two latencies overlap
    const int sl = wave_slice();
    const int lane = threadIdx.x & 63;
    const int g = sl * eolc::kSlice + lane;
    const bool active = sl < nslices && g < N;
    int row = 0, len = 0, base = 0;
    T zi0 = 0, zi1 = 0, zi2 = 0, po0 = 0, po1 = 0, po2 = 0, mb = 0;
    if (active) {
        row = srow[g];
        // every row holds its diagonal block as the last entry; the FP32 difference
        // form skips it (its S' block is 0 and the row mass is applied below)
        len = slen[g] - (DIFF ? 1 : 0);
        base = slice_off[sl] + lane;
        zi0 = z[3 * row]; zi1 = z[3 * row + 1]; zi2 = z[3 * row + 2];
        po0 = pold[3 * row]; po1 = pold[3 * row + 1]; po2 = pold[3 * row + 2];
        if constexpr (DIFF) mb = Mm[row];
    }
    // the window's p_old and z: loaded now, combined with beta below
    constexpr int kPer = STG ? (3 * eolc::kStageVerts + WG - 1) / WG : 1;
    int wlo = 0, nf = 0;
    T spo[kPer], sz[kPer];
    if constexpr (STG) {
        const int lw = logical_wg();
        wlo = wwin[2 * lw];
        nf = 3 * (wwin[2 * lw + 1] - wlo);
        const T *__restrict__ pw = pold + 3 * (int64_t)wlo;
        const T *__restrict__ zw = z + 3 * (int64_t)wlo;
#pragma unroll
        for (int t = 0; t < kPer; ++t) {
            const int f = t * WG + (int)threadIdx.x;
            spo[t] = f < nf ? pw[f] : (T)0;
            sz[t] = f < nf ? zw[f] : (T)0;
        }
    }
    double cur[2];
    const bool restart_beta0 = (k == scal->first_iter);
    double prev[1] = {1.0};
    if (restart_beta0) reduce_partials<2>(PB(partials, k & 1), cur, lds);
    else reduce_partials2<2, 1>(PB(partials, k & 1), PB(partials, (k + 1) & 1), cur, prev, lds);
    const double rzk = cur[0], rrk = cur[1];
    const double thr = fmax(scal->tol2 * scal->bnorm2, scal->delta2 * scal->rmax2);
    if (!isfinite(rrk) || !isfinite(rzk) || rrk <= thr) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            scal->done = 1;
            scal->reason = (isfinite(rrk) && isfinite(rzk)) ? 1 : 4;
            scal->rr = rrk;
            scal->rz = rzk;
        }
        return;
    }
    const double beta = restart_beta0 ? 0.0 : rzk / prev[0];
    const T betaT = (T)beta;   // p_new = z + beta p is formed in T, identically everywhere
    if constexpr (STG) {
#pragma unroll
        for (int t = 0; t < kPer; ++t) {
            const int f = t * WG + (int)threadIdx.x;
            if (f < nf) su[f] = fma(betaT, spo[t], sz[t]);
        }
        __syncthreads();
    }
    // p_new of column j: from LDS (staged) or formed from global p_old and z
    auto gat = [&](int j, T &u0, T &u1, T &u2) {
        if constexpr (STG) {
            const T *sj = su + 3 * (j - wlo);
            u0 = sj[0]; u1 = sj[1]; u2 = sj[2];
        } else {
            u0 = fma(betaT, pold[3 * j], z[3 * j]);
            u1 = fma(betaT, pold[3 * j + 1], z[3 * j + 1]);
            u2 = fma(betaT, pold[3 * j + 2], z[3 * j + 2]);
        }
    };
    double dot = 0.0;
    if (active) {
        const T *__restrict__ a = A + base;
        const int *__restrict__ cl = scol + base;
        T ui0 = fma(betaT, po0, zi0);
        T ui1 = fma(betaT, po1, zi1);
        T ui2 = fma(betaT, po2, zi2);
        if constexpr (!DIFF) {
            // the own row's p is not needed inside the loop: store it now, reload after
            pnew[3 * row] = ui0; pnew[3 * row + 1] = ui1; pnew[3 * row + 2] = ui2;
        }
        double y0 = 0, y1 = 0, y2 = 0;
        if constexpr (DIFF) {
            y0 = (double)(mb * ui0);
            y1 = (double)(mb * ui1);
            y2 = (double)(mb * ui2);
        }
        int kk = 0;
        if constexpr (DIFF && LAY == 2) {
            static_assert(std::is_same<T, float>::value, "record layout is FP32");
            typedef float f2v __attribute__((ext_vector_type(2)));
            const float *__restrict__ rb = sym.S32a + (size_t)10 * base;
#pragma unroll 1
            for (kk = 0; kk < len; ++kk) {
                const f2v *rec = reinterpret_cast<const f2v *>(rb + (size_t)640 * kk);
                const f2v r0 = ldm<f2v, NT>(rec), r1 = ldm<f2v, NT>(rec + 1), r2 = ldm<f2v, NT>(rec + 2),
                          r3 = ldm<f2v, NT>(rec + 3), r4 = ldm<f2v, NT>(rec + 4);
                const int j = __float_as_int(r0.x);
                T u0, u1, u2;
                gat(j, u0, u1, u2);
                block_apply<T, DIFF>(r0.y, r1.x, r1.y, r2.x, r2.y, r3.x, r3.y, r4.x, r4.y, u0, u1, u2, ui0, ui1, ui2,
                                     y0, y1, y2);
            }
        } else if constexpr (DIFF && (LAY == 1 || LAY == 4)) {
            using R = SymRec<T>;
            typedef typename R::v2 v2;
            const T *__restrict__ ubase = R::base(sym);
            const int nu = sym.supper[g];
            const int nl = len - nu;
            const T *__restrict__ ur = ubase + (size_t)10 * (sym.uoff[sl] + lane);
#pragma unroll 1
            for (kk = 0; kk < nu; ++kk) {
                const v2 *rec = reinterpret_cast<const v2 *>(ur + (size_t)640 * kk);
                const v2 r0 = ldm<v2, NT>(rec), r1 = ldm<v2, NT>(rec + 1), r2 = ldm<v2, NT>(rec + 2),
                         r3 = ldm<v2, NT>(rec + 3), r4 = ldm<v2, NT>(rec + 4);
                const int j = R::col(r0);
                T u0, u1, u2;
                gat(j, u0, u1, u2);
                block_apply<T, DIFF>(r0.y, r1.x, r1.y, r2.x, r2.y, r3.x, r3.y, r4.x, r4.y, u0, u1, u2, ui0, ui1, ui2,
                                     y0, y1, y2);
            }
            typedef int i2v __attribute__((ext_vector_type(2)));
            const i2v *__restrict__ lr = reinterpret_cast<const i2v *>(sym.lrec) + (sym.loff[sl] + lane);
#pragma unroll 1
            for (kk = 0; kk < nl; ++kk) {
                const i2v jt = ldm<i2v, NT>(lr + (size_t)64 * kk);
                const int j = jt.x;
                // block (j, i), record {column, 9 values}: apply its transpose
                const v2 *rec = reinterpret_cast<const v2 *>(ubase + (size_t)10 * jt.y);
                const v2 r0 = rec[0], r1 = rec[1], r2 = rec[2], r3 = rec[3], r4 = rec[4];
                T u0, u1, u2;
                gat(j, u0, u1, u2);
                block_apply<T, DIFF>(r0.y, r2.x, r3.y, r1.x, r2.y, r4.x, r1.y, r3.x, r4.y, u0, u1, u2, ui0, ui1, ui2,
                                     y0, y1, y2);
            }
        } else if constexpr (LAY == 3) {
            // SoA, software-pipelined: the next entry's column and nine values are
            // loaded before the current entry's gathers and products (twice the loads
            // in flight per wavefront)
            if (len > 0) {
                int jn = ldm<int, NT>(cl);
                T n0 = ldm<T, NT>(a), n1 = ldm<T, NT>(a + nbs), n2 = ldm<T, NT>(a + 2 * nbs),
                  n3 = ldm<T, NT>(a + 3 * nbs), n4 = ldm<T, NT>(a + 4 * nbs), n5 = ldm<T, NT>(a + 5 * nbs),
                  n6 = ldm<T, NT>(a + 6 * nbs), n7 = ldm<T, NT>(a + 7 * nbs), n8 = ldm<T, NT>(a + 8 * nbs);
#pragma unroll 1
                for (; kk < len; ++kk) {
                    const int j = jn;
                    const T c0 = n0, c1 = n1, c2 = n2, c3 = n3, c4 = n4, c5 = n5, c6 = n6, c7 = n7, c8 = n8;
                    const int en = (kk + 1 < len ? kk + 1 : kk) * eolc::kSlice;   // clamped: a valid entry
                    jn = ldm<int, NT>(cl + en);
                    const T *__restrict__ an = a + en;
                    n0 = ldm<T, NT>(an); n1 = ldm<T, NT>(an + nbs); n2 = ldm<T, NT>(an + 2 * nbs);
                    n3 = ldm<T, NT>(an + 3 * nbs); n4 = ldm<T, NT>(an + 4 * nbs); n5 = ldm<T, NT>(an + 5 * nbs);
                    n6 = ldm<T, NT>(an + 6 * nbs); n7 = ldm<T, NT>(an + 7 * nbs); n8 = ldm<T, NT>(an + 8 * nbs);
                    T u0, u1, u2;
                    gat(j, u0, u1, u2);
                    block_apply<T, DIFF>(c0, c1, c2, c3, c4, c5, c6, c7, c8, u0, u1, u2, ui0, ui1, ui2, y0, y1, y2);
                }
            }
        } else if constexpr (U == 1) {
            int jn = ldm<int, NT>(cl);
#pragma unroll 1
            for (; kk < len; ++kk) {
                const int e = kk * eolc::kSlice;
                int j;
                if constexpr (PF) {
                    j = jn;
                    const int kn = kk + 1 < len ? kk + 1 : kk;   // clamped: always a valid entry
                    jn = ldm<int, NT>(cl + kn * eolc::kSlice);
                } else {
                    j = ldm<int, NT>(cl + e);
                }
                T u0, u1, u2;
                gat(j, u0, u1, u2);
                const T *__restrict__ ae = a + e;
                block_apply<T, DIFF>(ldm<T, NT>(ae), ldm<T, NT>(ae + nbs), ldm<T, NT>(ae + 2 * nbs),
                                     ldm<T, NT>(ae + 3 * nbs), ldm<T, NT>(ae + 4 * nbs), ldm<T, NT>(ae + 5 * nbs),
                                     ldm<T, NT>(ae + 6 * nbs), ldm<T, NT>(ae + 7 * nbs), ldm<T, NT>(ae + 8 * nbs),
                                     u0, u1, u2, ui0, ui1, ui2, y0, y1, y2);
            }
        } else {
            int jn0 = ldm<int, NT>(cl), jn1 = len > 1 ? ldm<int, NT>(cl + eolc::kSlice) : 0;
#pragma unroll 1
            for (; kk + 1 < len; kk += 2) {
                const int e = kk * eolc::kSlice;
                int j0, j1;
                if constexpr (PF) {
                    j0 = jn0;
                    j1 = jn1;
                    const int k0 = kk + 2 < len ? kk + 2 : kk, k1 = kk + 3 < len ? kk + 3 : kk;
                    jn0 = ldm<int, NT>(cl + k0 * eolc::kSlice);
                    jn1 = ldm<int, NT>(cl + k1 * eolc::kSlice);
                } else {
                    j0 = ldm<int, NT>(cl + e);
                    j1 = ldm<int, NT>(cl + e + eolc::kSlice);
                }
                SpmvEntry<T, DIFF> x0, x1;
                x0.template load<NT>(a, nbs, e, j0, betaT, z, pold);
                x1.template load<NT>(a, nbs, e + eolc::kSlice, j1, betaT, z, pold);
                x0.accumulate(ui0, ui1, ui2, y0, y1, y2);
                x1.accumulate(ui0, ui1, ui2, y0, y1, y2);
            }
            if (kk < len) {
                const int e = kk * eolc::kSlice;
                SpmvEntry<T, DIFF> x0;
                x0.template load<NT>(a, nbs, e, ldm<int, NT>(cl + e), betaT, z, pold);
                x0.accumulate(ui0, ui1, ui2, y0, y1, y2);
            }
        }
        const unsigned char fm = fixmask[row];
        if (fm & 1) y0 = 0;
        if (fm & 2) y1 = 0;
        if (fm & 4) y2 = 0;
        if constexpr (DIFF) {
            pnew[3 * row] = ui0; pnew[3 * row + 1] = ui1; pnew[3 * row + 2] = ui2;
        } else {
            ui0 = pnew[3 * row]; ui1 = pnew[3 * row + 1]; ui2 = pnew[3 * row + 2];
        }
        const T q0 = (T)y0, q1 = (T)y1, q2 = (T)y2;
        q[3 * row] = q0; q[3 * row + 1] = q1; q[3 * row + 2] = q2;
        dot = ((double)ui0 * (double)q0 + (double)ui1 * (double)q1) + (double)ui2 * (double)q2;
    }
    double v[1] = {dot};
    write_partials<1>(v, PA(partials), lds);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal->rr = rrk;
        scal->rz = rzk;
    }
}

// ---------------------------------------------------------------------------
// Patch layout (LAY 5; difference form in either precision).  Workgroup lw owns the
// rows of patch lw, a compact set of at most 1024 vertices (Topology::pt_*), and
// stages p_new = z + beta p_old of the patch's window (its rows, then the vertices of
// other patches they reference) in LDS; records name their column by window index,
// so no product gathers from global memory.
//   Phase 1: every row applies its upper records (column > row, increasing
//   column); for an upper block whose column lies in the patch it also leaves the
//   transposed product for that column's row in an LDS slot.
//   Phase 2, after a workgroup barrier: every row adds its lower blocks in
//   increasing column order, from the LDS slot when the column is in the patch, else
//   from its next cross record (read from HBM like an upper record).
// An off-diagonal block inside a patch is therefore read from HBM once (layout 1
// reads it a second time, transposed and scattered through the caches).  Every
// row's sum takes the same terms in the same order as every other difference-form
// variant: upper entries, then lower entries, and the lower-row product
// S'_ji (p_j - p_i) = S'_ij^T (p_j - p_i) is -(S'_ij^T (p_i - p_j)) exactly (IEEE
// negation is exact and rounding is symmetric), so q is bitwise the q of layout 4.
// The workgroup partial of p.q sums the patch's rows instead of 1,024 consecutive
// rows, so p.q, and with it alpha, can differ in the last bits.
//
// LDS pool (kPatchPool entries of 3 values): transposed products, component c of
// slot s at c S + s (S = slots of the largest patch), then the window, vertex w at
// 3 S + 3 w.
struct PatchLayout {
    const int *row, *cnt, *roff, *soff, *loff, *woff, *win;
    const unsigned short *lo;
    const float *R32;
    const double *R64;
    int S;
    // blocks too stiff for FP32 (their FP32 records are zero): the row's mask of
    // stiff SELL entries, its first SELL entry, the SELL columns and the FP64 values
    const unsigned long long *smask;
    const int *sbase, *scol;
    const double *A64s;   // the stiff blocks' FP64 values, 9 contiguous per SELL entry
};
template <class T> __device__ __forceinline__ const T *patch_records(const PatchLayout &pl);
template <> __device__ __forceinline__ const float *patch_records<float>(const PatchLayout &pl) { return pl.R32; }
template <> __device__ __forceinline__ const double *patch_records<double>(const PatchLayout &pl) { return pl.R64; }

constexpr int kPatchLo = 8;   // lower entries per row loaded before phase 1 (the rest in phase 2)

// PD: how many records of a row are in flight ahead of the one in use (2 fits the
// FP64 kernel's 128 VGPRs at 4 waves per SIMD; in FP32 at 8 waves it spills).
// T: vectors and products; TR: the stored records (the same type in every launch
// variant; an FP32 record with FP64 vectors would need reliable updates like the
// mixed solve, and measured no faster per FP64 solve).  With FP32 records, a block whose
// entries exceed the context's stiff threshold is applied from its FP64 values by
// both of its rows (its FP32 record is zero): rounded to FP32, such a block (a
// collapsed element, up to 1e10 times the others) adds errors of 6e-8 of its size
// in every direction, also in its own null directions where the rest of the
// matrix is soft, and no FP32 iteration then converges.
template <class T, class TR, bool NT, int W, int PD>
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(W, W)))
k_pcg_spmv_patch(int k, PatchLayout pl, const T *__restrict__ Mm, const unsigned char *__restrict__ fixmask,
                 const T *__restrict__ z, const T *__restrict__ pold, T *__restrict__ pnew, T *__restrict__ q,
                 SolverScalars *scal, double *partials)
{
    using R = SymRec<TR>;
    typedef typename R::v2 v2;
    __shared__ double lds[3 * (WG / 64) + 3];
    __shared__ T pool[3 * eolc::kPatchPool];
    if (scal->done) return;
    const int S = pl.S;
    T *const su = pool + 3 * S;   // window values
    // Everything that does not depend on beta is loaded before the partial sums are
    // re-reduced: the row's counts, mass and fixed mask, its first lower entries, and
    // p_old and z of the window.
    const int lw = logical_wg();
    const int lane = threadIdx.x & 63;
    const int sl = lw * (WG / 64) + (threadIdx.x >> 6);
    const int g = sl * eolc::kSlice + lane;
    const int row = pl.row[g];
    const bool active = row >= 0;
    int nup = 0, nlo = 0;
    T mb = 0;
    unsigned char fm = 0;
    unsigned short le[kPatchLo];
    const unsigned short *__restrict__ lo = pl.lo + pl.loff[sl] + lane;
    if (active) {
        const unsigned cnt = (unsigned)pl.cnt[g];
        nup = (int)(cnt & 0xffffu);
        nlo = (int)(cnt >> 16);
        mb = Mm[row];
        fm = fixmask[row];
#pragma unroll
        for (int t = 0; t < kPatchLo; ++t) le[t] = t < nlo ? lo[t * eolc::kSlice] : eolc::kPatchCross;
    }
    const TR *__restrict__ rb = patch_records<TR>(pl) + (size_t)10 * (pl.roff[sl] + lane);
    const int s0 = pl.soff[sl] + lane;
    const int w0 = pl.woff[lw], nf = 3 * (pl.woff[lw + 1] - w0);
    constexpr int kPer = (3 * eolc::kPatchWin + WG - 1) / WG;
    T spo[kPer], sz[kPer];
#pragma unroll
    for (int t = 0; t < kPer; ++t) {
        const int f = t * WG + (int)threadIdx.x;
        spo[t] = 0;
        sz[t] = 0;
        if (f < nf) {
            const int v = pl.win[w0 + f / 3], c = f % 3;
            spo[t] = pold[3 * v + c];
            sz[t] = z[3 * v + c];
        }
    }
    double cur[2];
    const bool restart_beta0 = (k == scal->first_iter);
    double prev[1] = {1.0};
    if (restart_beta0) reduce_partials<2>(PB(partials, k & 1), cur, lds);
    else reduce_partials2<2, 1>(PB(partials, k & 1), PB(partials, (k + 1) & 1), cur, prev, lds);
    const double rzk = cur[0], rrk = cur[1];
    const double thr = fmax(scal->tol2 * scal->bnorm2, scal->delta2 * scal->rmax2);
    if (!isfinite(rrk) || !isfinite(rzk) || rrk <= thr) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            scal->done = 1;
            scal->reason = (isfinite(rrk) && isfinite(rzk)) ? 1 : 4;
            scal->rr = rrk;
            scal->rz = rzk;
        }
        return;
    }
    const double beta = restart_beta0 ? 0.0 : rzk / prev[0];
    const T betaT = (T)beta;   // p_new = z + beta p is formed in T, identically everywhere
#pragma unroll
    for (int t = 0; t < kPer; ++t) {
        const int f = t * WG + (int)threadIdx.x;
        if (f < nf) su[f] = fma(betaT, spo[t], sz[t]);
    }
    __syncthreads();
    // the row's own p_new: window index = position in the patch = thread index
    T ui0 = 0, ui1 = 0, ui2 = 0;
    if (active) {
        ui0 = su[3 * threadIdx.x];
        ui1 = su[3 * threadIdx.x + 1];
        ui2 = su[3 * threadIdx.x + 2];
    }
    double y0 = (double)(mb * ui0), y1 = (double)(mb * ui1), y2 = (double)(mb * ui2);
    if (active && nup > 0) {
        // records streamed PD ahead: record kk + PD's loads are issued before record
        // kk's LDS reads and products
        v2 nb[PD][5];
#pragma unroll
        for (int d = 0; d < PD; ++d) {
            if (d == 0 || d < nup) {
                const v2 *rn = reinterpret_cast<const v2 *>(rb + (size_t)640 * d);
#pragma unroll
                for (int h = 0; h < 5; ++h) nb[d][h] = ldm<v2, NT>(rn + h);
            }
        }
#pragma unroll 1
        for (int kk = 0; kk < nup; ++kk) {
            const v2 r0 = nb[0][0], r1 = nb[0][1], r2 = nb[0][2], r3 = nb[0][3], r4 = nb[0][4];
#pragma unroll
            for (int d = 0; d + 1 < PD; ++d)
#pragma unroll
                for (int h = 0; h < 5; ++h) nb[d][h] = nb[d + 1][h];
            if (kk + PD < nup) {
                const v2 *rn = reinterpret_cast<const v2 *>(rb + (size_t)640 * (kk + PD));
#pragma unroll
                for (int h = 0; h < 5; ++h) nb[PD - 1][h] = ldm<v2, NT>(rn + h);
            }
            const T *sj = su + 3 * R::col(r0);
            const T u0 = sj[0], u1 = sj[1], u2 = sj[2];
            const T a0 = (T)r0.y, a1 = (T)r1.x, a2 = (T)r1.y, a3 = (T)r2.x, a4 = (T)r2.y, a5 = (T)r3.x,
                    a6 = (T)r3.y, a7 = (T)r4.x, a8 = (T)r4.y;
            block_apply<T, true>(a0, a1, a2, a3, a4, a5, a6, a7, a8, u0, u1, u2, ui0, ui1, ui2, y0, y1, y2);
            // S'^T (p_j - p_i) for row j; stored negated: -(S'^T (p_j - p_i)) = S'^T (p_i - p_j).
            // Written for cross upper blocks too (their slots are never read).
            const T d0 = u0 - ui0, d1 = u1 - ui1, d2 = u2 - ui2;
            const int slot = s0 + kk * eolc::kSlice;
            pool[slot] = -fma(a6, d2, fma(a3, d1, a0 * d0));
            pool[S + slot] = -fma(a7, d2, fma(a4, d1, a1 * d0));
            pool[2 * S + slot] = -fma(a8, d2, fma(a5, d1, a2 * d0));
        }
    }
    __syncthreads();
    double dot = 0.0;
    if (active) {
        int cr = nup;   // next cross record
#pragma unroll 1
        for (int kk = 0; kk < nlo; ++kk) {
            unsigned e = eolc::kPatchCross;
#pragma unroll
            for (int t = 0; t < kPatchLo; ++t)
                if (t == kk) e = le[t];
            if (kk >= kPatchLo) e = lo[kk * eolc::kSlice];
            if (e != eolc::kPatchCross) {
                y0 += (double)pool[e];
                y1 += (double)pool[S + e];
                y2 += (double)pool[2 * S + e];
            } else {
                const v2 *rec = reinterpret_cast<const v2 *>(rb + (size_t)640 * cr);
                ++cr;
                const v2 r0 = ldm<v2, NT>(rec), r1 = ldm<v2, NT>(rec + 1), r2 = ldm<v2, NT>(rec + 2),
                         r3 = ldm<v2, NT>(rec + 3), r4 = ldm<v2, NT>(rec + 4);
                const T *sj = su + 3 * R::col(r0);
                block_apply<T, true>((T)r0.y, (T)r1.x, (T)r1.y, (T)r2.x, (T)r2.y, (T)r3.x, (T)r3.y, (T)r4.x,
                                     (T)r4.y, sj[0], sj[1], sj[2], ui0, ui1, ui2, y0, y1, y2);
            }
        }
        if constexpr (std::is_same<TR, float>::value) {
            unsigned long long m = pl.smask[row];
            if (m) {
                const int base = pl.sbase[row];
                const double di0 = (double)ui0, di1 = (double)ui1, di2 = (double)ui2;
                while (m) {
                    const int e = base + __builtin_ctzll(m) * eolc::kSlice;
                    m &= m - 1;
                    const int j = pl.scol[e];
                    const double d0 = (double)fma(betaT, pold[3 * j], z[3 * j]) - di0;
                    const double d1 = (double)fma(betaT, pold[3 * j + 1], z[3 * j + 1]) - di1;
                    const double d2 = (double)fma(betaT, pold[3 * j + 2], z[3 * j + 2]) - di2;
                    const double *a = pl.A64s + 9 * (int64_t)e;
                    y0 += fma(a[2], d2, fma(a[1], d1, a[0] * d0));
                    y1 += fma(a[5], d2, fma(a[4], d1, a[3] * d0));
                    y2 += fma(a[8], d2, fma(a[7], d1, a[6] * d0));
                }
            }
        }
        if (fm & 1) y0 = 0;
        if (fm & 2) y1 = 0;
        if (fm & 4) y2 = 0;
        pnew[3 * row] = ui0; pnew[3 * row + 1] = ui1; pnew[3 * row + 2] = ui2;
        const T q0 = (T)y0, q1 = (T)y1, q2 = (T)y2;
        q[3 * row] = q0; q[3 * row + 1] = q1; q[3 * row + 2] = q2;
        dot = ((double)ui0 * (double)q0 + (double)ui1 * (double)q1) + (double)ui2 * (double)q2;
    }
    double v[1] = {dot};
    write_partials<1>(v, PA(partials), lds);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal->rr = rrk;
        scal->rz = rzk;
    }
}

// ---------------------------------------------------------------------------
// The multilevel modes' update (x += alpha p, r -= alpha q, r.r) touches no 3x3
// block, so it runs over the flat DOF arrays: workgroup lw owns DOFs
// [lw c, lw c + c) with c = ceil(3N / G) rounded up to 4, and every thread handles 4
// consecutive DOFs per trip with 16-byte loads (FP32) or two 16-byte loads (FP64).
// The same arithmetic per DOF as k_pcg_update; r.r is summed per DOF range.
template <class T, class TX>
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_pcg_update_flat(int k, int G, int n3, const T *__restrict__ pnew, const T *__restrict__ q, TX *__restrict__ x,
                  T *__restrict__ r, SolverScalars *scal, double *partials)
{
    __shared__ double lds[2 * (WG / 64) + 2];
    typedef T v4 __attribute__((ext_vector_type(4)));
    typedef TX x4 __attribute__((ext_vector_type(4)));
    if (scal->done) return;
    const int chunk = (((n3 + G - 1) / G) + 3) & ~3;
    const int lw0 = logical_wg(), lwx = extra_logical_wg(G);
    // mixed precision: the first trip's FP64 accumulator (half the bytes read) loads
    // before the partial sums are re-reduced (it does not depend on alpha), so the two
    // latencies overlap.  The other vectors too, or the FP64 mode's, would spill at
    // eight waves per SIMD.
    constexpr bool PRE = sizeof(T) == 4;
    const int b0 = min(lw0 * chunk, n3), e0 = min(b0 + chunk, n3), i0 = b0 + 4 * (int)threadIdx.x;
    x4 xv0 = {};
    if (PRE && i0 + 3 < e0) xv0 = *reinterpret_cast<const x4 *>(x + i0);
    double pq[1], rz[1];
    reduce_partials2<1, 1>(PA(partials), PB(partials, k & 1), pq, rz, lds, G);
    if (!(pq[0] > 0.0)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            scal->done = 1;
            scal->reason = isfinite(pq[0]) ? 2 : 4;
            scal->pq = pq[0];
        }
        return;
    }
    const double alpha = rz[0] / pq[0];
    const T alphaT = (T)alpha;
    // the G logical ranges may run on fewer workgroups: the first ones then take a
    // second range
#pragma unroll 1
    for (int pass = 0; pass < 2; ++pass) {
        const int lw = pass == 0 ? lw0 : lwx;
        if (lw < 0) break;   // uniform in the workgroup
        const int b = min(lw * chunk, n3), e = min(b + chunk, n3);
        double srr = 0.0;
        for (int i = b + 4 * (int)threadIdx.x; i < e; i += 4 * WG) {
            if (i + 3 < e) {
                // i is a multiple of 4: aligned 4-element vector loads and stores
                const v4 pv = *reinterpret_cast<const v4 *>(pnew + i), qv = *reinterpret_cast<const v4 *>(q + i);
                const v4 rv = *reinterpret_cast<const v4 *>(r + i);
                const x4 xv = (PRE && pass == 0 && i == i0) ? xv0 : *reinterpret_cast<const x4 *>(x + i);
                x4 xo;
                v4 ro;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if constexpr (std::is_same<TX, T>::value) xo[u] = fma(alphaT, pv[u], xv[u]);
                    else xo[u] = fma((TX)alpha, (TX)pv[u], xv[u]);
                    ro[u] = fma(-alphaT, qv[u], rv[u]);
                    srr += (double)ro[u] * (double)ro[u];
                }
                *reinterpret_cast<x4 *>(x + i) = xo;
                *reinterpret_cast<v4 *>(r + i) = ro;
            } else {
                for (int u = i; u < e; ++u) {
                    if constexpr (std::is_same<TX, T>::value) x[u] = fma(alphaT, pnew[u], x[u]);
                    else x[u] = fma((TX)alpha, (TX)pnew[u], x[u]);
                    const T rn = fma(-alphaT, q[u], r[u]);
                    r[u] = rn;
                    srr += (double)rn * (double)rn;
                }
            }
        }
        double v[1] = {srr};
        write_partials<1>(v, PB(partials, (k + 1) & 1) + eolc::kMaxPartials, lds, lw);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal->iters = k + 1;
        scal->pq = pq[0];
        scal->alpha = alpha;
    }
}

// ---------------------------------------------------------------------------
// TX is the solution accumulator (double in FP64 mode, float in mixed mode).  ML:
// the multilevel kernels form z and r.z afterwards (k_ml.hip); only r.r here.
template <class T, class TX, bool ML>
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_pcg_update(int k, int N, const T *__restrict__ D, const T *__restrict__ pnew, const T *__restrict__ q,
             TX *__restrict__ x, T *__restrict__ r, T *__restrict__ z, SolverScalars *scal, double *partials)
{
    __shared__ double lds[2 * (WG / 64) + 2];
    if (scal->done) return;
    // the thread's first vertex is loaded before the partial sums are re-reduced
    // (the loads do not depend on alpha), so the two latencies overlap
    int rb, re;
    wg_rows(N, &rb, &re);
    const int i0 = rb + (int)threadIdx.x;
    T pv[3] = {0, 0, 0}, qv[3] = {0, 0, 0}, rv[3] = {0, 0, 0};
    TX xv[3] = {0, 0, 0};
    if (i0 < re) {
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            pv[c] = pnew[3 * i0 + c]; qv[c] = q[3 * i0 + c]; rv[c] = r[3 * i0 + c]; xv[c] = x[3 * i0 + c];
        }
    }
    double pq[1], rz[1];
    reduce_partials2<1, 1>(PA(partials), PB(partials, k & 1), pq, rz, lds);
    if (!(pq[0] > 0.0)) {
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            scal->done = 1;
            scal->reason = isfinite(pq[0]) ? 2 : 4;
            scal->pq = pq[0];
        }
        return;
    }
    const double alpha = rz[0] / pq[0];
    const T alphaT = (T)alpha;
    double srz = 0.0, srr = 0.0;
    for (int i = i0; i < re; i += WG) {
        if (i != i0) {
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                pv[c] = pnew[3 * i + c]; qv[c] = q[3 * i + c]; rv[c] = r[3 * i + c]; xv[c] = x[3 * i + c];
            }
        }
        // the correction accumulates in TX: FP64 also in the mixed modes, whose FP32
        // rounding of x would break the constraints of very stiff blocks (errors of
        // 6e-8 |x| times the block) at every reliable update
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            if constexpr (std::is_same<TX, T>::value) x[3 * i + c] = fma(alphaT, pv[c], xv[c]);
            else x[3 * i + c] = fma((TX)alpha, (TX)pv[c], xv[c]);
        }
        const T r0 = fma(-alphaT, qv[0], rv[0]);
        const T r1 = fma(-alphaT, qv[1], rv[1]);
        const T r2 = fma(-alphaT, qv[2], rv[2]);
        r[3 * i] = r0; r[3 * i + 1] = r1; r[3 * i + 2] = r2;
        if constexpr (!ML) {
            T z0, z1, z2;
            eolc::apply_sym3<T>(D + i, N, r0, r1, r2, z0, z1, z2);
            z[3 * i] = z0; z[3 * i + 1] = z1; z[3 * i + 2] = z2;
            srz += ((double)r0 * (double)z0 + (double)r1 * (double)z1) + (double)r2 * (double)z2;
        }
        srr += ((double)r0 * (double)r0 + (double)r1 * (double)r1) + (double)r2 * (double)r2;
    }
    if constexpr (ML) {
        double v[1] = {srr};
        write_partials<1>(v, PB(partials, (k + 1) & 1) + eolc::kMaxPartials, lds);
    } else {
        double v[2] = {srz, srr};
        write_partials<2>(v, PB(partials, (k + 1) & 1), lds);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        scal->iters = k + 1;
        scal->pq = pq[0];
        scal->alpha = alpha;
    }
}

// ---------------------------------------------------------------------------
// out = b - A x (FP64), masked at fixed DOFs; b may be null (0).  Partial
// ||out||^2 -> partials (one array of G).
__global__ void __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(8, 8)))
k_resid64(int N, int nslices, const int *__restrict__ slice_off, const int *__restrict__ srow,
          const int *__restrict__ slen, const int *__restrict__ scol, const double *__restrict__ A, int64_t nbs,
          const unsigned char *__restrict__ fixmask, const double *__restrict__ xin, const double *__restrict__ b,
          double *__restrict__ out, double *partials)
{
    __shared__ double lds[2 * (WG / 64) + 2];
    const int sl = wave_slice();
    const int lane = threadIdx.x & 63;
    const int g = sl * eolc::kSlice + lane;
    double acc = 0.0;
    if (sl < nslices && g < N) {
        const int row = srow[g], len = slen[g];
        const double *__restrict__ a = A + slice_off[sl] + lane;
        const int *__restrict__ cl = scol + slice_off[sl] + lane;
        double y0 = 0, y1 = 0, y2 = 0;
#pragma unroll 1
        for (int kk = 0; kk < len; ++kk) {
            const int e = kk * eolc::kSlice;
            const int j = cl[e];
            const double u0 = xin[3 * j], u1 = xin[3 * j + 1], u2 = xin[3 * j + 2];
            block_apply<double, false>(a[0 * nbs + e], a[1 * nbs + e], a[2 * nbs + e], a[3 * nbs + e],
                                       a[4 * nbs + e], a[5 * nbs + e], a[6 * nbs + e], a[7 * nbs + e],
                                       a[8 * nbs + e], u0, u1, u2, 0.0, 0.0, 0.0, y0, y1, y2);
        }
        double o0 = (b ? b[3 * row] : 0.0) - y0;
        double o1 = (b ? b[3 * row + 1] : 0.0) - y1;
        double o2 = (b ? b[3 * row + 2] : 0.0) - y2;
        const unsigned char fm = fixmask[row];
        if (fm & 1) o0 = 0;
        if (fm & 2) o1 = 0;
        if (fm & 4) o2 = 0;
        out[3 * row] = o0; out[3 * row + 1] = o1; out[3 * row + 2] = o2;
        acc = (o0 * o0 + o1 * o1) + o2 * o2;
    }
    double v[1] = {acc};
    write_partials<1>(v, partials, lds);
}

// One workgroup: *dst = fixed-order sum of partials[0..G).
__global__ void __launch_bounds__(WG) k_reduce_to(int G, const double *__restrict__ partials, double *dst)
{
    __shared__ double lds[2 * (WG / 64) + 2];
    double s[1] = {0.0};
    for (int i = threadIdx.x; i < G; i += WG) s[0] += partials[i];
    block_allsum<1>(s, lds);
    if (threadIdx.x == 0) *dst = s[0];
}

// z = D r (r converted from FP64 to T), partials r.z, r.r -> PB[parity].  ML: r and
// r.r only; the multilevel kernels form z and r.z.
template <class T, bool ML>
__global__ void __launch_bounds__(WG)
k_precond_init(int parity, int N, const T *__restrict__ D, const double *r64, T *r, T *__restrict__ z,
               double *partials)
{
    __shared__ double lds[2 * (WG / 64) + 2];
    int rb, re;
    wg_rows(N, &rb, &re);
    double rz = 0.0, rr = 0.0;
    for (int i = rb + threadIdx.x; i < re; i += WG) {
        const T r0 = (T)r64[3 * i], r1 = (T)r64[3 * i + 1], r2 = (T)r64[3 * i + 2];
        r[3 * i] = r0; r[3 * i + 1] = r1; r[3 * i + 2] = r2;
        const double s0 = r0, s1 = r1, s2 = r2;
        if constexpr (!ML) {
            const T *d = D + i;
            const double e00 = d[0], e01 = d[N], e02 = d[2 * N], e11 = d[3 * N], e12 = d[4 * N], e22 = d[5 * N];
            const T z0 = (T)fma(e02, s2, fma(e01, s1, e00 * s0));
            const T z1 = (T)fma(e12, s2, fma(e11, s1, e01 * s0));
            const T z2 = (T)fma(e22, s2, fma(e12, s1, e02 * s0));
            z[3 * i] = z0; z[3 * i + 1] = z1; z[3 * i + 2] = z2;
            rz += (s0 * (double)z0 + s1 * (double)z1) + s2 * (double)z2;
        }
        rr += (s0 * s0 + s1 * s1) + s2 * s2;
    }
    if constexpr (ML) {
        double v[1] = {rr};
        write_partials<1>(v, PB(partials, parity) + eolc::kMaxPartials, lds);
    } else {
        double v[2] = {rz, rr};
        write_partials<2>(v, PB(partials, parity), lds);
    }
}

// Block-Jacobi inverse of the diagonal blocks, with fixed DOFs replaced by
// identity rows/columns.  kind 0: 3x3 blocks; kind 1: diagonal only (Jacobi).
__global__ void __launch_bounds__(256)
k_block_jacobi(int N, const int *__restrict__ diagpos, const double *__restrict__ A, int64_t nbs,
               const unsigned char *__restrict__ fixmask, int kind, double *__restrict__ D64,
               float *__restrict__ D32)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    int k = diagpos[i];
    double a[9];
    unsigned char fm = fixmask[i];
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        int r = c / 3, s = c % 3;
        double val = A[c * nbs + k];
        if (kind == 1 && r != s) val = 0.0;
        if (((fm >> r) & 1) || ((fm >> s) & 1)) val = (r == s) ? 1.0 : 0.0;
        a[c] = val;
    }
    double c00 = a[4] * a[8] - a[5] * a[7], c01 = a[5] * a[6] - a[3] * a[8], c02 = a[3] * a[7] - a[4] * a[6];
    double det = a[0] * c00 + a[1] * c01 + a[2] * c02;
    double o[9];
    if (det != 0.0 && isfinite(det)) {
        double id = 1.0 / det;
        o[0] = c00 * id; o[1] = (a[2] * a[7] - a[1] * a[8]) * id; o[2] = (a[1] * a[5] - a[2] * a[4]) * id;
        o[3] = c01 * id; o[4] = (a[0] * a[8] - a[2] * a[6]) * id; o[5] = (a[2] * a[3] - a[0] * a[5]) * id;
        o[6] = c02 * id; o[7] = (a[1] * a[6] - a[0] * a[7]) * id; o[8] = (a[0] * a[4] - a[1] * a[3]) * id;
    } else {
#pragma unroll
        for (int c = 0; c < 9; ++c) o[c] = (c % 4 == 0) ? 1.0 : 0.0;
    }
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        int r = c / 3, s = c % 3;
        if (((fm >> r) & 1) || ((fm >> s) & 1)) o[c] = 0.0;   // z stays 0 on fixed DOFs
    }
    const int up[eolc::kDsym] = {0, 1, 2, 4, 5, 8};           // upper entries (the inverse is symmetric)
#pragma unroll
    for (int k = 0; k < eolc::kDsym; ++k) {
        D64[k * N + i] = o[up[k]];
        D32[k * N + i] = (float)o[up[k]];
    }
}

// Fixed values of Constraints.cpp:114-119 from the device-resident velocity:
// vals[t] = (1 - 0.01) v_old[dofs[t]] + vals[t], vals holding the scripted velocity.
__global__ void k_fixed_from_vold(int n, const int *__restrict__ dofs, const double *__restrict__ v,
                                  double *__restrict__ vals)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    vals[t] = (1 - 0.01) * v[dofs[t]] + vals[t];
}

__global__ void k_set_fixed(int n, const int *__restrict__ dofs, const double *__restrict__ vals,
                            unsigned char *__restrict__ fixmask, double *__restrict__ vfix, int set)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int d = dofs[t];
    int vtx = d / 3, c = d % 3;
    if (set) {
        __hip_atomic_fetch_or(reinterpret_cast<unsigned int *>(fixmask) + (vtx >> 2), 1u << ((vtx & 3) * 8 + c),
                              __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        vfix[d] = vals[t];
    } else {
        fixmask[vtx] = 0;
        vfix[d] = 0.0;
    }
}

__global__ void k_masked_copy(int N, const double *__restrict__ vin, const unsigned char *__restrict__ fixmask,
                              double *__restrict__ xout)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * N) return;
    int vtx = t / 3, c = t - 3 * vtx;
    xout[t] = ((fixmask[vtx] >> c) & 1) ? 0.0 : vin[t];
}

__global__ void k_accumulate(int n, double *__restrict__ xs, double *__restrict__ xl)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    xs[t] += xl[t];
    xl[t] = 0.0;
}

__global__ void k_write_solution(int N, const double *__restrict__ xs, const unsigned char *__restrict__ fixmask,
                                 const double *__restrict__ vfix, double *__restrict__ v)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * N) return;
    int vtx = t / 3, c = t - 3 * vtx;
    v[t] = ((fixmask[vtx] >> c) & 1) ? vfix[t] : xs[t];
}

// reset: set the tolerances and clear the counters (iteration 0 starts with
// beta = 0); otherwise only clear `done` so the iteration can go on.
__global__ void k_scal_set(SolverScalars *s, double tol2, double delta2, int reset)
{
    if (reset) {
        s->tol2 = tol2;
        s->delta2 = delta2;
        s->iters = 0;
        s->first_iter = 0;
    }
    s->done = 0;
    s->reason = 0;
}

// The iteration k starts over with beta = 0 (preconditioner changed).
__global__ void k_scal_restart(SolverScalars *s, int k)
{
    s->first_iter = k;
    s->done = 0;
    s->reason = 0;
}

__global__ void k_scal_delta(SolverScalars *s, double delta2) { s->delta2 = delta2; }

inline int blocks_for(int64_t n, int t = 256) { return (int)((n + t - 1) / t); }

// Streaming copy (16-byte non-temporal loads and stores, four in flight per
// thread, grid stride): the known-size kernel that calibrates the memory counters
// and measures achievable HBM bandwidth.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__global__ void __launch_bounds__(256) k_stream_copy(int64_t n16, const u32x4 *__restrict__ src, u32x4 *__restrict__ dst)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    for (; i + 3 * stride < n16; i += 4 * stride) {
        u32x4 v[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
        for (int u = 0; u < 4; ++u) __builtin_nontemporal_store(v[u], dst + i + u * stride);
    }
    for (; i < n16; i += stride) __builtin_nontemporal_store(__builtin_nontemporal_load(src + i), dst + i);
}

// Streaming read (4-byte non-temporal loads, 16 in flight per thread, grid stride):
// the achievable read bandwidth that read-dominated kernels such as the SpMV are
// measured against.  The XOR of the data is stored only when it equals a sentinel,
// which keeps the loads live.
__global__ void __launch_bounds__(256) k_stream_read(int64_t n4, const unsigned *__restrict__ src,
                                                     unsigned *__restrict__ out)
{
    const int64_t stride = (int64_t)gridDim.x * blockDim.x;
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    unsigned acc = 0;
    for (; i + 15 * stride < n4; i += 16 * stride) {
        unsigned v[16];
#pragma unroll
        for (int u = 0; u < 16; ++u) v[u] = __builtin_nontemporal_load(src + i + u * stride);
#pragma unroll
        for (int u = 0; u < 16; ++u) acc ^= v[u];
    }
    for (; i < n4; i += stride) acc ^= src[i];
    if (acc == 0x9e3779b9u) out[0] = acc;
}

__global__ void k_fill(int n, double v, double *__restrict__ out)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) out[t] = v;
}

// SpMV launch variants (see k_pcg_spmv); the order is the one eolc_autotune reports.
// Columns: id, entries per trip, column prefetch, non-temporal loads, waves per SIMD,
// block layout, LDS-staged gathers.  The symmetric (1) and record (2) layouts apply
// to the FP32 difference form, and the FP64 product runs its SoA kernel for them;
// the pipelined SoA loop (3) serves both precisions; layout 4 is the symmetric
// records in either precision.  The staged variants (19-24) need every workgroup's
// vertex window within kStageVerts; their LDS (65 KB FP32, 131 KB FP64) allows two
// workgroups per CU in FP32 and one in FP64.  Layout 5 (25-28) is the patch layout
// in either precision (k_pcg_spmv_patch; 81 KB of LDS in FP32, 162 KB in FP64).
#define EOLC_SPMV_VARIANT_LIST(X)      \
    X(0, 1, false, false, 8, 0, false) \
    X(1, 1, true, false, 8, 0, false) \
    X(2, 1, true, true, 8, 0, false) \
    X(3, 1, false, true, 8, 0, false) \
    X(4, 1, true, true, 6, 0, false) \
    X(5, 2, true, true, 6, 0, false) \
    X(6, 2, true, false, 6, 0, false) \
    X(7, 2, true, true, 4, 0, false) \
    X(8, 1, true, true, 4, 0, false) \
    X(9, 1, false, false, 8, 1, false) \
    X(10, 1, false, false, 6, 1, false) \
    X(11, 1, false, true, 8, 1, false) \
    X(12, 1, false, false, 8, 2, false) \
    X(13, 1, false, true, 8, 2, false) \
    X(14, 1, false, false, 6, 2, false) \
    X(15, 1, false, false, 8, 3, false) \
    X(16, 1, false, true, 8, 3, false) \
    X(17, 1, false, true, 6, 3, false) \
    X(18, 1, false, false, 8, 4, false) \
    X(19, 1, false, true, 8, 0, true) \
    X(20, 1, false, true, 4, 0, true) \
    X(21, 1, false, false, 8, 1, true) \
    X(22, 1, false, false, 4, 4, true) \
    X(23, 1, false, true, 8, 3, true) \
    X(24, 1, false, true, 4, 3, true) \
    X(25, 1, false, false, 8, 5, false) \
    X(26, 2, false, false, 8, 5, false) \
    X(27, 1, false, false, 4, 5, false) \
    X(28, 2, false, false, 4, 5, false)

// layout column of a variant -> the FP32 array it reads
constexpr int lay_bit(int lay)
{
    return (lay == 1 || lay == 4) ? eolc::kLaySym : lay == 2 ? eolc::kLayRecord : lay == 5 ? eolc::kLayPatch
                                                                                           : eolc::kLaySoA;
}
#define EOLC_SPMV_LAYOUT_CASE(id, U, PF, NT, W, LAY, STG) \
    case id:                                         \
        return lay_bit(LAY);
#define EOLC_SPMV_LAYOUT64_CASE(id, U, PF, NT, W, LAY, STG) \
    case id:                                           \
        return LAY == 4 ? eolc::kLaySym64 : LAY == 5 ? eolc::kLayPatch64 : 0;

template <class T, int U, bool PF, bool NT, int W, int LAY, bool STG>
void launch_spmv_one(eolc_ctx *c, int k, int G, const T *A, const T *Mm, const T *z, const T *pold, T *pnew, T *q)
{
    const eolc::Topology &t = c->topo;
    constexpr bool diff = std::is_same<T, float>::value;
    if constexpr (LAY == 5) {
        using TR = T;
        const PatchLayout pl{c->pt_row, c->pt_cnt, c->pt_roff, c->pt_soff, c->pt_loff, c->pt_woff, c->pt_win,
                             c->pt_lo,  c->P32,    c->P64,     t.pt_slots, c->smask,   c->sellbase, c->scol,
                             c->A64s};
        hipLaunchKernelGGL((k_pcg_spmv_patch<T, TR, NT, W, U>), dim3(G), dim3(WG), 0, c->stream, k, pl, Mm, c->fixmask,
                           z, pold, pnew, q, c->scal, c->partials);
    } else {
        const SymLayout sym{c->supper, c->uoff, c->loff, c->lrec, c->S32u, c->S32a, c->S64u};
        hipLaunchKernelGGL((k_pcg_spmv<T, diff || LAY == 4, U, PF, NT, W, (diff || LAY >= 3) ? LAY : 0, STG>),
                           dim3(G), dim3(WG), 0, c->stream, k, t.N, t.nslices, c->slice_off, c->srow, c->slen, c->scol,
                           A, Mm, t.nbs, c->fixmask, z, pold, pnew, q, c->scal, c->partials, sym, c->wwin);
    }
}

template <class T>
void launch_spmv(eolc_ctx *c, int variant, int k, int G, const T *A, const T *Mm, const T *z, const T *pold, T *pnew,
                 T *q)
{
    // a staged or patch variant chosen for an earlier mesh that this mesh cannot run:
    // its twin with the same bits
    if (!eolc::spmv_usable(c->topo, variant)) variant = eolc::spmv_fallback(variant);
#define EOLC_SPMV_CASE(id, U, PF, NT, W, LAY, STG)                                      \
    case id:                                                                       \
        launch_spmv_one<T, U, PF, NT, W, LAY, STG>(c, k, G, A, Mm, z, pold, pnew, q); \
        break;
    switch (variant) {
        EOLC_SPMV_VARIANT_LIST(EOLC_SPMV_CASE)
    default:
        launch_spmv_one<T, 1, false, false, 8, 0, false>(c, k, G, A, Mm, z, pold, pnew, q);
    }
#undef EOLC_SPMV_CASE
}

template <class T, class TX>
void launch_iteration(eolc_ctx *c, int k, int G, const T *A, const T *Mm, const T *D, T *z, T *const p[2], T *q, T *r,
                      TX *x)
{
    const eolc::Topology &t = c->topo;
    const int pin = k & 1, pout = pin ^ 1;
    constexpr bool diff = std::is_same<T, float>::value;
    launch_spmv<T>(c, c->spmv_variant[diff ? 0 : 1], k, G, A, Mm, z, p[pin], p[pout], q);
    if (c->ml_on) {
        hipLaunchKernelGGL((k_pcg_update_flat<T, TX>), dim3(eolc::pcg_grid_phys(G)), dim3(WG), 0, c->stream, k, G,
                           3 * t.N, p[pout], q, x, r, c->scal, c->partials);
        (void)eolc::ml_apply(c, diff, (k + 1) & 1, 1);   // launch errors surface at the chunk's hipGetLastError
    } else {
        hipLaunchKernelGGL((k_pcg_update<T, TX, false>), dim3(G), dim3(WG), 0, c->stream, k, t.N, D, p[pout], q, x,
                           r, z, c->scal, c->partials);
    }
}

// z = P r for the residual r64 (FP64): mixed -> r32, z32; FP64 -> z64 (r64 in place);
// r.z, r.r -> PB[parity].
hipError_t precondition(eolc_ctx *c, bool mixed, int parity)
{
    const int N = c->topo.N;
    const int G = eolc::pcg_grid(N);
    if (mixed) {
        if (c->ml_on)
            hipLaunchKernelGGL((k_precond_init<float, true>), dim3(G), dim3(WG), 0, c->stream, parity, N, c->D32,
                               c->r64, c->r32, c->z32, c->partials);
        else
            hipLaunchKernelGGL((k_precond_init<float, false>), dim3(G), dim3(WG), 0, c->stream, parity, N, c->D32,
                               c->r64, c->r32, c->z32, c->partials);
    } else {
        if (c->ml_on)
            hipLaunchKernelGGL((k_precond_init<double, true>), dim3(G), dim3(WG), 0, c->stream, parity, N, c->D64,
                               c->r64, c->r64, c->z64, c->partials);
        else
            hipLaunchKernelGGL((k_precond_init<double, false>), dim3(G), dim3(WG), 0, c->stream, parity, N, c->D64,
                               c->r64, c->r64, c->z64, c->partials);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && c->ml_on) e = eolc::ml_apply(c, mixed, parity, 0);
    return e;
}

void launch_resid64(eolc_ctx *c, int G, const double *xin, const double *b, double *out)
{
    const eolc::Topology &t = c->topo;
    hipLaunchKernelGGL(k_resid64, dim3(G), dim3(WG), 0, c->stream, t.N, t.nslices, c->slice_off, c->srow, c->slen,
                       c->scol, c->A64, t.nbs, c->fixmask, xin, b, out, c->partials + 5 * eolc::kMaxPartials);
}

}  // namespace

namespace eolc {

#define EOLC_SPMV_STAGED_CASE(id, U, PF, NT, W, LAY, STG) \
    case id:                                             \
        return STG;
bool spmv_staged(int variant)
{
    switch (variant) {
        EOLC_SPMV_VARIANT_LIST(EOLC_SPMV_STAGED_CASE)
    default:
        return false;
    }
}

// 19, 20 read the SoA arrays like 3; 21 the symmetric FP32 records like 9; 22 the
// symmetric records of both precisions like 18; 23, 24 the SoA arrays like 16
int spmv_unstaged(int variant)
{
    switch (variant) {
    case 19:
    case 20: return 3;
    case 21: return 9;
    case 22: return 18;
    case 23:
    case 24: return 16;
    default: return variant;
    }
}

#define EOLC_SPMV_PATCH_CASE(id, U, PF, NT, W, LAY, STG) \
    case id:                                            \
        return LAY == 5;
bool spmv_patch(int variant)
{
    switch (variant) {
        EOLC_SPMV_VARIANT_LIST(EOLC_SPMV_PATCH_CASE)
    default:
        return false;
    }
}

bool spmv_usable(const Topology &t, int variant)
{
    if (spmv_staged(variant) && t.wmax > kStageVerts) return false;
    if (spmv_patch(variant) && !t.pt_ok) return false;
    return true;
}

// the staged variants' unstaged twins; the patch variants' symmetric-record twin (18:
// the same sums in the same order, in both precisions)
int spmv_fallback(int variant) { return spmv_patch(variant) ? 18 : spmv_unstaged(variant); }

int fp32_layout(int variant)
{
    switch (variant) {
        EOLC_SPMV_VARIANT_LIST(EOLC_SPMV_LAYOUT_CASE)
    default:
        return kLaySoA;
    }
}

// FP64 variants read A64 (always written) and, for layout 4, the FP64 records
int fp64_layout(int variant)
{
    switch (variant) {
        EOLC_SPMV_VARIANT_LIST(EOLC_SPMV_LAYOUT64_CASE)
    default:
        return 0;
    }
}

// Physical workgroups of the row kernels that can take a second logical workgroup
// (the patch SpMV, the multilevel update and prolongation): when G exceeds a
// multiple of 512 (two resident 1,024-thread workgroups on each of the 256 CUs; four
// rounds of one for the FP64 patch SpMV) by at most 64, the excess rides on the first
// workgroups, which start first, instead of running as a nearly empty last round.
int pcg_grid_phys(int G)
{
    const int E = G % 512;
    return (G > 512 && E > 0 && E <= 64) ? G - E : G;
}

int pcg_grid(int N)
{
    const int nslices = (N + kSlice - 1) / kSlice;
    int g = (nslices + WG / 64 - 1) / (WG / 64);
    g = (g + 7) & ~7;
    return g;
}

hipError_t pcg_setup(eolc_ctx *c, bool /*mixed*/, int precond_kind, bool ml_linear)
{
    int N = c->topo.N;
    hipLaunchKernelGGL(k_block_jacobi, dim3(blocks_for(N)), dim3(256), 0, c->stream, N, c->diagpos, c->A64,
                       c->topo.nbs, c->fixmask, precond_kind, c->D64, c->D32);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess && c->ml_on) e = ml_setup(c, ml_linear);
    return e;
}

hipError_t fixed_from_vold(eolc_ctx *c, int n, const int *d_dofs, double *d_vals)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_fixed_from_vold, dim3(blocks_for(n)), dim3(256), 0, c->stream, n, d_dofs, c->v, d_vals);
    return hipGetLastError();
}

hipError_t set_fixed(eolc_ctx *c, int n, const int *d_dofs, const double *d_vals, int set)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_set_fixed, dim3(blocks_for(n)), dim3(256), 0, c->stream, n, d_dofs, d_vals, c->fixmask,
                       c->vfix, set);
    return hipGetLastError();
}

// b_red = rhs - A vfix (-> bnorm2); x0 = masked v_old; r = b_red - A x0 (-> rmax2);
// z, r.z, r.r -> PB[0].
hipError_t pcg_init(eolc_ctx *c, bool mixed)
{
    const int N = c->topo.N;
    const int G = pcg_grid(N);
    const double *pr = c->partials + 5 * kMaxPartials;
    launch_resid64(c, G, c->vfix, c->rhs, c->b64);
    hipLaunchKernelGGL(k_reduce_to, dim3(1), dim3(WG), 0, c->stream, G, pr, &c->scal->bnorm2);
    hipLaunchKernelGGL(k_masked_copy, dim3(blocks_for(3 * N)), dim3(256), 0, c->stream, N, c->v, c->fixmask, c->xs64);
    launch_resid64(c, G, c->xs64, c->b64, c->r64);
    hipLaunchKernelGGL(k_reduce_to, dim3(1), dim3(WG), 0, c->stream, G, pr, &c->scal->rmax2);
    hipError_t e = hipMemsetAsync(mixed ? (void *)c->p32[0] : (void *)c->p64[0], 0,
                                  (size_t)3 * N * (mixed ? sizeof(float) : sizeof(double)), c->stream);
    if (e != hipSuccess) return e;
    if (mixed) {
        e = hipMemsetAsync(c->xl64, 0, (size_t)3 * N * sizeof(double), c->stream);
        if (e != hipSuccess) return e;
    }
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    return precondition(c, mixed, 0);
}

hipError_t pcg_chunk(eolc_ctx *c, bool mixed, int k0, int iters)
{
    const int G = pcg_grid(c->topo.N);
    for (int k = k0; k < k0 + iters; ++k) {
        if (mixed)
            launch_iteration<float, double>(c, k, G, c->S32, c->M32, c->D32, c->z32, c->p32, c->q32, c->r32, c->xl64);
        else
            launch_iteration<double, double>(c, k, G, c->A64, c->M64, c->D64, c->z64, c->p64, c->q64, c->r64, c->xs64);
    }
    return hipGetLastError();
}

// Mixed mode: fold the FP32 correction into the FP64 solution and form the true
// FP64 residual (Clark et al., reliable updates); ||r||^2 -> scal->rmax2.
hipError_t pcg_reliable_update(eolc_ctx *c)
{
    const int N = c->topo.N;
    const int G = pcg_grid(N);
    hipLaunchKernelGGL(k_accumulate, dim3(blocks_for(3 * N)), dim3(256), 0, c->stream, 3 * N, c->xs64, c->xl64);
    launch_resid64(c, G, c->xs64, c->b64, c->r64);
    hipLaunchKernelGGL(k_reduce_to, dim3(1), dim3(WG), 0, c->stream, G, c->partials + 5 * kMaxPartials,
                       &c->scal->rmax2);
    return hipGetLastError();
}

// A mixed solve that stalled with linear coarse fields restarts in FP64 from its
// initial guess: x = the resident v on the free DOFs, the FP32 correction dropped,
// r = b - A x; ||r||^2 -> scal->rmax2.  (Continued from such a stalled iterate, the
// FP64 iteration on a 1024^2 state with collapsed elements stagnated at a true
// residual of 1e-7, where the same preconditioner from the initial guess converged in
// 2,242 iterations.)
hipError_t pcg_restart_from_guess(eolc_ctx *c)
{
    const int N = c->topo.N;
    const int G = pcg_grid(N);
    hipLaunchKernelGGL(k_masked_copy, dim3(blocks_for(3 * N)), dim3(256), 0, c->stream, N, c->v, c->fixmask, c->xs64);
    hipError_t e = hipMemsetAsync(c->xl64, 0, (size_t)3 * N * sizeof(double), c->stream);
    if (e != hipSuccess) return e;
    launch_resid64(c, G, c->xs64, c->b64, c->r64);
    hipLaunchKernelGGL(k_reduce_to, dim3(1), dim3(WG), 0, c->stream, G, c->partials + 5 * kMaxPartials,
                       &c->scal->rmax2);
    return hipGetLastError();
}

// Continue the iteration at index k from the replaced residual, keeping p.
hipError_t pcg_restart_inner(eolc_ctx *c, bool mixed, int k) { return precondition(c, mixed, k & 1); }

// z = P r of the current mode for r in r64 (the preconditioner-application entry point)
hipError_t pcg_precondition(eolc_ctx *c, bool mixed) { return precondition(c, mixed, 0); }

// FP64 mode: the recursive residual met the tolerance; replace it with the true
// FP64 residual b - A x of the current iterate (the search direction is kept, as in
// the mixed mode's reliable updates); ||r||^2 -> scal->rmax2.
hipError_t pcg_replace_residual64(eolc_ctx *c)
{
    const int G = pcg_grid(c->topo.N);
    launch_resid64(c, G, c->xs64, c->b64, c->r64);
    hipLaunchKernelGGL(k_reduce_to, dim3(1), dim3(WG), 0, c->stream, G, c->partials + 5 * kMaxPartials,
                       &c->scal->rmax2);
    return hipGetLastError();
}

// The solution with its fixed values -> out (c->v on success; xs64 itself, in
// place, when the solve failed and the resident velocity is kept).
hipError_t pcg_finish(eolc_ctx *c, double *out)
{
    const int N = c->topo.N;
    hipLaunchKernelGGL(k_write_solution, dim3(blocks_for(3 * N)), dim3(256), 0, c->stream, N, c->xs64, c->fixmask,
                       c->vfix, out);
    return hipGetLastError();
}

hipError_t scal_restart(eolc_ctx *c, int k)
{
    hipLaunchKernelGGL(k_scal_restart, dim3(1), dim3(1), 0, c->stream, c->scal, k);
    return hipGetLastError();
}

hipError_t scal_delta(eolc_ctx *c, double delta2)
{
    hipLaunchKernelGGL(k_scal_delta, dim3(1), dim3(1), 0, c->stream, c->scal, delta2);
    return hipGetLastError();
}

hipError_t scal_set(eolc_ctx *c, double tol2, double delta2, int reset)
{
    hipLaunchKernelGGL(k_scal_set, dim3(1), dim3(1), 0, c->stream, c->scal, tol2, delta2, reset);
    return hipGetLastError();
}

// ||rhs - A v||^2 over free DOFs (v = a solution incl. fixed values) -> scal->true_rr
hipError_t spmv_true_residual(eolc_ctx *c, const double *v)
{
    const int N = c->topo.N;
    const int G = pcg_grid(N);
    launch_resid64(c, G, v, c->rhs, c->r64);
    hipLaunchKernelGGL(k_reduce_to, dim3(1), dim3(WG), 0, c->stream, G, c->partials + 5 * kMaxPartials,
                       &c->scal->true_rr);
    return hipGetLastError();
}

// Timed launches of the streaming copy (read_only = 0: `bytes` read and `bytes`
// written per launch) or of the streaming read (`bytes` read per launch).
hipError_t stream_copy(eolc_ctx *c, size_t bytes, int reps, float *ms, int read_only)
{
    const int64_t n16 = (int64_t)(bytes / 16);
    u32x4 *a = nullptr, *b = nullptr;
    hipError_t e = hipMalloc((void **)&a, n16 * 16);
    if (e != hipSuccess) return e;
    e = hipMalloc((void **)&b, n16 * 16);
    if (e != hipSuccess) {
        (void)hipFree(a);
        return e;
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    const int grid = read_only ? 65536 : 256 * 32;   // fewer workgroups reach 10-20% less
    auto launch = [&] {
        if (read_only)
            hipLaunchKernelGGL(k_stream_read, dim3(grid), dim3(256), 0, c->stream, 4 * n16,
                               reinterpret_cast<const unsigned *>(a), reinterpret_cast<unsigned *>(b));
        else
            hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(256), 0, c->stream, n16, a, b);
    };
    launch();   // first touch
    (void)hipEventRecord(e0, c->stream);
    for (int r = 0; r < reps; ++r) launch();
    (void)hipEventRecord(e1, c->stream);
    e = hipEventSynchronize(e1);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    return e;
}

// Back-to-back launches of one SpMV variant (-1: the context's current choice) for
// timing (iteration index 1, `done` cleared first).  Leaves the solver state undefined.
hipError_t time_spmv(eolc_ctx *c, bool mixed, int variant, int reps, float *ms)
{
    const int N = c->topo.N;
    const int G = pcg_grid(N);
    if (variant < 0) variant = c->spmv_variant[mixed ? 0 : 1];
    {
        const hipError_t el = mixed ? ensure_fp32_layout(c, variant) : ensure_fp64_layout(c, variant);
        if (el != hipSuccess) return el;
    }
    hipEvent_t a, b;
    hipError_t e = hipEventCreate(&a);
    if (e != hipSuccess) return e;
    e = hipEventCreate(&b);
    if (e != hipSuccess) {
        (void)hipEventDestroy(a);
        return e;
    }
    e = scal_set(c, 0.0, 0.0, 1);
    if (e != hipSuccess) {
        (void)hipEventDestroy(a);
        (void)hipEventDestroy(b);
        return e;
    }
    hipLaunchKernelGGL(k_fill, dim3(blocks_for(4 * kMaxPartials)), dim3(256), 0, c->stream, 4 * kMaxPartials, 1.0,
                       c->partials + kMaxPartials);
    // one untimed launch (first-touch, instruction cache)
    if (mixed)
        launch_spmv<float>(c, variant, 1, G, c->S32, c->M32, c->z32, c->p32[1], c->p32[0], c->q32);
    else
        launch_spmv<double>(c, variant, 1, G, c->A64, c->M64, c->z64, c->p64[1], c->p64[0], c->q64);
    e = hipEventRecord(a, c->stream);
    for (int r = 0; r < reps && e == hipSuccess; ++r) {
        if (mixed)
            launch_spmv<float>(c, variant, 1, G, c->S32, c->M32, c->z32, c->p32[1], c->p32[0], c->q32);
        else
            launch_spmv<double>(c, variant, 1, G, c->A64, c->M64, c->z64, c->p64[1], c->p64[0], c->q64);
    }
    if (e == hipSuccess) e = hipEventRecord(b, c->stream);
    if (e == hipSuccess) e = hipEventSynchronize(b);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, a, b);
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    return e == hipSuccess ? hipGetLastError() : e;
}

}  // namespace eolc
