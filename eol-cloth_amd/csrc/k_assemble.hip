// Deterministic, atomic-free assembly (gfx950).
//
// Replaces the triplet push + setFromTriplets of Forces::fill (Forces.cpp:103-125,
// 522-539, 928-929).  The host builds, once per topology, the list of element
// slots feeding every 3x3 block in the order the reference inserts its triplets
// (faces in face order, then hinges in edge order), so each block is summed in
// the same order Eigen's collapseDuplicates folds duplicates -- without atomics.
//
//   k_assemble : one thread per SELL entry (padding entries have no slots and
//                come out 0); MDK block = sum of its staged slots (a slot
//                flagged "transposed" is the (w,v) image of an element's (v,w)
//                block); the mass scalar sums the triangle slots.  Writes the FP64
//                SoA values and, for the mixed solver, the FP32 difference form
//                S'_ij = dampingB h^2 K_ij + m_ij I off the diagonal (0 on it), with
//                which (A p)_i = Mbar_i p_i + sum_{j != i} S'_ij (p_j - p_i),
//                Mbar_i = sum_j m_ij (since sum_j m_ij p_j = Mbar_i p_i +
//                sum_{j != i} m_ij (p_j - p_i)).  The FP32 values go only to the
//                layouts in `lay` (SoA, symmetric upper records, per-entry records,
//                patch records in FP32 or FP64): the one the chosen SpMV variant
//                reads, the others on demand.
//   k_rhs      : one thread per vertex (SELL storage order); f_i = sum of its element force slots
//                (faces, then hinges when the bending force is included),
//                rhs_i = sum_j m_ij v_j + h f_i, the right-hand side M v + h f of
//                Cloth::solve (Cloth.cpp:345), and the FP32 row mass Mbar_i.
#include <algorithm>

#include "eolc_internal.h"

namespace {

__global__ void __launch_bounds__(256)
k_assemble(int64_t nbs, int ntrislots, const int *__restrict__ cptr, const int *__restrict__ clist,
           const int *__restrict__ entry_is_diag, const double *__restrict__ stageK, const double *__restrict__ stageM,
           double *__restrict__ A64, int lay, float *__restrict__ S32, const int *__restrict__ e2u,
           const int *__restrict__ scol, float *__restrict__ S32u, float *__restrict__ S32a,
           double *__restrict__ S64u, const int *__restrict__ e2r, const int *__restrict__ e2l,
           float *__restrict__ P32, double *__restrict__ P64, double tau, unsigned char *__restrict__ stiff,
           double *__restrict__ A64s, double *__restrict__ Mval)
{
    int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nbs) return;
    double acc[9], stf[9];
#pragma unroll
    for (int c = 0; c < 9; ++c) { acc[c] = 0.0; stf[c] = 0.0; }
    double m = 0.0;
    const int e0 = cptr[b], e1 = cptr[b + 1];
    for (int e = e0; e < e1; ++e) {
        const int code = clist[e];
        const int slot = code >> 1;
        const double *sk = stageK + (size_t)9 * slot;
        const double mm = slot < ntrislots ? stageM[slot] : 0.0;
#pragma unroll
        for (int j = 0; j < 3; ++j)
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double kv = (code & 1) ? sk[3 * k + j] : sk[3 * j + k];
                // the triplet value of the reference: M_vw + dampingB h^2 K_vw (hinges: no mass)
                acc[3 * j + k] += (slot < ntrislots ? (j == k ? mm : 0.0) + kv : kv);
                stf[3 * j + k] += kv;
            }
        m += mm;
    }
#pragma unroll
    for (int c = 0; c < 9; ++c) A64[c * nbs + b] = acc[c];
    if (lay) {
        const bool diag = entry_is_diag[b] != 0;
        float sv[9];
#pragma unroll
        for (int c = 0; c < 9; ++c) sv[c] = diag ? 0.0f : (float)(stf[c] + (c % 4 == 0 ? m : 0.0));
        if (lay & eolc::kLaySoA) {
#pragma unroll
            for (int c = 0; c < 9; ++c) S32[c * nbs + b] = sv[c];
        }
        if (lay & eolc::kLayRecord) {   // record layout: every entry
            float *ra = S32a + (size_t)10 * b;
            ra[0] = __int_as_float(scol[b]);
#pragma unroll
            for (int c = 0; c < 9; ++c) ra[1 + c] = sv[c];
        }
        const int u = (lay & eolc::kLaySym) ? e2u[b] : -1;
        if (u >= 0) {   // upper block: its record in the symmetric layout
            float *rec = S32u + (size_t)10 * u;
            rec[0] = __int_as_float(scol[b]);
#pragma unroll
            for (int c = 0; c < 9; ++c) rec[1 + c] = sv[c];
        }
        const int u64 = (lay & eolc::kLaySym64) ? e2u[b] : -1;
        if (u64 >= 0) {   // the same record in FP64 (the column's bits in the first slot)
            double *rec = S64u + (size_t)10 * u64;
            rec[0] = __longlong_as_double((long long)scol[b]);
#pragma unroll
            for (int c = 0; c < 9; ++c) rec[1 + c] = diag ? 0.0 : stf[c] + (c % 4 == 0 ? m : 0.0);
        }
        // patch layout: upper blocks and cross lower blocks, in either precision, the
        // column as its index in the patch's window
        const int pr = (lay & (eolc::kLayPatch | eolc::kLayPatch64)) ? e2r[b] : -1;
        if (pr >= 0 && (lay & eolc::kLayPatch)) {
            float *rec = P32 + (size_t)10 * pr;
            rec[0] = __int_as_float(e2l[b]);
#pragma unroll
            for (int c = 0; c < 9; ++c) rec[1 + c] = sv[c];
        }
        if (lay & eolc::kLayPatch) {
            // a block too stiff for FP32 (k_stiff_mask moves it to FP64 in both rows:
            // (i, j) and (j, i) have the same largest entry, so both are flagged)
            double mx = 0.0;
#pragma unroll
            for (int c = 0; c < 9; ++c) mx = fmax(mx, fabs(stf[c] + (c % 4 == 0 ? m : 0.0)));
            const bool st = !diag && tau > 0.0 && mx > tau;
            stiff[b] = st ? 1 : 0;
            if (st)   // the FP64 values once more, contiguous: the SpMV reads them as one 72-byte piece
#pragma unroll
                for (int c = 0; c < 9; ++c) A64s[9 * b + c] = acc[c];
        }
        if (pr >= 0 && (lay & eolc::kLayPatch64)) {
            double *rec = P64 + (size_t)10 * pr;
            rec[0] = __longlong_as_double((long long)e2l[b]);
#pragma unroll
            for (int c = 0; c < 9; ++c) rec[1 + c] = diag ? 0.0 : stf[c] + (c % 4 == 0 ? m : 0.0);
        }
    }
    Mval[b] = m;
}

// One thread per SELL storage position: the row's stiff-entry mask (a row holds at
// most 64 off-diagonal blocks where the patch layout applies) and the FP32 patch records of its stiff entries
// zeroed, so that the FP32 product skips them and the patch SpMV applies them from
// their contiguous FP64 copies (A64s, written by k_assemble) in both rows.
__global__ void __launch_bounds__(256)
k_stiff_mask(int N, const int *__restrict__ slice_off, const int *__restrict__ srow, const int *__restrict__ slen,
             const unsigned char *__restrict__ stiff, const int *__restrict__ e2r, float *__restrict__ P32,
             unsigned long long *__restrict__ smask)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N) return;
    const int base = slice_off[g / eolc::kSlice] + (g % eolc::kSlice);
    const int n = min(slen[g] - 1, 64);   // the diagonal block is the row's last entry
    unsigned long long m = 0;
    for (int k = 0; k < n; ++k) {
        const int e = base + k * eolc::kSlice;
        if (!stiff[e]) continue;
        m |= 1ull << k;
        // the entry's record, if it has one (a lower block inside a patch has none: its
        // upper twin's record is zeroed by that row)
        const int r = e2r[e];
        if (r < 0) continue;
        float *rec = P32 + (size_t)10 * r;
#pragma unroll
        for (int c = 1; c < 10; ++c) rec[c] = 0.0f;
    }
    smask[srow[g]] = m;
}

// one thread per SELL storage position (row srow[g])
__global__ void __launch_bounds__(256)
k_rhs(int N, int ntriforce, int include_bending, double h, const int *__restrict__ fptr,
      const int *__restrict__ flist, const double *__restrict__ stageF, const int *__restrict__ slice_off,
      const int *__restrict__ srow, const int *__restrict__ slen, const int *__restrict__ scol,
      const double *__restrict__ Mval, const double *__restrict__ v, double *__restrict__ f, double *__restrict__ rhs,
      float *__restrict__ Mrow, double *__restrict__ Mrow64)
{
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= N) return;
    const int i = srow[g], len = slen[g];
    const int base = slice_off[g / eolc::kSlice] + (g % eolc::kSlice);
    double f0 = 0.0, f1 = 0.0, f2 = 0.0;
    for (int e = fptr[i]; e < fptr[i + 1]; ++e) {
        const int slot = flist[e];
        if (slot >= ntriforce && !include_bending) continue;
        const double *sf = stageF + (size_t)3 * slot;
        f0 += sf[0]; f1 += sf[1]; f2 += sf[2];
    }
    f[3 * i] = f0; f[3 * i + 1] = f1; f[3 * i + 2] = f2;
    // M v with the columns in increasing order (SELL keeps the CSR order within a row)
    double y0 = 0.0, y1 = 0.0, y2 = 0.0, msum = 0.0;
    for (int k = 0; k < len; ++k) {
        const int e = base + k * eolc::kSlice;
        const int j = scol[e];
        const double m = Mval[e];
        y0 += m * v[3 * j]; y1 += m * v[3 * j + 1]; y2 += m * v[3 * j + 2];
        msum += m;
    }
    if (Mrow) Mrow[i] = (float)msum;
    if (Mrow64) Mrow64[i] = msum;
    rhs[3 * i] = y0 + h * f0;
    rhs[3 * i + 1] = y1 + h * f1;
    rhs[3 * i + 2] = y2 + h * f2;
}

inline int blocks_for(int64_t n, int t = 256) { return (int)((n + t - 1) / t); }

}  // namespace

namespace eolc {

hipError_t launch_assemble(eolc_ctx *c, int lay)
{
    const int64_t nb = c->topo.nbs;
    c->s32_valid = 0;
    if (nb == 0) return hipSuccess;
    if (!c->topo.pt_ok) lay &= ~(eolc::kLayPatch | eolc::kLayPatch64);
    // stiff threshold of the FP32 patch records: stiff_rel x dampingB h^2 E (E the
    // largest Young's modulus on the context)
    const double emax = std::max(c->E, c->batch.B > 0 ? c->e_batch_max : 0.0);
    const double tau = c->stiff_rel > 0.0 ? c->stiff_rel * c->dampB * c->h_last * c->h_last * emax : 0.0;
    hipLaunchKernelGGL(k_assemble, dim3(blocks_for(nb)), dim3(256), 0, c->stream, nb, 6 * c->topo.F, c->cptr, c->clist,
                       c->entry_is_diag, c->stageK, c->stageM, c->A64, lay, c->S32, c->e2u, c->scol, c->S32u,
                       c->S32a, c->S64u, c->pt_e2r, c->pt_e2l, c->P32, c->P64, tau, c->stiff, c->A64s, c->Mval);
    if (lay & eolc::kLayPatch)
        hipLaunchKernelGGL(k_stiff_mask, dim3(blocks_for(c->topo.N)), dim3(256), 0, c->stream, c->topo.N,
                           c->slice_off, c->srow, c->slen, c->stiff, c->pt_e2r, c->P32, c->smask);
    const hipError_t e = hipGetLastError();
    if (e == hipSuccess) c->s32_valid = lay;
    return e;
}

// The layout that SpMV variant `variant` reads, written from the staged element
// blocks of the last assembly if that assembly did not write it (same values, same
// bits: the fold is the same; the FP64 values are rewritten unchanged).
static hipError_t ensure_layout(eolc_ctx *c, int need)
{
    if ((c->s32_valid & need) == need) return hipSuccess;
    const int have = c->s32_valid;
    const hipError_t e = launch_assemble(c, need);
    if (e == hipSuccess) c->s32_valid = have | need;
    return e;
}

// (a variant this mesh cannot run is replaced by its twin, as the SpMV launch does)
hipError_t ensure_fp32_layout(eolc_ctx *c, int variant)
{
    return ensure_layout(c, fp32_layout(spmv_usable(c->topo, variant) ? variant : spmv_fallback(variant)));
}
hipError_t ensure_fp64_layout(eolc_ctx *c, int variant)
{
    return ensure_layout(c, fp64_layout(spmv_usable(c->topo, variant) ? variant : spmv_fallback(variant)));
}

hipError_t launch_rhs(eolc_ctx *c, double h)
{
    const int N = c->topo.N;
    hipLaunchKernelGGL(k_rhs, dim3(blocks_for(N)), dim3(256), 0, c->stream, N, 3 * c->topo.F,
                       (c->flags & EOLC_OPT_INCLUDE_BENDING_FORCE) ? 1 : 0, h, c->fptr, c->flist, c->stageF,
                       c->slice_off, c->srow, c->slen, c->scol, c->Mval, c->v, c->f, c->rhs, c->M32, c->M64);
    return hipGetLastError();
}

}  // namespace eolc
