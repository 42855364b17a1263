// Learned block preconditioner for the velocity solve (gfx950).  No counterpart in
// the reference: it replaces the block-Jacobi inverse D_i^-1 used by the PCG with
//
//     P_i = G_i G_i^T,   G_i = L_i (I + T_i),   L_i L_i^T = D_i^-1 (Cholesky),
//
// T_i lower-triangular from 0.5 tanh(o_i), where o_i (6 numbers per vertex) is the
// output of a message-passing network over the vertex graph of M + h^2 K.  P_i is
// SPD for every output and equals block-Jacobi at o = 0.  Fixed DOFs keep zero rows
// and columns (z stays 0 there).
//
// Network (features 16 -> 64, three message-passing layers 128 -> 64, 64 -> 6):
//   h0 = relu(X W_in + b_in);  a_l = sum_j w_ij h_l[j];  h_l+1 = relu([h_l | a_l] W_l + b_l);
//   o  = h3 W_out + b_out.
// w_ij = |A_ij|_F / sum_k |A_ik|_F over the off-diagonal blocks of row i.  X, h_l,
// a_l, the backward pre-activation gradients and the weight copies used by the
// matrix units are stored in bfloat16 (round to nearest even); biases, outputs,
// gradients and the Adam state in FP32; all products accumulate in FP32 on
// v_mfma_f32_16x16x16_bf16.  The loss and everything touching A stay in FP64.
//
// Training loss on the current system for an error vector e (r = A e):
//     z = P r,   loss = 1 - (z.r)^2 / ((z.A z)(r.e))
// (with eolc_learned_set_multilevel: z = P r + the multilevel coarse corrections of r,
// the preconditioner the blocks are used in; the gradient flows through P r only)
// -- the A-norm error left after one optimally scaled step along z relative to
// the initial one (0 <= loss <= 1 by Cauchy-Schwarz).  Every reduction over
// vertices goes through per-workgroup partials summed in a fixed order, so the
// loss, the gradient and the trained weights are bitwise reproducible.
//
// Matrix-unit operand layout of the 16x16x16 bf16 instruction (wave64): lane l
// supplies A[l%16][4(l/16)+t] and B[4(l/16)+t][l%16] (t = 0..3) and receives
// D[4(l/16)+t][l%16].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "eolc_internal.h"

namespace {

typedef short short4v __attribute__((ext_vector_type(4)));
typedef float float4v __attribute__((ext_vector_type(4)));

constexpr int HID = eolc::kLpHidden;   // 64
constexpr int NF = eolc::kLpFeat;      // 16
constexpr int NO = eolc::kLpOut;       // 6
constexpr int RT = 256;                // threads of the per-vertex kernels
constexpr int SL = eolc::kSlice;

inline int blocks_for(int64_t n, int t = RT) { return (int)((n + t - 1) / t); }

__device__ __forceinline__ unsigned short f2bf(float f)
{
    unsigned u = __float_as_uint(f);
    if ((u & 0x7F800000u) == 0x7F800000u) return (unsigned short)((u >> 16) | ((u & 0xFFFFu) ? 0x40u : 0u));
    u += 0x7FFFu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float(((unsigned)b) << 16); }

// Fixed-order block sum of one double per thread (RT threads); result in thread 0.
__device__ __forceinline__ double block_sum(double v, double *s)
{
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = RT / 2; o > 0; o >>= 1) {
        if (threadIdx.x < o) s[threadIdx.x] += s[threadIdx.x + o];
        __syncthreads();
    }
    return s[0];
}

// ---------------------------------------------------------------------------
// 3x3 helpers.  Blocks of the reduced system: fixed DOFs' rows and columns are
// those of the identity (diagonal block) or zero (off-diagonal blocks).
__device__ __forceinline__ void masked_block(const double *__restrict__ A, int64_t nbs, int64_t e, unsigned fr,
                                             unsigned fc, bool diag, double b[9])
{
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        const int r = c / 3, s = c % 3;
        double v = A[c * nbs + e];
        if (((fr >> r) & 1) || ((fc >> s) & 1)) v = (diag && r == s) ? 1.0 : 0.0;
        b[c] = v;
    }
}

__device__ __forceinline__ double fro(const double b[9])
{
    double s = 0.0;
#pragma unroll
    for (int c = 0; c < 9; ++c) s += b[c] * b[c];
    return sqrt(s);
}

// Per-vertex factors: D^-1, L (Cholesky of D^-1), G = L (I + T), th = tanh(o),
// L and G as full 3x3 (zeros above the diagonal).  ok = false when D^-1 is not SPD
// or the outputs are not finite; the block then falls back to block-Jacobi
// (P = D^-1) and contributes no gradient.
struct LpBlock {
    double Di[9];
    double L[3][3], G[3][3];
    double th[6];
    bool ok;
};

__device__ void lp_block(const double D[9], const float *__restrict__ o, LpBlock &b)
{
    const double c00 = D[4] * D[8] - D[5] * D[7], c01 = D[5] * D[6] - D[3] * D[8], c02 = D[3] * D[7] - D[4] * D[6];
    const double det = D[0] * c00 + D[1] * c01 + D[2] * c02;
    if (det != 0.0 && isfinite(det)) {
        const double id = 1.0 / det;
        b.Di[0] = c00 * id; b.Di[1] = (D[2] * D[7] - D[1] * D[8]) * id; b.Di[2] = (D[1] * D[5] - D[2] * D[4]) * id;
        b.Di[3] = c01 * id; b.Di[4] = (D[0] * D[8] - D[2] * D[6]) * id; b.Di[5] = (D[2] * D[3] - D[0] * D[5]) * id;
        b.Di[6] = c02 * id; b.Di[7] = (D[1] * D[6] - D[0] * D[7]) * id; b.Di[8] = (D[0] * D[4] - D[1] * D[3]) * id;
    } else {
#pragma unroll
        for (int c = 0; c < 9; ++c) b.Di[c] = (c % 4 == 0) ? 1.0 : 0.0;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) b.L[r][c] = b.G[r][c] = 0.0;
#pragma unroll
    for (int k = 0; k < 6; ++k) b.th[k] = 0.0;
    const double *a = b.Di;
    b.ok = false;
    const double l00 = a[0] > 0.0 ? sqrt(a[0]) : 0.0;
    if (!(l00 > 0.0)) return;
    const double l10 = a[3] / l00, l20 = a[6] / l00;
    const double s11 = a[4] - l10 * l10;
    if (!(s11 > 0.0)) return;
    const double l11 = sqrt(s11);
    const double l21 = (a[7] - l20 * l10) / l11;
    const double s22 = a[8] - l20 * l20 - l21 * l21;
    if (!(s22 > 0.0)) return;
    const double l22 = sqrt(s22);
    bool fin = true;
#pragma unroll
    for (int k = 0; k < 6; ++k) {
        const double ok = (double)o[k];
        fin = fin && isfinite(ok);
        b.th[k] = tanh(ok);
    }
    if (!fin) return;
    b.L[0][0] = l00; b.L[1][0] = l10; b.L[1][1] = l11; b.L[2][0] = l20; b.L[2][1] = l21; b.L[2][2] = l22;
    // U = I + T, T[r][c] = 0.5 th_k for (r, c) = (0,0) (1,0) (1,1) (2,0) (2,1) (2,2)
    const double U[3][3] = {{1.0 + 0.5 * b.th[0], 0.0, 0.0},
                            {0.5 * b.th[1], 1.0 + 0.5 * b.th[2], 0.0},
                            {0.5 * b.th[3], 0.5 * b.th[4], 1.0 + 0.5 * b.th[5]}};
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double s = 0.0;
#pragma unroll
            for (int k = 0; k < 3; ++k) s += b.L[r][k] * U[k][c];
            b.G[r][c] = s;
        }
    b.ok = true;
}

// P = G G^T (or D^-1 when !ok), fixed rows/columns zeroed.
__device__ __forceinline__ void lp_P(const LpBlock &b, unsigned fm, double P[9])
{
#pragma unroll
    for (int r = 0; r < 3; ++r)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double s;
            if (b.ok) {
                s = 0.0;
#pragma unroll
                for (int k = 0; k < 3; ++k) s += b.G[r][k] * b.G[c][k];
            } else {
                s = b.Di[3 * r + c];
            }
            if (((fm >> r) & 1) || ((fm >> c) & 1)) s = 0.0;
            P[3 * r + c] = s;
        }
}

// ---------------------------------------------------------------------------
// Features.  Pass 1: partial sums of log(tr D_i / 3) (natural vertex order).
__global__ void __launch_bounds__(RT)
k_lp_logt(int N, const int *__restrict__ diagpos, const double *__restrict__ A, int64_t nbs,
          const unsigned char *__restrict__ fixmask, double *__restrict__ part)
{
    __shared__ double s[RT];
    const int i = blockIdx.x * RT + threadIdx.x;
    double lt = 0.0;
    if (i < N) {
        double D[9];
        masked_block(A, nbs, diagpos[i], fixmask[i], fixmask[i], true, D);
        lt = log((D[0] + D[4] + D[8]) / 3.0);
    }
    const double t = block_sum(lt, s);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// One workgroup: out = fixed-order sum of part[0..n).
__global__ void __launch_bounds__(RT) k_lp_reduce(int n, const double *__restrict__ part, double *__restrict__ out)
{
    __shared__ double s[RT];
    double v = 0.0;
    for (int k = threadIdx.x; k < n; k += RT) v += part[k];
    const double t = block_sum(v, s);
    if (threadIdx.x == 0) *out = t;
}

// Pass 2 (SELL storage order g): the 16 features of vertex srow[g] and the
// normalised weights of its off-diagonal entries (same storage positions as A).
__global__ void __launch_bounds__(RT)
k_lp_features(int N, const int *__restrict__ srow, const int *__restrict__ slice_off, const int *__restrict__ slen,
              const int *__restrict__ scol, const int *__restrict__ diagpos, const double *__restrict__ A,
              int64_t nbs, const double *__restrict__ Mval, const unsigned char *__restrict__ fixmask,
              const double *__restrict__ logt_sum, unsigned short *__restrict__ X, float *__restrict__ w)
{
    const int g = blockIdx.x * RT + threadIdx.x;
    if (g >= N) return;
    const int i = srow[g];
    const int64_t base = slice_off[g / SL] + (g % SL);
    const int len = slen[g];
    const unsigned fi = fixmask[i];
    double D[9];
    masked_block(A, nbs, diagpos[i], fi, fi, true, D);
    const double t = (D[0] + D[4] + D[8]) / 3.0;
    double ssum = 0.0;
    int deg = 0;
    for (int k = 0; k < len; ++k) {
        const int64_t e = base + (int64_t)k * SL;
        const int j = scol[e];
        if (j == i) continue;
        double b[9];
        masked_block(A, nbs, e, fi, fixmask[j], false, b);
        ssum += fro(b);
        ++deg;
    }
    for (int k = 0; k < len; ++k) {
        const int64_t e = base + (int64_t)k * SL;
        const int j = scol[e];
        float wv = 0.0f;
        if (j != i && ssum > 0.0) {
            double b[9];
            masked_block(A, nbs, e, fi, fixmask[j], false, b);
            wv = (float)(fro(b) / ssum);
        }
        w[e] = wv;
    }
    double F[NF];
    F[0] = D[0] / t; F[1] = D[1] / t; F[2] = D[2] / t; F[3] = D[4] / t; F[4] = D[5] / t; F[5] = D[8] / t;
    F[6] = log(t) - *logt_sum / N;
    F[7] = Mval[diagpos[i]] / t;
    F[8] = ssum / fro(D);
    F[9] = deg / 16.0;
    F[10] = (double)(fi & 1); F[11] = (double)((fi >> 1) & 1); F[12] = (double)((fi >> 2) & 1);
    F[13] = F[14] = F[15] = 0.0;
    unsigned short *dst = X + (size_t)NF * i;
#pragma unroll
    for (int c = 0; c < NF; ++c) dst[c] = f2bf((float)F[c]);
}

// wT[e(i, j)] = w[e(j, i)]
__global__ void __launch_bounds__(RT)
k_lp_transpose_w(int64_t nbs, const int *__restrict__ tpos, const float *__restrict__ w, float *__restrict__ wT)
{
    const int64_t e = (int64_t)blockIdx.x * RT + threadIdx.x;
    if (e >= nbs) return;
    const int t = tpos[e];
    wT[e] = t >= 0 ? w[t] : 0.0f;
}

// ---------------------------------------------------------------------------
// Dense layer on the matrix units.  Y[N][NOUT] = [A1 | A2] W (+ bias), where the
// input rows are A1[N][K1] | A2[N][K2] (bf16) and W (K x NOUT) is given transposed,
// Wt[NOUT][K] (bf16).  RELU_BF16: ReLU, bf16 output; otherwise FP32 output.
// Each wave keeps its B fragments (the whole of W) in registers and walks 16-row
// tiles with a grid stride.
template <int K1, int K2, int NOUT, bool RELU_BF16>
__global__ void __launch_bounds__(256)
k_lp_dense(int N, const unsigned short *__restrict__ A1, const unsigned short *__restrict__ A2,
           const unsigned short *__restrict__ Wt, const float *__restrict__ bias, void *__restrict__ Y)
{
    constexpr int K = K1 + K2, KB = K / 16, NB = NOUT / 16;
    static_assert(K1 % 16 == 0 && K2 % 16 == 0 && NOUT % 16 == 0, "tile sizes");
    const int lane = threadIdx.x & 63, q = lane >> 4, c16 = lane & 15;
    short4v bf[KB][NB];
#pragma unroll
    for (int kb = 0; kb < KB; ++kb)
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) {
            const uint2 v = *reinterpret_cast<const uint2 *>(Wt + (size_t)(nb * 16 + c16) * K + kb * 16 + 4 * q);
            bf[kb][nb] = short4v{(short)(v.x & 0xFFFFu), (short)(v.x >> 16), (short)(v.y & 0xFFFFu), (short)(v.y >> 16)};
        }
    float bv[NB];
#pragma unroll
    for (int nb = 0; nb < NB; ++nb) bv[nb] = bias ? bias[nb * 16 + c16] : 0.0f;
    const int nwaves = gridDim.x * 4;
    for (int tile = blockIdx.x * 4 + (threadIdx.x >> 6); tile * 16 < N; tile += nwaves) {
        const int row0 = tile * 16, ar = row0 + c16;
        float4v acc[NB];
#pragma unroll
        for (int nb = 0; nb < NB; ++nb) acc[nb] = float4v{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int kb = 0; kb < KB; ++kb) {
            uint2 v = make_uint2(0u, 0u);
            if (ar < N) {
                const unsigned short *src = (kb * 16 < K1) ? A1 + (size_t)K1 * ar + kb * 16 + 4 * q
                                                           : A2 + (size_t)K2 * ar + (kb * 16 - K1) + 4 * q;
                v = *reinterpret_cast<const uint2 *>(src);
            }
            const short4v a = short4v{(short)(v.x & 0xFFFFu), (short)(v.x >> 16), (short)(v.y & 0xFFFFu), (short)(v.y >> 16)};
#pragma unroll
            for (int nb = 0; nb < NB; ++nb) acc[nb] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bf[kb][nb], acc[nb], 0, 0, 0);
        }
#pragma unroll
        for (int nb = 0; nb < NB; ++nb)
#pragma unroll
            for (int t = 0; t < 4; ++t) {
                const int r = row0 + 4 * q + t, col = nb * 16 + c16;
                if (r >= N) continue;
                const float y = acc[nb][t] + bv[nb];
                if constexpr (RELU_BF16) static_cast<unsigned short *>(Y)[(size_t)NOUT * r + col] = f2bf(y > 0.f ? y : 0.f);
                else static_cast<float *>(Y)[(size_t)NOUT * r + col] = y;
            }
    }
}

// Aggregation a[i] = sum_e w[e] h[col_e] over row i's entries (bf16 in and out,
// FP32 sums in entry order).  8 threads per row, 8 channels each.
__global__ void __launch_bounds__(RT)
k_lp_aggregate(int N, const int *__restrict__ srow, const int *__restrict__ slice_off, const int *__restrict__ slen,
               const int *__restrict__ scol, const float *__restrict__ w, const unsigned short *__restrict__ h,
               unsigned short *__restrict__ a)
{
    const int t = blockIdx.x * RT + threadIdx.x;
    const int g = t >> 3, cg = (t & 7) * 8;
    if (g >= N) return;
    const int i = srow[g];
    const int64_t base = slice_off[g / SL] + (g % SL);
    const int len = slen[g];
    float acc[8];
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[c] = 0.0f;
    for (int k = 0; k < len; ++k) {
        const int64_t e = base + (int64_t)k * SL;
        const float wv = w[e];
        if (wv == 0.0f) continue;
        const uint4 v = *reinterpret_cast<const uint4 *>(h + (size_t)HID * scol[e] + cg);
        const unsigned vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int c = 0; c < 4; ++c) {
            acc[2 * c] += wv * bf2f((unsigned short)(vv[c] & 0xFFFFu));
            acc[2 * c + 1] += wv * bf2f((unsigned short)(vv[c] >> 16));
        }
    }
    unsigned pk[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) pk[c] = (unsigned)f2bf(acc[2 * c]) | ((unsigned)f2bf(acc[2 * c + 1]) << 16);
    *reinterpret_cast<uint4 *>(a + (size_t)HID * i + cg) = make_uint4(pk[0], pk[1], pk[2], pk[3]);
}

// Output layer o = h3 W_out + b_out (FP32 sums of bf16 products).
__global__ void __launch_bounds__(RT)
k_lp_out(int N, const unsigned short *__restrict__ h, const unsigned short *__restrict__ Wq /* [64][6] */,
         const float *__restrict__ b, float *__restrict__ o)
{
    __shared__ float sW[HID * NO];
    for (int k = threadIdx.x; k < HID * NO; k += RT) sW[k] = bf2f(Wq[k]);
    __syncthreads();
    const int i = blockIdx.x * RT + threadIdx.x;
    if (i >= N) return;
    float acc[NO];
#pragma unroll
    for (int k = 0; k < NO; ++k) acc[k] = 0.f;
    const uint4 *src = reinterpret_cast<const uint4 *>(h + (size_t)HID * i);
    for (int c8 = 0; c8 < HID / 8; ++c8) {
        const uint4 v = src[c8];
        const unsigned vv[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const float hv = bf2f((unsigned short)((u & 1) ? (vv[u >> 1] >> 16) : (vv[u >> 1] & 0xFFFFu)));
            const int c = 8 * c8 + u;
#pragma unroll
            for (int k = 0; k < NO; ++k) acc[k] += hv * sW[c * NO + k];
        }
    }
#pragma unroll
    for (int k = 0; k < NO; ++k) o[(size_t)NO * i + k] = acc[k] + b[k];
}

// P blocks (natural vertex order).  Inference: P -> D32 (6 upper entries, SoA).
// Training (r != null): z = P r and partial sums of z.r.
__global__ void __launch_bounds__(RT)
k_lp_precond(int N, const int *__restrict__ diagpos, const double *__restrict__ A, int64_t nbs,
             const unsigned char *__restrict__ fixmask, const float *__restrict__ o, float *__restrict__ D32,
             const double *__restrict__ r, double *__restrict__ z, double *__restrict__ part)
{
    __shared__ double s[RT];
    const int i = blockIdx.x * RT + threadIdx.x;
    double zr = 0.0;
    if (i < N) {
        const unsigned fm = fixmask[i];
        double D[9];
        masked_block(A, nbs, diagpos[i], fm, fm, true, D);
        LpBlock b;
        lp_block(D, o + (size_t)NO * i, b);
        double P[9];
        lp_P(b, fm, P);
        if (D32) {
            const int up[eolc::kDsym] = {0, 1, 2, 4, 5, 8};   // P is symmetric (G G^T)
#pragma unroll
            for (int k = 0; k < eolc::kDsym; ++k) D32[(size_t)k * N + i] = (float)P[up[k]];
        }
        if (r) {
            const double r0 = r[3 * i], r1 = r[3 * i + 1], r2 = r[3 * i + 2];
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                const double zz = (P[3 * a] * r0 + P[3 * a + 1] * r1) + P[3 * a + 2] * r2;
                z[3 * i + a] = zz;
                zr += zz * r[3 * i + a];
            }
        }
    }
    if (r) {
        const double t = block_sum(zr, s);
        if (threadIdx.x == 0) part[blockIdx.x] = t;
    }
}

// z += zc (the multilevel coarse corrections of r) and partial sums of the new z.r:
// the training loss of the blocks used inside the multilevel preconditioner.
__global__ void __launch_bounds__(RT)
k_lp_add_coarse(int N, const double *__restrict__ zc, const double *__restrict__ r, double *__restrict__ z,
                double *__restrict__ part)
{
    __shared__ double s[RT];
    const int i = blockIdx.x * RT + threadIdx.x;
    double zr = 0.0;
    if (i < N) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            const double zz = z[3 * i + a] + zc[3 * i + a];
            z[3 * i + a] = zz;
            zr += zz * r[3 * i + a];
        }
    }
    const double t = block_sum(zr, s);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// y = A_red x (FP64; fixed rows give 0) in SELL order, partial sums of y.u.
__global__ void __launch_bounds__(RT)
k_lp_spmv_dot(int N, const int *__restrict__ srow, const int *__restrict__ slice_off, const int *__restrict__ slen,
              const int *__restrict__ scol, const double *__restrict__ A, int64_t nbs,
              const unsigned char *__restrict__ fixmask, const double *__restrict__ x, double *__restrict__ y,
              const double *__restrict__ u, double *__restrict__ part)
{
    __shared__ double s[RT];
    const int g = blockIdx.x * RT + threadIdx.x;
    double acc = 0.0;
    if (g < N) {
        const int i = srow[g];
        const int64_t base = slice_off[g / SL] + (g % SL);
        const int len = slen[g];
        double y0 = 0.0, y1 = 0.0, y2 = 0.0;
        for (int k = 0; k < len; ++k) {
            const int64_t e = base + (int64_t)k * SL;
            const int j = scol[e];
            const double u0 = x[3 * j], u1 = x[3 * j + 1], u2 = x[3 * j + 2];
            y0 += A[0 * nbs + e] * u0 + A[1 * nbs + e] * u1 + A[2 * nbs + e] * u2;
            y1 += A[3 * nbs + e] * u0 + A[4 * nbs + e] * u1 + A[5 * nbs + e] * u2;
            y2 += A[6 * nbs + e] * u0 + A[7 * nbs + e] * u1 + A[8 * nbs + e] * u2;
        }
        const unsigned fm = fixmask[i];
        if (fm & 1) y0 = 0.0;
        if (fm & 2) y1 = 0.0;
        if (fm & 4) y2 = 0.0;
        y[3 * i] = y0; y[3 * i + 1] = y1; y[3 * i + 2] = y2;
        acc = (y0 * u[3 * i] + y1 * u[3 * i + 1]) + y2 * u[3 * i + 2];
    }
    const double t = block_sum(acc, s);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

__global__ void k_lp_loss(double *scal)
{
    // scal: [1] r.e, [2] z.r, [3] z.Az -> [4] loss
    if (blockIdx.x == 0 && threadIdx.x == 0) scal[4] = 1.0 - scal[2] * scal[2] / (scal[3] * scal[1]);
}

__global__ void __launch_bounds__(RT)
k_lp_mask(int N, const double *__restrict__ in, const unsigned char *__restrict__ fixmask, double *__restrict__ out)
{
    const int t = blockIdx.x * RT + threadIdx.x;
    if (t >= 3 * N) return;
    const int i = t / 3, c = t - 3 * i;
    out[t] = ((fixmask[i] >> c) & 1) ? 0.0 : in[t];
}

// ---------------------------------------------------------------------------
// Backward.  dO_i from dloss/dz_i = c1 (r_i - c2 Az_i), c1 = -2 zr / (zAz re),
// c2 = zr / zAz:  S = dz r^T (fixed rows/cols 0), dG = (S + S^T) G, dT = L^T dG,
// dO_k = dT[r_k][c_k] * 0.5 (1 - th_k^2).
__global__ void __launch_bounds__(RT)
k_lp_dO(int N, const int *__restrict__ diagpos, const double *__restrict__ A, int64_t nbs,
        const unsigned char *__restrict__ fixmask, const float *__restrict__ o, const double *__restrict__ r,
        const double *__restrict__ Az, const double *__restrict__ scal, float *__restrict__ dO)
{
    const int i = blockIdx.x * RT + threadIdx.x;
    if (i >= N) return;
    const double re = scal[1], zr = scal[2], zAz = scal[3];
    const double c1 = -2.0 * zr / (zAz * re), c2 = zr / zAz;
    const unsigned fm = fixmask[i];
    double D[9];
    masked_block(A, nbs, diagpos[i], fm, fm, true, D);
    LpBlock b;
    lp_block(D, o + (size_t)NO * i, b);
    float *out = dO + (size_t)NO * i;
    if (!b.ok) {
#pragma unroll
        for (int k = 0; k < NO; ++k) out[k] = 0.0f;
        return;
    }
    double rv[3], dz[3];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        rv[a] = r[3 * i + a];
        dz[a] = c1 * (rv[a] - c2 * Az[3 * i + a]);
    }
    double S2[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double s = dz[a] * rv[c];
            if (((fm >> a) & 1) || ((fm >> c) & 1)) s = 0.0;
            S2[a][c] = s;
        }
    double Ssym[3][3], dG[3][3];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) Ssym[a][c] = S2[a][c] + S2[c][a];
#pragma unroll
    for (int a = 0; a < 3; ++a)
#pragma unroll
        for (int c = 0; c < 3; ++c) {
            double s = 0.0;
#pragma unroll
            for (int m = 0; m < 3; ++m) s += Ssym[a][m] * b.G[m][c];
            dG[a][c] = s;
        }
    const int TR[6] = {0, 1, 1, 2, 2, 2}, TC[6] = {0, 0, 1, 0, 1, 2};
#pragma unroll
    for (int k = 0; k < NO; ++k) {
        double s = 0.0;
#pragma unroll
        for (int m = 0; m < 3; ++m) s += b.L[m][TR[k]] * dG[m][TC[k]];
        out[k] = (float)(s * 0.5 * (1.0 - b.th[k] * b.th[k]));
    }
}

// Output layer backward: dpre = bf16((dO W_out^T) * (h3 > 0)); per-workgroup
// partials of dW_out = h3^T dO (64 x 6) and db_out (6), stride 390.
__global__ void __launch_bounds__(RT)
k_lp_out_bwd(int N, const unsigned short *__restrict__ h, const unsigned short *__restrict__ Wq,
             const float *__restrict__ dO, unsigned short *__restrict__ dpre, float *__restrict__ part)
{
    constexpr int NG = HID * NO + NO;   // 390
    constexpr int HP = HID + 2;         // padded LDS row
    __shared__ float sW[HID * NO];
    __shared__ float sd[RT * NO];
    __shared__ unsigned short sh[RT * HP];
    for (int k = threadIdx.x; k < HID * NO; k += RT) sW[k] = bf2f(Wq[k]);
    const int r0 = blockIdx.x * RT;
    for (int k = threadIdx.x; k < RT * NO; k += RT) {
        const int r = r0 + k / NO;
        sd[k] = r < N ? dO[(size_t)NO * r + (k % NO)] : 0.0f;
    }
    for (int k = threadIdx.x; k < RT * HID; k += RT) {
        const int rr = k / HID, r = r0 + rr;
        sh[rr * HP + (k % HID)] = r < N ? h[(size_t)HID * r + (k % HID)] : (unsigned short)0;
    }
    __syncthreads();
    const int i = r0 + threadIdx.x;
    if (i < N) {
        const float *d = sd + NO * threadIdx.x;
        const unsigned short *hr = sh + threadIdx.x * HP;
        unsigned short *dst = dpre + (size_t)HID * i;
        for (int c8 = 0; c8 < HID / 8; ++c8) {
            unsigned pk[4];
#pragma unroll
            for (int u = 0; u < 8; u += 2) {
                float v[2];
#pragma unroll
                for (int w2 = 0; w2 < 2; ++w2) {
                    const int c = 8 * c8 + u + w2;
                    float acc = 0.f;
#pragma unroll
                    for (int k = 0; k < NO; ++k) acc += d[k] * sW[c * NO + k];
                    v[w2] = bf2f(hr[c]) > 0.f ? acc : 0.f;
                }
                pk[u >> 1] = (unsigned)f2bf(v[0]) | ((unsigned)f2bf(v[1]) << 16);
            }
            *reinterpret_cast<uint4 *>(dst + 8 * c8) = make_uint4(pk[0], pk[1], pk[2], pk[3]);
        }
    }
    for (int g = threadIdx.x; g < NG; g += RT) {
        float acc = 0.f;
        if (g < HID * NO) {
            const int c = g / NO, k = g % NO;
            for (int rr = 0; rr < RT; ++rr) acc += bf2f(sh[rr * HP + c]) * sd[rr * NO + k];
        } else {
            const int k = g - HID * NO;
            for (int rr = 0; rr < RT; ++rr) acc += sd[rr * NO + k];
        }
        part[(size_t)blockIdx.x * NG + g] = acc;
    }
}

// Weight gradient on the matrix units: per workgroup, over its row range,
//   part[K][NOUT] = [A1 | A2]^T dY   and   part[K * NOUT + n] = sum of dY[:, n],
// bf16 inputs staged transposed through LDS (64-row tiles), FP32 sums.
template <int K1, int K2, int NOUT>
__global__ void __launch_bounds__(256)
k_lp_wgrad(int N, int rows_per_wg, const unsigned short *__restrict__ A1, const unsigned short *__restrict__ A2,
           const unsigned short *__restrict__ dY, float *__restrict__ part)
{
    constexpr int K = K1 + K2;
    constexpr int TR = 64, TP = TR + 4;   // rows per tile, padded LDS row
    __shared__ unsigned short sA[K * TP];
    __shared__ unsigned short sD[NOUT * TP];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, c16 = lane & 15;
    constexpr int NBLK = (K / 16) * (NOUT / 16);
    constexpr int PER = (NBLK + 3) / 4;
    float4v acc[PER];
#pragma unroll
    for (int p = 0; p < PER; ++p) acc[p] = float4v{0.f, 0.f, 0.f, 0.f};
    float bacc = 0.0f;
    const int r_begin = blockIdx.x * rows_per_wg;
    const int r_end = min(N, r_begin + rows_per_wg);
    for (int r0 = r_begin; r0 < r_end; r0 += TR) {
        __syncthreads();
        for (int idx = threadIdx.x; idx < TR * K; idx += 256) {
            const int rr = idx / K, kk = idx % K, r = r0 + rr;
            unsigned short v = 0;
            if (r < r_end) v = (kk < K1) ? A1[(size_t)K1 * r + kk] : A2[(size_t)K2 * r + (kk - K1)];
            sA[kk * TP + rr] = v;
        }
        for (int idx = threadIdx.x; idx < TR * NOUT; idx += 256) {
            const int rr = idx / NOUT, n = idx % NOUT, r = r0 + rr;
            sD[n * TP + rr] = r < r_end ? dY[(size_t)NOUT * r + n] : (unsigned short)0;
        }
        __syncthreads();
        if (threadIdx.x < NOUT) {
            for (int rr = 0; rr < TR; ++rr) bacc += bf2f(sD[threadIdx.x * TP + rr]);
        }
#pragma unroll
        for (int p = 0; p < PER; ++p) {
            const int blk = wave * PER + p;
            if (blk < NBLK) {
                const int mb = blk / (NOUT / 16), nb = blk % (NOUT / 16);
#pragma unroll
                for (int rb = 0; rb < TR; rb += 16) {
                    const unsigned short *pa = sA + (mb * 16 + c16) * TP + rb + 4 * q;
                    const unsigned short *pb = sD + (nb * 16 + c16) * TP + rb + 4 * q;
                    const short4v a = short4v{(short)pa[0], (short)pa[1], (short)pa[2], (short)pa[3]};
                    const short4v bb = short4v{(short)pb[0], (short)pb[1], (short)pb[2], (short)pb[3]};
                    acc[p] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bb, acc[p], 0, 0, 0);
                }
            }
        }
    }
    float *out = part + (size_t)blockIdx.x * (K * NOUT + NOUT);
#pragma unroll
    for (int p = 0; p < PER; ++p) {
        const int blk = wave * PER + p;
        if (blk < NBLK) {
            const int mb = blk / (NOUT / 16), nb = blk % (NOUT / 16);
#pragma unroll
            for (int t = 0; t < 4; ++t) out[(size_t)(mb * 16 + 4 * q + t) * NOUT + nb * 16 + c16] = acc[p][t];
        }
    }
    if (threadIdx.x < NOUT) out[K * NOUT + threadIdx.x] = bacc;
}

// Transposed aggregation for the backward of a message-passing layer:
//   dh[i] = dinp[i][0:64] + sum_e wT[e] dinp[col_e][64:128],
//   dpre[i] = bf16(dh[i] * (h[i] > 0)).
__global__ void __launch_bounds__(RT)
k_lp_agg_bwd(int N, const int *__restrict__ srow, const int *__restrict__ slice_off, const int *__restrict__ slen,
             const int *__restrict__ scol, const float *__restrict__ wT, const float *__restrict__ dinp,
             const unsigned short *__restrict__ h, unsigned short *__restrict__ dpre)
{
    const int t = blockIdx.x * RT + threadIdx.x;
    const int g = t >> 3, cg = (t & 7) * 8;
    if (g >= N) return;
    const int i = srow[g];
    const int64_t base = slice_off[g / SL] + (g % SL);
    const int len = slen[g];
    float acc[8];
    const float *own = dinp + (size_t)2 * HID * i + cg;
#pragma unroll
    for (int c = 0; c < 8; ++c) acc[c] = own[c];
    for (int k = 0; k < len; ++k) {
        const int64_t e = base + (int64_t)k * SL;
        const float wv = wT[e];
        if (wv == 0.0f) continue;
        const float *src = dinp + (size_t)2 * HID * scol[e] + HID + cg;
#pragma unroll
        for (int c = 0; c < 8; ++c) acc[c] += wv * src[c];
    }
    const uint4 hv = *reinterpret_cast<const uint4 *>(h + (size_t)HID * i + cg);
    const unsigned hh[4] = {hv.x, hv.y, hv.z, hv.w};
    unsigned pk[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) {
        const float h0 = bf2f((unsigned short)(hh[c] & 0xFFFFu)), h1 = bf2f((unsigned short)(hh[c] >> 16));
        const float v0 = h0 > 0.f ? acc[2 * c] : 0.f, v1 = h1 > 0.f ? acc[2 * c + 1] : 0.f;
        pk[c] = (unsigned)f2bf(v0) | ((unsigned)f2bf(v1) << 16);
    }
    *reinterpret_cast<uint4 *>(dpre + (size_t)HID * i + cg) = make_uint4(pk[0], pk[1], pk[2], pk[3]);
}

// out[k] = sum over G partial rows (stride L) of part[g * L + k], in row order.
__global__ void __launch_bounds__(RT)
k_lp_sum_partials(int G, int L, const float *__restrict__ part, float *__restrict__ out)
{
    const int k = blockIdx.x * RT + threadIdx.x;
    if (k >= L) return;
    float s = 0.0f;
    for (int g = 0; g < G; ++g) s += part[(size_t)g * L + k];
    out[k] = s;
}

// Adam (Kingma & Ba 2015) with bias correction.
__global__ void __launch_bounds__(RT)
k_lp_adam(int n, float *__restrict__ theta, const float *__restrict__ g, float *__restrict__ m, float *__restrict__ v,
          float lr, float b1, float b2, float eps, float bc1, float bc2)
{
    const int k = blockIdx.x * RT + threadIdx.x;
    if (k >= n) return;
    const float gk = g[k];
    const float mk = b1 * m[k] + (1.0f - b1) * gk;
    const float vk = b2 * v[k] + (1.0f - b2) * gk * gk;
    m[k] = mk;
    v[k] = vk;
    theta[k] -= lr * (mk / bc1) / (sqrtf(vk / bc2) + eps);
}

// bf16 copies of the weight matrices in the layouts the kernels read.
__global__ void __launch_bounds__(RT) k_lp_quantize(const float *__restrict__ th, unsigned short *__restrict__ wq)
{
    const int t = blockIdx.x * RT + threadIdx.x;
    if (t >= eolc::kLqSize) return;
    float v;
    if (t < eolc::kLqLayerT) {                       // W_in^T[n][k] = W_in[k][n]
        const int n = t / NF, k = t % NF;
        v = th[eolc::kLpOffIn + k * HID + n];
    } else if (t < eolc::kLqLayer) {                 // W_l^T[n][k] = W_l[k][n]
        const int u = t - eolc::kLqLayerT, l = u / (2 * HID * HID), rem = u % (2 * HID * HID);
        const int n = rem / (2 * HID), k = rem % (2 * HID);
        v = th[eolc::kLpOffLayer0 + l * eolc::kLpLayerSize + k * HID + n];
    } else if (t < eolc::kLqOut) {                   // W_l as stored, [128][64]
        const int u = t - eolc::kLqLayer, l = u / (2 * HID * HID), rem = u % (2 * HID * HID);
        v = th[eolc::kLpOffLayer0 + l * eolc::kLpLayerSize + rem];
    } else {                                         // W_out [64][6]
        v = th[eolc::kLpOffOut + (t - eolc::kLqOut)];
    }
    wq[t] = f2bf(v);
}

inline int dense_grid(int N) { return std::max(1, std::min(2048, (N + 63) / 64)); }

}  // namespace

namespace eolc {

void learned_wgrad_shape(int N, int *G, int *rows)
{
    int g = std::min(kLpGradWG, std::max(1, (N + 63) / 64));
    int per = (N + g - 1) / g;
    per = (per + 63) & ~63;
    *rows = per;
    *G = (N + per - 1) / per;
}

hipError_t learned_quantize(eolc_ctx *c)
{
    hipLaunchKernelGGL(k_lp_quantize, dim3(blocks_for(kLqSize)), dim3(RT), 0, c->stream, c->lp.theta, c->lp.wq);
    return hipGetLastError();
}

hipError_t learned_mask_error(eolc_ctx *c)
{
    const int N = c->topo.N;
    hipLaunchKernelGGL(k_lp_mask, dim3(blocks_for(3 * N)), dim3(RT), 0, c->stream, N, c->lp.e, c->fixmask, c->lp.e);
    return hipGetLastError();
}

hipError_t learned_forward(eolc_ctx *c)
{
    LearnedState &L = c->lp;
    const Topology &t = c->topo;
    const int N = t.N;
    const int nblk = blocks_for(N);
    hipLaunchKernelGGL(k_lp_logt, dim3(nblk), dim3(RT), 0, c->stream, N, c->diagpos, c->A64, t.nbs, c->fixmask, L.dpart);
    hipLaunchKernelGGL(k_lp_reduce, dim3(1), dim3(RT), 0, c->stream, nblk, L.dpart, L.scal + 0);
    hipLaunchKernelGGL(k_lp_features, dim3(nblk), dim3(RT), 0, c->stream, N, c->srow, c->slice_off, c->slen, c->scol,
                       c->diagpos, c->A64, t.nbs, c->Mval, c->fixmask, L.scal + 0, L.X, L.w);
    hipLaunchKernelGGL(k_lp_transpose_w, dim3(blocks_for(t.nbs)), dim3(RT), 0, c->stream, t.nbs, L.tpos, L.w, L.wT);
    const int dg = dense_grid(N);
    const float *th = L.theta;
    hipLaunchKernelGGL((k_lp_dense<NF, 0, HID, true>), dim3(dg), dim3(256), 0, c->stream, N, L.X,
                       (const unsigned short *)nullptr, L.wq + kLqInT, th + kLpOffIn + NF * HID, (void *)L.h[0]);
    for (int l = 0; l < kLpLayers; ++l) {
        hipLaunchKernelGGL(k_lp_aggregate, dim3(blocks_for((int64_t)N * 8)), dim3(RT), 0, c->stream, N, c->srow,
                           c->slice_off, c->slen, c->scol, L.w, L.h[l], L.a[l]);
        hipLaunchKernelGGL((k_lp_dense<HID, HID, HID, true>), dim3(dg), dim3(256), 0, c->stream, N, L.h[l], L.a[l],
                           L.wq + kLqLayerT + l * 2 * HID * HID,
                           th + kLpOffLayer0 + l * kLpLayerSize + 2 * HID * HID, (void *)L.h[l + 1]);
    }
    hipLaunchKernelGGL(k_lp_out, dim3(nblk), dim3(RT), 0, c->stream, N, L.h[kLpLayers], L.wq + kLqOut,
                       th + kLpOffOut + HID * NO, L.o);
    return hipGetLastError();
}

hipError_t learned_precond(eolc_ctx *c)
{
    hipError_t e = learned_forward(c);
    if (e != hipSuccess) return e;
    const int N = c->topo.N;
    hipLaunchKernelGGL(k_lp_precond, dim3(blocks_for(N)), dim3(RT), 0, c->stream, N, c->diagpos, c->A64, c->topo.nbs,
                       c->fixmask, c->lp.o, c->D32, (const double *)nullptr, (double *)nullptr, (double *)nullptr);
    return hipGetLastError();
}

hipError_t learned_loss_grad(eolc_ctx *c)
{
    LearnedState &L = c->lp;
    const Topology &t = c->topo;
    const int N = t.N;
    const int nblk = blocks_for(N);
    hipError_t err = learned_forward(c);
    if (err != hipSuccess) return err;
    // r = A e (r.e), z = P r (z.r), Az (z.Az), loss
    hipLaunchKernelGGL(k_lp_spmv_dot, dim3(nblk), dim3(RT), 0, c->stream, N, c->srow, c->slice_off, c->slen, c->scol,
                       c->A64, t.nbs, c->fixmask, L.e, L.r, L.e, L.dpart);
    hipLaunchKernelGGL(k_lp_reduce, dim3(1), dim3(RT), 0, c->stream, nblk, L.dpart, L.scal + 1);
    hipLaunchKernelGGL(k_lp_precond, dim3(nblk), dim3(RT), 0, c->stream, N, c->diagpos, c->A64, t.nbs, c->fixmask, L.o,
                       (float *)nullptr, L.r, L.z, L.dpart);
    if (c->ml_on) {
        // z = P r + coarse(r): the coarse part does not depend on the parameters, so
        // the backward pass below is unchanged (it reads r, Az and the scalars)
        err = hipMemcpyAsync(c->r64, L.r, sizeof(double) * 3 * (size_t)N, hipMemcpyDeviceToDevice, c->stream);
        if (err != hipSuccess) return err;
        err = ml_coarse64(c);
        if (err != hipSuccess) return err;
        hipLaunchKernelGGL(k_lp_add_coarse, dim3(nblk), dim3(RT), 0, c->stream, N, c->z64, L.r, L.z, L.dpart);
    }
    hipLaunchKernelGGL(k_lp_reduce, dim3(1), dim3(RT), 0, c->stream, nblk, L.dpart, L.scal + 2);
    hipLaunchKernelGGL(k_lp_spmv_dot, dim3(nblk), dim3(RT), 0, c->stream, N, c->srow, c->slice_off, c->slen, c->scol,
                       c->A64, t.nbs, c->fixmask, L.z, L.Az, L.z, L.dpart);
    hipLaunchKernelGGL(k_lp_reduce, dim3(1), dim3(RT), 0, c->stream, nblk, L.dpart, L.scal + 3);
    hipLaunchKernelGGL(k_lp_loss, dim3(1), dim3(1), 0, c->stream, L.scal);
    // backward
    hipLaunchKernelGGL(k_lp_dO, dim3(nblk), dim3(RT), 0, c->stream, N, c->diagpos, c->A64, t.nbs, c->fixmask, L.o, L.r,
                       L.Az, L.scal, L.dO);
    constexpr int NGO = HID * NO + NO;
    hipLaunchKernelGGL(k_lp_out_bwd, dim3(nblk), dim3(RT), 0, c->stream, N, L.h[kLpLayers], L.wq + kLqOut, L.dO,
                       L.dpre[0], L.gpart);
    hipLaunchKernelGGL(k_lp_sum_partials, dim3(blocks_for(NGO)), dim3(RT), 0, c->stream, nblk, NGO, L.gpart,
                       L.grad + kLpOffOut);
    int Gw, rows;
    learned_wgrad_shape(N, &Gw, &rows);
    const int dg = dense_grid(N);
    int cur = 0;
    constexpr int NGL = 2 * HID * HID + HID;
    for (int l = kLpLayers - 1; l >= 0; --l) {
        hipLaunchKernelGGL((k_lp_wgrad<HID, HID, HID>), dim3(Gw), dim3(256), 0, c->stream, N, rows, L.h[l], L.a[l],
                           L.dpre[cur], L.gpart);
        hipLaunchKernelGGL(k_lp_sum_partials, dim3(blocks_for(NGL)), dim3(RT), 0, c->stream, Gw, NGL, L.gpart,
                           L.grad + kLpOffLayer0 + l * kLpLayerSize);
        // dinp = dpre W_l^T: W_l [128][64] is the transposed operand of this product
        hipLaunchKernelGGL((k_lp_dense<HID, 0, 2 * HID, false>), dim3(dg), dim3(256), 0, c->stream, N, L.dpre[cur],
                           (const unsigned short *)nullptr, L.wq + kLqLayer + l * 2 * HID * HID,
                           (const float *)nullptr, (void *)L.dinp);
        hipLaunchKernelGGL(k_lp_agg_bwd, dim3(blocks_for((int64_t)N * 8)), dim3(RT), 0, c->stream, N, c->srow,
                           c->slice_off, c->slen, c->scol, L.wT, L.dinp, L.h[l], L.dpre[cur ^ 1]);
        cur ^= 1;
    }
    constexpr int NGI = NF * HID + HID;
    hipLaunchKernelGGL((k_lp_wgrad<NF, 0, HID>), dim3(Gw), dim3(256), 0, c->stream, N, rows, L.X,
                       (const unsigned short *)nullptr, L.dpre[cur], L.gpart);
    hipLaunchKernelGGL(k_lp_sum_partials, dim3(blocks_for(NGI)), dim3(RT), 0, c->stream, Gw, NGI, L.gpart,
                       L.grad + kLpOffIn);
    return hipGetLastError();
}

hipError_t learned_adam(eolc_ctx *c, float lr)
{
    LearnedState &L = c->lp;
    L.adam_t += 1;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float bc1 = 1.0f - (float)std::pow((double)b1, (double)L.adam_t);
    const float bc2 = 1.0f - (float)std::pow((double)b2, (double)L.adam_t);
    hipLaunchKernelGGL(k_lp_adam, dim3(blocks_for(kLpParams)), dim3(RT), 0, c->stream, kLpParams, L.theta, L.grad,
                       L.adam_m, L.adam_v, lr, b1, b2, eps, bc1, bc2);
    return learned_quantize(c);
}

}  // namespace eolc
