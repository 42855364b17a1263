// Block-Jacobi PCG on the 3x3 BSR matrix MDK (gfx950).
//
// Solver "none" of GeneralizedSolver::velocitySolve (GeneralizedSolver.cpp:107-172)
// as the Dirichlet-reduced solve of MDK v = M v_old + h f: fixed DOFs are masked
// out of every matrix-vector product and carry their prescribed values.
//
// One iteration = two kernels:
//   k_pcg_spmv   : p_new = z + beta p_old (rows and, on the fly, neighbours),
//                  q = A p_new (16 lanes per block row), partial p_new.q;
//                  the last workgroup turns the partials into alpha.
//   k_pcg_update : x += alpha p_new, r -= alpha q, z = D^-1 r, partial r.z, r.r;
//                  the last workgroup forms beta and the stop decision.
// Scalars live in device memory (SolverScalars); a set `done` flag turns every
// later launch into a no-op, so the host enqueues chunks of iterations and only
// synchronises between chunks.  Reductions are deterministic: fixed in-workgroup
// trees and a fixed-order sum over the per-workgroup partials.
//
// Workgroup -> rows: launch w is mapped to logical chunk (w % 8) * (G / 8) + w / 8,
// so the 8 XCDs (which receive workgroups round-robin) each stream one
// contiguous eighth of the rows and keep their neighbour gathers in their own L2.
//
// Matrix values are stored SoA: A[c * nbp + k] is entry c (row-major 3x3) of
// block k, so the 16 lanes of a row read consecutive blocks fully coalesced.
#include "eolc_internal.h"

using eolc::SolverScalars;

namespace {

__device__ __forceinline__ int logical_wg()
{
    int G = gridDim.x;
    int w = blockIdx.x;
    return (w & 7) * (G >> 3) + (w >> 3);
}

// rows [begin, end) owned by this workgroup (G is a multiple of 8)
__device__ __forceinline__ void wg_rows(int N, int *begin, int *end)
{
    int G = gridDim.x;
    int per = (N + G - 1) / G;
    int lw = logical_wg();
    int b = lw * per;
    int e = b + per;
    *begin = b < N ? b : N;
    *end = e < N ? e : N;
}

template <class T> __device__ __forceinline__ T shfl_xor16(T v, int m) { return __shfl_xor(v, m, 16); }

__device__ __forceinline__ double wave_sum(double v)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Block sum of up to 4 per-thread values; result valid in thread 0.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double *lds /* 4*K */)
{
    int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = wave_sum(v[k]);
        if (lane == 0) lds[wid * K + k] = s;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) v[k] = ((lds[k] + lds[K + k]) + lds[2 * K + k]) + lds[3 * K + k];
    }
}

// Publishes this workgroup's partials and returns true in the last workgroup
// to arrive; that workgroup then sees every partial.
template <int K>
__device__ bool publish_partials(const double (&v)[K], double *partials, unsigned *counter, int *lds_flag)
{
    if (threadIdx.x == 0) {
        int lw = logical_wg();
#pragma unroll
        for (int k = 0; k < K; ++k) partials[k * eolc::kMaxPartials + lw] = v[k];
        __threadfence();
        unsigned prev = __hip_atomic_fetch_add(counter, 1u, __ATOMIC_ACQ_REL, __HIP_MEMORY_SCOPE_AGENT);
        *lds_flag = (prev == gridDim.x - 1);
    }
    __syncthreads();
    bool last = *lds_flag != 0;
    if (last) __threadfence();
    return last;
}

// In the last workgroup: fixed-order sum of partials[k][0..G) into out[k] (thread 0).
template <int K>
__device__ void reduce_partials(const double *partials, double (&out)[K], double *lds)
{
    int G = gridDim.x;
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
        for (int i = threadIdx.x; i < G; i += blockDim.x) s += partials[k * eolc::kMaxPartials + i];
        acc[k] = s;
    }
    block_sum<K>(acc, lds);
#pragma unroll
    for (int k = 0; k < K; ++k) out[k] = acc[k];
}

template <class T>
__device__ __forceinline__ void load3(const T *__restrict__ a, int i, T &x, T &y, T &z)
{
    x = a[3 * i]; y = a[3 * i + 1]; z = a[3 * i + 2];
}

// ---------------------------------------------------------------------------
template <class T>
__global__ void __launch_bounds__(256)
k_pcg_spmv(int N, const int *__restrict__ rowptr, const int *__restrict__ col,
           const T *__restrict__ A, int64_t nbp, const unsigned char *__restrict__ fixmask,
           const T *__restrict__ z, const T *__restrict__ pold, T *__restrict__ pnew,
           T *__restrict__ q, SolverScalars *scal, double *partials, unsigned *counter)
{
    __shared__ double lds[8];
    __shared__ int flag;
    if (scal->done) return;
    const T beta = (T)scal->beta;
    const int lane = threadIdx.x & 15;
    const int sub = threadIdx.x >> 4;
    int rb, re;
    wg_rows(N, &rb, &re);
    double dot = 0.0;
    for (int base = rb; base < re; base += eolc::kRowsPerWG) {
        int i = base + sub;
        T y0 = 0, y1 = 0, y2 = 0;
        if (i < re) {
            int kb = rowptr[i], ke = rowptr[i + 1];
            for (int k = kb + lane; k < ke; k += 16) {
                int j = col[k];
                T zj0, zj1, zj2, pj0, pj1, pj2;
                load3(z, j, zj0, zj1, zj2);
                load3(pold, j, pj0, pj1, pj2);
                T u0 = fma(beta, pj0, zj0), u1 = fma(beta, pj1, zj1), u2 = fma(beta, pj2, zj2);
                const T *a = A + k;
                y0 += a[0 * nbp] * u0 + a[1 * nbp] * u1 + a[2 * nbp] * u2;
                y1 += a[3 * nbp] * u0 + a[4 * nbp] * u1 + a[5 * nbp] * u2;
                y2 += a[6 * nbp] * u0 + a[7 * nbp] * u1 + a[8 * nbp] * u2;
            }
        }
#pragma unroll
        for (int m = 8; m > 0; m >>= 1) {
            y0 += shfl_xor16(y0, m);
            y1 += shfl_xor16(y1, m);
            y2 += shfl_xor16(y2, m);
        }
        if (i < re && lane == 0) {
            unsigned char fm = fixmask[i];
            if (fm & 1) y0 = 0;
            if (fm & 2) y1 = 0;
            if (fm & 4) y2 = 0;
            T zi0, zi1, zi2, pi0, pi1, pi2;
            load3(z, i, zi0, zi1, zi2);
            load3(pold, i, pi0, pi1, pi2);
            T u0 = fma(beta, pi0, zi0), u1 = fma(beta, pi1, zi1), u2 = fma(beta, pi2, zi2);
            pnew[3 * i] = u0; pnew[3 * i + 1] = u1; pnew[3 * i + 2] = u2;
            q[3 * i] = y0; q[3 * i + 1] = y1; q[3 * i + 2] = y2;
            dot += ((double)u0 * (double)y0 + (double)u1 * (double)y1) + (double)u2 * (double)y2;
        }
    }
    double v[1] = { dot };
    block_sum<1>(v, lds);
    if (publish_partials<1>(v, partials, counter, &flag)) {
        double tot[1];
        reduce_partials<1>(partials, tot, lds);
        if (threadIdx.x == 0) {
            double pq = tot[0];
            scal->pq = pq;
            if (!(pq > 0.0)) {
                scal->done = 1;
                scal->reason = isfinite(pq) ? 2 : 4;
            } else {
                scal->alpha = scal->rz / pq;
            }
            *counter = 0u;
        }
    }
}

// ---------------------------------------------------------------------------
// X is the solution accumulator (double in FP64 mode, float in mixed mode).
template <class T, class TX>
__global__ void __launch_bounds__(256)
k_pcg_update(int N, const T *__restrict__ D, const T *__restrict__ pnew, const T *__restrict__ q,
             TX *__restrict__ x, T *__restrict__ r, T *__restrict__ z,
             SolverScalars *scal, double *partials, unsigned *counter)
{
    __shared__ double lds[8];
    __shared__ int flag;
    if (scal->done) return;
    const T alpha = (T)scal->alpha;
    int rb, re;
    wg_rows(N, &rb, &re);
    double rz = 0.0, rr = 0.0;
    for (int i = rb + threadIdx.x; i < re; i += blockDim.x) {
        T p0, p1, p2, q0, q1, q2, r0, r1, r2;
        load3(pnew, i, p0, p1, p2);
        load3(q, i, q0, q1, q2);
        load3(r, i, r0, r1, r2);
        x[3 * i] += (TX)(alpha * p0);
        x[3 * i + 1] += (TX)(alpha * p1);
        x[3 * i + 2] += (TX)(alpha * p2);
        r0 = fma(-alpha, q0, r0);
        r1 = fma(-alpha, q1, r1);
        r2 = fma(-alpha, q2, r2);
        r[3 * i] = r0; r[3 * i + 1] = r1; r[3 * i + 2] = r2;
        const T *d = D + i;
        T z0 = d[0 * N] * r0 + d[1 * N] * r1 + d[2 * N] * r2;
        T z1 = d[3 * N] * r0 + d[4 * N] * r1 + d[5 * N] * r2;
        T z2 = d[6 * N] * r0 + d[7 * N] * r1 + d[8 * N] * r2;
        z[3 * i] = z0; z[3 * i + 1] = z1; z[3 * i + 2] = z2;
        rz += ((double)r0 * (double)z0 + (double)r1 * (double)z1) + (double)r2 * (double)z2;
        rr += ((double)r0 * (double)r0 + (double)r1 * (double)r1) + (double)r2 * (double)r2;
    }
    double v[2] = { rz, rr };
    block_sum<2>(v, lds);
    if (publish_partials<2>(v, partials, counter, &flag)) {
        double tot[2];
        reduce_partials<2>(partials, tot, lds);
        if (threadIdx.x == 0) {
            double rzn = tot[0], rrn = tot[1];
            scal->beta = rzn / scal->rz;
            scal->rz_prev = scal->rz;
            scal->rz = rzn;
            scal->rr = rrn;
            scal->iters += 1;
            double thr = fmax(scal->tol2 * scal->bnorm2, scal->delta2 * scal->rmax2);
            if (!isfinite(rrn) || !isfinite(rzn)) {
                scal->done = 1; scal->reason = 4;
            } else if (rrn <= thr) {
                scal->done = 1; scal->reason = 1;
            }
            *counter = 0u;
        }
    }
}

// ---------------------------------------------------------------------------
// out = b - A x (FP64), masked at fixed DOFs; b may be null (0).  Last
// workgroup stores ||out||^2 into *rr_out.
__global__ void __launch_bounds__(256)
k_resid64(int N, const int *__restrict__ rowptr, const int *__restrict__ col,
          const double *__restrict__ A, int64_t nbp, const unsigned char *__restrict__ fixmask,
          const double *__restrict__ xin, const double *__restrict__ b, double *__restrict__ out,
          double *rr_out, double *partials, unsigned *counter)
{
    __shared__ double lds[8];
    __shared__ int flag;
    const int lane = threadIdx.x & 15;
    const int sub = threadIdx.x >> 4;
    int rb, re;
    wg_rows(N, &rb, &re);
    double acc = 0.0;
    for (int base = rb; base < re; base += eolc::kRowsPerWG) {
        int i = base + sub;
        double y0 = 0, y1 = 0, y2 = 0;
        if (i < re) {
            int kb = rowptr[i], ke = rowptr[i + 1];
            for (int k = kb + lane; k < ke; k += 16) {
                int j = col[k];
                double u0 = xin[3 * j], u1 = xin[3 * j + 1], u2 = xin[3 * j + 2];
                const double *a = A + k;
                y0 += a[0 * nbp] * u0 + a[1 * nbp] * u1 + a[2 * nbp] * u2;
                y1 += a[3 * nbp] * u0 + a[4 * nbp] * u1 + a[5 * nbp] * u2;
                y2 += a[6 * nbp] * u0 + a[7 * nbp] * u1 + a[8 * nbp] * u2;
            }
        }
#pragma unroll
        for (int m = 8; m > 0; m >>= 1) {
            y0 += shfl_xor16(y0, m);
            y1 += shfl_xor16(y1, m);
            y2 += shfl_xor16(y2, m);
        }
        if (i < re && lane == 0) {
            double o0 = (b ? b[3 * i] : 0.0) - y0;
            double o1 = (b ? b[3 * i + 1] : 0.0) - y1;
            double o2 = (b ? b[3 * i + 2] : 0.0) - y2;
            unsigned char fm = fixmask[i];
            if (fm & 1) o0 = 0;
            if (fm & 2) o1 = 0;
            if (fm & 4) o2 = 0;
            out[3 * i] = o0; out[3 * i + 1] = o1; out[3 * i + 2] = o2;
            acc += (o0 * o0 + o1 * o1) + o2 * o2;
        }
    }
    double v[1] = { acc };
    block_sum<1>(v, lds);
    if (publish_partials<1>(v, partials, counter, &flag)) {
        double tot[1];
        reduce_partials<1>(partials, tot, lds);
        if (threadIdx.x == 0) {
            *rr_out = tot[0];
            *counter = 0u;
        }
    }
}

// z = D r (r converted from FP64 to T when T is float), partial r.z and r.r;
// last workgroup sets rz (and beta = rz_new / rz_old when keep_p).
template <class T>
__global__ void __launch_bounds__(256)
k_precond_init(int N, const T *__restrict__ D, const double *r64, T *r,
               T *__restrict__ z, int keep_p, SolverScalars *scal, double *partials, unsigned *counter)
{
    __shared__ double lds[8];
    __shared__ int flag;
    int rb, re;
    wg_rows(N, &rb, &re);
    double rz = 0.0, rr = 0.0;
    for (int i = rb + threadIdx.x; i < re; i += blockDim.x) {
        T r0 = (T)r64[3 * i], r1 = (T)r64[3 * i + 1], r2 = (T)r64[3 * i + 2];
        r[3 * i] = r0; r[3 * i + 1] = r1; r[3 * i + 2] = r2;
        const T *d = D + i;
        T z0 = d[0 * N] * r0 + d[1 * N] * r1 + d[2 * N] * r2;
        T z1 = d[3 * N] * r0 + d[4 * N] * r1 + d[5 * N] * r2;
        T z2 = d[6 * N] * r0 + d[7 * N] * r1 + d[8 * N] * r2;
        z[3 * i] = z0; z[3 * i + 1] = z1; z[3 * i + 2] = z2;
        rz += ((double)r0 * (double)z0 + (double)r1 * (double)z1) + (double)r2 * (double)z2;
        rr += ((double)r0 * (double)r0 + (double)r1 * (double)r1) + (double)r2 * (double)r2;
    }
    double v[2] = { rz, rr };
    block_sum<2>(v, lds);
    if (publish_partials<2>(v, partials, counter, &flag)) {
        double tot[2];
        reduce_partials<2>(partials, tot, lds);
        if (threadIdx.x == 0) {
            if (keep_p) {
                scal->beta = tot[0] / scal->rz;
            } else {
                scal->beta = 0.0;
            }
            scal->rz_prev = scal->rz;
            scal->rz = tot[0];
            scal->rr = tot[1];
            *counter = 0u;
        }
    }
}

// Block-Jacobi inverse of the diagonal blocks, with fixed DOFs replaced by
// identity rows/columns.  kind 0: 3x3 blocks; kind 1: diagonal only (Jacobi).
__global__ void __launch_bounds__(256)
k_block_jacobi(int N, const int *__restrict__ diagpos, const double *__restrict__ A, int64_t nbp,
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
        double val = A[c * nbp + k];
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
    // fixed DOFs: z must stay 0 there
#pragma unroll
    for (int c = 0; c < 9; ++c) {
        int r = c / 3, s = c % 3;
        if (((fm >> r) & 1) || ((fm >> s) & 1)) o[c] = 0.0;
        D64[c * N + i] = o[c];
        D32[c * N + i] = (float)o[c];
    }
}

__global__ void k_set_fixed(int n, const int *__restrict__ dofs, const double *__restrict__ vals,
                            unsigned char *__restrict__ fixmask, double *__restrict__ vfix, int set)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    int d = dofs[t];
    int vtx = d / 3, c = d % 3;
    if (set) {
        __hip_atomic_fetch_or(reinterpret_cast<unsigned int *>(fixmask) + (vtx >> 2),
                              1u << ((vtx & 3) * 8 + c), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        vfix[d] = vals[t];
    } else {
        fixmask[vtx] = 0;
        vfix[d] = 0.0;
    }
}

// x0 (free DOFs) = warm start v_old, fixed DOFs 0
__global__ void k_masked_copy(int N, const double *__restrict__ vin, const unsigned char *__restrict__ fixmask,
                              double *__restrict__ xout)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * N) return;
    int vtx = t / 3, c = t - 3 * vtx;
    xout[t] = ((fixmask[vtx] >> c) & 1) ? 0.0 : vin[t];
}

// xs64 += xl32; xl32 = 0 (mixed mode, before a reliable update)
__global__ void k_accumulate(int n, double *__restrict__ xs, float *__restrict__ xl)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    xs[t] += (double)xl[t];
    xl[t] = 0.0f;
}

// v = xs on free DOFs, vfix on fixed DOFs
__global__ void k_write_solution(int N, const double *__restrict__ xs, const unsigned char *__restrict__ fixmask,
                                 const double *__restrict__ vfix, double *__restrict__ v)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 3 * N) return;
    int vtx = t / 3, c = t - 3 * vtx;
    v[t] = ((fixmask[vtx] >> c) & 1) ? vfix[t] : xs[t];
}

// reset: set the tolerances and clear the counters; otherwise only clear `done`.
__global__ void k_scal_set(SolverScalars *s, double tol2, double delta2, int reset)
{
    if (reset) {
        s->tol2 = tol2;
        s->delta2 = delta2;
        s->iters = 0;
    }
    s->done = 0;
    s->reason = 0;
}

int grid_rows(int N)
{
    int g = (N + 4 * eolc::kRowsPerWG - 1) / (4 * eolc::kRowsPerWG);  // >= 4 passes per workgroup
    if (g > eolc::kMaxPartials) g = eolc::kMaxPartials;
    g = (g + 7) & ~7;
    if (g < 8) g = 8;
    if (g > eolc::kMaxPartials) g = eolc::kMaxPartials;
    return g;
}

inline int blocks_for(int64_t n, int t = 256) { return (int)((n + t - 1) / t); }

}  // namespace

namespace eolc {

hipError_t pcg_setup(eolc_ctx *c, bool /*mixed*/, int precond_kind)
{
    int N = c->topo.N;
    hipLaunchKernelGGL(k_block_jacobi, dim3(blocks_for(N)), dim3(256), 0, c->stream, N, c->diagpos,
                       c->A64, c->topo.nbp, c->fixmask, precond_kind, c->D64, c->D32);
    return hipGetLastError();
}

hipError_t set_fixed(eolc_ctx *c, int n, const int *d_dofs, const double *d_vals, int set)
{
    if (n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_set_fixed, dim3(blocks_for(n)), dim3(256), 0, c->stream, n, d_dofs, d_vals,
                       c->fixmask, c->vfix, set);
    return hipGetLastError();
}

// b_red = rhs - A vfix; x0 = masked v_old; r = b_red - A x0; z, rz, beta = 0.
hipError_t pcg_init(eolc_ctx *c, bool mixed, double tol)
{
    int N = c->topo.N;
    int G = grid_rows(N);
    const int64_t nbp = c->topo.nbp;
    double *rr_slot = &c->scal->bnorm2;
    hipLaunchKernelGGL(k_resid64, dim3(G), dim3(256), 0, c->stream, N, c->rowptr, c->col, c->A64, nbp,
                       c->fixmask, c->vfix, c->rhs, c->b64, rr_slot, c->partials, c->counter);
    hipLaunchKernelGGL(k_masked_copy, dim3(blocks_for(3 * N)), dim3(256), 0, c->stream, N, c->v, c->fixmask,
                       c->xs64);
    hipLaunchKernelGGL(k_resid64, dim3(G), dim3(256), 0, c->stream, N, c->rowptr, c->col, c->A64, nbp,
                       c->fixmask, c->xs64, c->b64, c->r64, &c->scal->rmax2, c->partials, c->counter);
    hipError_t e = hipMemsetAsync(mixed ? (void *)c->p32[0] : (void *)c->p64[0], 0,
                                  (size_t)3 * N * (mixed ? sizeof(float) : sizeof(double)), c->stream);
    if (e != hipSuccess) return e;
    if (mixed) {
        e = hipMemsetAsync(c->xl32, 0, (size_t)3 * N * sizeof(float), c->stream);
        if (e != hipSuccess) return e;
        hipLaunchKernelGGL(k_precond_init<float>, dim3(G), dim3(256), 0, c->stream, N, c->D32, c->r64, c->r32,
                           c->z32, 0, c->scal, c->partials, c->counter);
    } else {
        hipLaunchKernelGGL(k_precond_init<double>, dim3(G), dim3(256), 0, c->stream, N, c->D64, c->r64, c->r64,
                           c->z64, 0, c->scal, c->partials, c->counter);
    }
    (void)tol;
    return hipGetLastError();
}

hipError_t pcg_chunk(eolc_ctx *c, bool mixed, int iters, int *parity)
{
    int N = c->topo.N;
    int G = grid_rows(N);
    const int64_t nbp = c->topo.nbp;
    for (int it = 0; it < iters; ++it) {
        int pin = *parity & 1, pout = pin ^ 1;
        if (mixed) {
            hipLaunchKernelGGL(k_pcg_spmv<float>, dim3(G), dim3(256), 0, c->stream, N, c->rowptr, c->col, c->A32, nbp,
                               c->fixmask, c->z32, c->p32[pin], c->p32[pout], c->q32, c->scal, c->partials,
                               c->counter);
            hipLaunchKernelGGL((k_pcg_update<float, float>), dim3(G), dim3(256), 0, c->stream, N, c->D32, c->p32[pout],
                               c->q32, c->xl32, c->r32, c->z32, c->scal, c->partials, c->counter);
        } else {
            hipLaunchKernelGGL(k_pcg_spmv<double>, dim3(G), dim3(256), 0, c->stream, N, c->rowptr, c->col, c->A64, nbp,
                               c->fixmask, c->z64, c->p64[pin], c->p64[pout], c->q64, c->scal, c->partials,
                               c->counter);
            hipLaunchKernelGGL((k_pcg_update<double, double>), dim3(G), dim3(256), 0, c->stream, N, c->D64,
                               c->p64[pout], c->q64, c->xs64, c->r64, c->z64, c->scal, c->partials, c->counter);
        }
        *parity = pout;
    }
    return hipGetLastError();
}

// Mixed mode: fold the FP32 correction into the FP64 solution and replace the
// recursive residual by the true FP64 one (Clark et al., reliable updates).
// Stores ||r_true||^2 into scal->rmax2; the caller decides whether to go on.
hipError_t pcg_reliable_update(eolc_ctx *c, double /*tol*/, int /*parity*/)
{
    int N = c->topo.N;
    int G = grid_rows(N);
    hipLaunchKernelGGL(k_accumulate, dim3(blocks_for(3 * N)), dim3(256), 0, c->stream, 3 * N, c->xs64, c->xl32);
    hipLaunchKernelGGL(k_resid64, dim3(G), dim3(256), 0, c->stream, N, c->rowptr, c->col, c->A64, c->topo.nbp,
                       c->fixmask, c->xs64, c->b64, c->r64, &c->scal->rmax2, c->partials, c->counter);
    return hipGetLastError();
}

hipError_t pcg_restart_inner(eolc_ctx *c)
{
    int N = c->topo.N;
    int G = grid_rows(N);
    hipLaunchKernelGGL(k_precond_init<float>, dim3(G), dim3(256), 0, c->stream, N, c->D32, c->r64, c->r32, c->z32,
                       1, c->scal, c->partials, c->counter);
    return hipGetLastError();
}

hipError_t pcg_finish(eolc_ctx *c, bool /*mixed*/)
{
    int N = c->topo.N;
    hipLaunchKernelGGL(k_write_solution, dim3(blocks_for(3 * N)), dim3(256), 0, c->stream, N, c->xs64, c->fixmask,
                       c->vfix, c->v);
    return hipGetLastError();
}

hipError_t scal_set(eolc_ctx *c, double tol2, double delta2, int reset)
{
    hipLaunchKernelGGL(k_scal_set, dim3(1), dim3(1), 0, c->stream, c->scal, tol2, delta2, reset);
    return hipGetLastError();
}

// ||rhs - A v||^2 over free DOFs (v = final solution incl. fixed values) -> scal->rr
hipError_t spmv_true_residual(eolc_ctx *c, double * /*unused*/)
{
    int N = c->topo.N;
    int G = grid_rows(N);
    hipLaunchKernelGGL(k_resid64, dim3(G), dim3(256), 0, c->stream, N, c->rowptr, c->col, c->A64, c->topo.nbp,
                       c->fixmask, c->v, c->rhs, c->r64, &c->scal->rr, c->partials, c->counter);
    return hipGetLastError();
}

}  // namespace eolc

namespace eolc {
// Back-to-back launches of the SpMV kernel for timing (scalars reset each launch).
hipError_t time_spmv(eolc_ctx *c, bool mixed, int reps, float *ms)
{
    int N = c->topo.N;
    int G = grid_rows(N);
    hipEvent_t a, b;
    hipError_t e = hipEventCreate(&a);
    if (e != hipSuccess) return e;
    e = hipEventCreate(&b);
    if (e != hipSuccess) return e;
    scal_set(c, 0.0, 0.0, 1);
    hipEventRecord(a, c->stream);
    for (int r = 0; r < reps; ++r) {
        if (mixed)
            hipLaunchKernelGGL(k_pcg_spmv<float>, dim3(G), dim3(256), 0, c->stream, N, c->rowptr, c->col, c->A32,
                               c->topo.nbp, c->fixmask, c->z32, c->p32[0], c->p32[1], c->q32, c->scal, c->partials,
                               c->counter);
        else
            hipLaunchKernelGGL(k_pcg_spmv<double>, dim3(G), dim3(256), 0, c->stream, N, c->rowptr, c->col, c->A64,
                               c->topo.nbp, c->fixmask, c->z64, c->p64[0], c->p64[1], c->q64, c->scal, c->partials,
                               c->counter);
    }
    hipEventRecord(b, c->stream);
    e = hipEventSynchronize(b);
    if (e == hipSuccess) e = hipEventElapsedTime(ms, a, b);
    hipEventDestroy(a);
    hipEventDestroy(b);
    return e == hipSuccess ? hipGetLastError() : e;
}
}  // namespace eolc
