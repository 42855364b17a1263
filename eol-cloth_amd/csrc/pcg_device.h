// Device helpers shared by the PCG kernels (k_pcg.hip, k_ml.hip): workgroup ->
// rows mapping, wavefront and workgroup sums, and the per-workgroup partial-sum
// arrays that the next kernel re-reduces in a fixed order.
#pragma once
#include "eolc_internal.h"

// no multiply-add contraction in the kernels that include this header
#pragma clang fp contract(off)

namespace {

constexpr int WG = eolc::kPcgWG;

__device__ __forceinline__ int logical_wg()
{
    int G = gridDim.x;
    int w = blockIdx.x;
    return (w & 7) * (G >> 3) + (w >> 3);
}

__device__ __forceinline__ void wg_rows(int N, int *begin, int *end)
{
    int G = gridDim.x;
    int per = (N + G - 1) / G;
    int lw = logical_wg();
    int b = lw * per, e = b + per;
    *begin = b < N ? b : N;
    *end = e < N ? e : N;
}

// A row kernel may run on fewer workgroups than the G logical ones (pcg_grid): the
// first G - gridDim.x workgroups (the first dispatched, one per XCD in turn) then
// also do logical workgroup gridDim.x + blockIdx.x.  Returns that index or -1.
__device__ __forceinline__ int extra_logical_wg(int G)
{
    const int x = (int)gridDim.x + (int)blockIdx.x;
    return x < G ? x : -1;
}

// rows of logical workgroup lw of G
__device__ __forceinline__ void wg_rows_at(int N, int G, int lw, int *begin, int *end)
{
    int per = (N + G - 1) / G;
    int b = lw * per, e = b + per;
    *begin = b < N ? b : N;
    *end = e < N ? e : N;
}

// slice owned by this wavefront (16 per workgroup)
__device__ __forceinline__ int wave_slice() { return logical_wg() * (WG / 64) + (threadIdx.x >> 6); }

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    return v;
}

// Sum over the workgroup of K values per thread; result broadcast to all threads.
template <int K>
__device__ __forceinline__ void block_allsum(double (&v)[K], double *lds /* K * WG/64 + K */)
{
    constexpr int NW = WG / 64;
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = wave_sum(v[k]);
        if (lane == 0) lds[k * NW + wid] = s;
    }
    __syncthreads();
    if (wid == 0) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            double s = lane < NW ? lds[k * NW + lane] : 0.0;
#pragma unroll
            for (int o = NW / 2; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
            if (lane == 0) lds[K * NW + k] = s;
        }
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; ++k) v[k] = lds[K * NW + k];
}

// Fixed-order sum of K partial arrays (stride kMaxPartials) of length G, done by
// every workgroup; the result is identical everywhere.
template <int K>
__device__ __forceinline__ void reduce_partials(const double *__restrict__ part, double (&out)[K], double *lds,
                                                int G = -1)
{
    if (G < 0) G = gridDim.x;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = 0.0;
        for (int i = threadIdx.x; i < G; i += WG) s += part[k * eolc::kMaxPartials + i];
        out[k] = s;
    }
    block_allsum<K>(out, lds);
}

// Two partial arrays in one pass (K1 sums of p1, then K2 of p2): the same per-array
// order as two reduce_partials calls, so the same bits, with one load round and one
// workgroup sum instead of two.  lds: (K1 + K2) * WG/64 + K1 + K2 doubles.
template <int K1, int K2>
__device__ __forceinline__ void reduce_partials2(const double *__restrict__ p1, const double *__restrict__ p2,
                                                 double (&o1)[K1], double (&o2)[K2], double *lds, int G = -1)
{
    if (G < 0) G = gridDim.x;
    double v[K1 + K2];
#pragma unroll
    for (int k = 0; k < K1 + K2; ++k) v[k] = 0.0;
    for (int i = threadIdx.x; i < G; i += WG) {
#pragma unroll
        for (int k = 0; k < K1; ++k) v[k] += p1[k * eolc::kMaxPartials + i];
#pragma unroll
        for (int k = 0; k < K2; ++k) v[K1 + k] += p2[k * eolc::kMaxPartials + i];
    }
    block_allsum<K1 + K2>(v, lds);
#pragma unroll
    for (int k = 0; k < K1; ++k) o1[k] = v[k];
#pragma unroll
    for (int k = 0; k < K2; ++k) o2[k] = v[K1 + k];
}

template <int K>
__device__ __forceinline__ void write_partials(double (&v)[K], double *__restrict__ part, double *lds, int lw = -1)
{
    block_allsum<K>(v, lds);
    if (threadIdx.x == 0) {
        if (lw < 0) lw = logical_wg();
#pragma unroll
        for (int k = 0; k < K; ++k) part[k * eolc::kMaxPartials + lw] = v[k];
    }
}

template <class T>
__device__ __forceinline__ void load3(const T *__restrict__ a, int i, double &x, double &y, double &z)
{
    x = (double)a[3 * i]; y = (double)a[3 * i + 1]; z = (double)a[3 * i + 2];
}

// Partials layout (doubles, stride kMaxPartials): PA at slot 0, PB[parity] (r.z, r.r)
// at slots 1 + 2 parity and 2 + 2 parity, residual-norm partials at slot 5.
__device__ __forceinline__ double *PA(double *p) { return p; }
__device__ __forceinline__ double *PB(double *p, int parity) { return p + (1 + 2 * parity) * eolc::kMaxPartials; }

}  // namespace
