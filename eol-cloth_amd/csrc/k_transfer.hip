// Velocity transfer after a remesh, non-EOL vertices (gfx950): the searches of
// Cloth::velocityTransfer (Cloth.cpp:195-230 and 265-281) on uniform grids.
//
// For every new vertex the reference scans
//   - all old vertices for the FIRST one (index order) whose material-space distance,
//     unsigned_vv_distance (geometry.cpp:228-230), is below 1e-6: the vertex existed
//     before and keeps that velocity;
//   - otherwise all old faces for the FIRST one that is_inside accepts
//     (get_enclosing_face, geometry.cpp:284-290; is_inside, geometry.cpp:276-281),
//     and takes the barycentric combination of its vertices' velocities
//     (get_barycentric_coords, geometry.cpp:253-271).
// That is O(N_new N_old + N_new F_old).  Here old vertices and old faces are binned
// once into a uniform grid over the old mesh's material-space bounding box (faces
// into every cell their bounding box, widened by the is_inside tolerance, overlaps),
// and each new vertex tests only the candidates in its cells, keeping the smallest
// index that passes: the same answer as the first hit of the full scan.  Bins are
// filled with atomics (order inside a bin is irrelevant to a minimum).  All the
// arithmetic is the reference's, in its order, with no contraction (this file is
// built with -ffp-contract=off), so indices, barycentric coordinates and velocities
// are bitwise those of the scan.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <vector>

#include "eolc_internal.h"

namespace {

struct Grid {
    double lox, loy, inv;   // origin and 1 / cell size
    int nx, ny;
};

__device__ __forceinline__ int cell_x(const Grid &g, double x)
{
    const double c = floor((x - g.lox) * g.inv);
    return c < 0.0 ? 0 : (c >= g.nx ? g.nx - 1 : (int)c);
}

__device__ __forceinline__ int cell_y(const Grid &g, double y)
{
    const double c = floor((y - g.loy) * g.inv);
    return c < 0.0 ? 0 : (c >= g.ny ? g.ny - 1 : (int)c);
}

__global__ void k_vbin(int n, const double *__restrict__ u, Grid g, int *__restrict__ cnt, int *__restrict__ slot,
                       int pass, const int *__restrict__ start, int *__restrict__ list)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int c = cell_y(g, u[2 * j + 1]) * g.nx + cell_x(g, u[2 * j]);
    if (pass == 0) {
        atomicAdd(&cnt[c], 1);
    } else {
        const int k = atomicAdd(&slot[c], 1);
        list[start[c] + k] = j;
    }
}

// Face bounding box widened so that every point is_inside accepts lies inside: the
// barycentric thresholds let a point leave the triangle by at most about 1.2e-5 of
// its longest edge.
__device__ __forceinline__ void face_box(const double *__restrict__ u, const int *__restrict__ t, double &x0,
                                         double &y0, double &x1, double &y1)
{
    const double ax = u[2 * t[0]], ay = u[2 * t[0] + 1], bx = u[2 * t[1]], by = u[2 * t[1] + 1];
    const double cx = u[2 * t[2]], cy = u[2 * t[2] + 1];
    x0 = fmin(ax, fmin(bx, cx)); x1 = fmax(ax, fmax(bx, cx));
    y0 = fmin(ay, fmin(by, cy)); y1 = fmax(ay, fmax(by, cy));
    const double m = 2e-5 * ((x1 - x0) + (y1 - y0)) + 1e-12;
    x0 -= m; y0 -= m; x1 += m; y1 += m;
}

__global__ void k_fbin(int F, const double *__restrict__ u, const int *__restrict__ tri, Grid g, int *__restrict__ cnt,
                       int *__restrict__ slot, int pass, const int *__restrict__ start, int *__restrict__ list)
{
    const int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    double x0, y0, x1, y1;
    face_box(u, tri + 3 * f, x0, y0, x1, y1);
    const int cx0 = cell_x(g, x0), cx1 = cell_x(g, x1), cy0 = cell_y(g, y0), cy1 = cell_y(g, y1);
    for (int cy = cy0; cy <= cy1; ++cy)
        for (int cx = cx0; cx <= cx1; ++cx) {
            const int c = cy * g.nx + cx;
            if (pass == 0) {
                atomicAdd(&cnt[c], 1);
            } else {
                const int k = atomicAdd(&slot[c], 1);
                list[start[c] + k] = f;
            }
        }
}

// Exclusive scan of n ints in three passes: per-block scans (1024 per block), a
// one-workgroup scan of the block totals, and the add-back.
constexpr int kScanWG = 1024;
__global__ void __launch_bounds__(kScanWG) k_scan_block(int n, const int *__restrict__ in, int *__restrict__ out,
                                                        int *__restrict__ totals)
{
    __shared__ int s[kScanWG];
    const int i = blockIdx.x * kScanWG + threadIdx.x;
    const int v = i < n ? in[i] : 0;
    s[threadIdx.x] = v;
    __syncthreads();
    for (int o = 1; o < kScanWG; o <<= 1) {
        const int t = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
        __syncthreads();
        s[threadIdx.x] += t;
        __syncthreads();
    }
    if (i < n) out[i] = s[threadIdx.x] - v;
    if (threadIdx.x == kScanWG - 1) totals[blockIdx.x] = s[threadIdx.x];
}

__global__ void __launch_bounds__(kScanWG) k_scan_totals(int nb, int *__restrict__ totals, int *__restrict__ sum)
{
    __shared__ int carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    __shared__ int s[kScanWG];
    for (int b0 = 0; b0 < nb; b0 += kScanWG) {
        const int i = b0 + threadIdx.x;
        const int v = i < nb ? totals[i] : 0;
        s[threadIdx.x] = v;
        __syncthreads();
        for (int o = 1; o < kScanWG; o <<= 1) {
            const int t = threadIdx.x >= o ? s[threadIdx.x - o] : 0;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < nb) totals[i] = carry + s[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 0) carry += s[kScanWG - 1];
        __syncthreads();
    }
    if (threadIdx.x == 0) *sum = carry;
}

__global__ void k_scan_add(int n, const int *__restrict__ totals, int *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] += totals[i / kScanWG];
}

// get_barycentric_coords (geometry.cpp:253-271) in the reference's order.
__device__ __forceinline__ void bary_coords(double px, double py, const double *__restrict__ u, const int *__restrict__ t,
                                            double &b0, double &b1, double &b2)
{
    const double cx = u[2 * t[2]], cy = u[2 * t[2] + 1];
    const double v0x = u[2 * t[0]] - cx, v0y = u[2 * t[0] + 1] - cy;
    const double v1x = u[2 * t[1]] - cx, v1y = u[2 * t[1] + 1] - cy;
    const double v2x = px - cx, v2y = py - cy;
    const double dot00 = v0x * v0x + v0y * v0y;
    const double dot01 = v0x * v1x + v0y * v1y;
    const double dot02 = v0x * v2x + v0y * v2y;
    const double dot11 = v1x * v1x + v1y * v1y;
    const double dot12 = v1x * v2x + v1y * v2y;
    const double invDenom = 1.0 / (dot00 * dot11 - dot01 * dot01);
    b0 = (dot11 * dot02 - dot01 * dot12) * invDenom;
    b1 = (dot00 * dot12 - dot01 * dot02) * invDenom;
    b2 = 1 - b0 - b1;
}

__global__ void k_transfer(int n_new, const double *__restrict__ u_new, const double *__restrict__ u_old,
                           const double *__restrict__ v_old, const int *__restrict__ tri, Grid g,
                           const int *__restrict__ vstart, const int *__restrict__ vlist,
                           const int *__restrict__ fstart, const int *__restrict__ flist, double *__restrict__ v_new,
                           int *__restrict__ src, int *__restrict__ face, double *__restrict__ bary,
                           int *__restrict__ missing)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_new) return;
    const double px = u_new[2 * n], py = u_new[2 * n + 1];
    // an old vertex within 1e-6 (cells overlapping the 1e-6 box around p)
    int best = INT_MAX;
    const int cx0 = cell_x(g, px - 1e-6), cx1 = cell_x(g, px + 1e-6);
    const int cy0 = cell_y(g, py - 1e-6), cy1 = cell_y(g, py + 1e-6);
    for (int cy = cy0; cy <= cy1; ++cy)
        for (int cx = cx0; cx <= cx1; ++cx) {
            const int c = cy * g.nx + cx;
            for (int k = vstart[c]; k < vstart[c + 1]; ++k) {
                const int j = vlist[k];
                if (j >= best) continue;
                const double dx = u_old[2 * j] - px, dy = u_old[2 * j + 1] - py;
                if (sqrt(dx * dx + dy * dy) < 1e-6) best = j;
            }
        }
    if (best != INT_MAX) {
        src[n] = best;
        face[n] = -1;
        for (int c = 0; c < 3; ++c) {
            v_new[3 * n + c] = v_old[3 * best + c];
            bary[3 * n + c] = 0.0;
        }
        return;
    }
    src[n] = -1;
    // is_inside (geometry.cpp:276-281): thresholds -10 EPSILON, -10 EPSILON,
    // -100 EPSILON with EPSILON = 1e-7f, products in float as in the C expression
    const float eps = 1e-7f;
    const double t01 = (double)(-10 * eps), t2 = (double)(-100 * eps);
    int bf = INT_MAX;
    double b0 = 0.0, b1 = 0.0, b2 = 0.0;
    const int c = cell_y(g, py) * g.nx + cell_x(g, px);
    for (int k = fstart[c]; k < fstart[c + 1]; ++k) {
        const int f = flist[k];
        if (f >= bf) continue;
        double q0, q1, q2;
        bary_coords(px, py, u_old, tri + 3 * f, q0, q1, q2);
        if (q0 >= t01 && q1 >= t01 && q2 >= t2) {
            bf = f;
            b0 = q0; b1 = q1; b2 = q2;
        }
    }
    if (bf == INT_MAX) {
        face[n] = -1;
        for (int cc = 0; cc < 3; ++cc) {
            v_new[3 * n + cc] = 0.0;
            bary[3 * n + cc] = 0.0;
        }
        atomicAdd(missing, 1);
        return;
    }
    face[n] = bf;
    const int *t = tri + 3 * bf;
    bary[3 * n] = b0; bary[3 * n + 1] = b1; bary[3 * n + 2] = b2;
    for (int cc = 0; cc < 3; ++cc)
        v_new[3 * n + cc] = b0 * v_old[3 * t[0] + cc] + b1 * v_old[3 * t[1] + cc] + b2 * v_old[3 * t[2] + cc];
}

// EOL correction of a new LAG vertex (Cloth.cpp:254-262): each vertex of its old face
// that was an EOL node contributes v - F V instead of v, F the old face's deformation
// gradient (deform_grad, UtilEOL.cpp:13-28, at the old world positions).  Recomputes
// v_new only for vertices whose face has an EOL vertex, in the reference's order:
// vw = v + (-F) V, then b0 vwA + b1 vwB + b2 vwC.
__global__ void k_transfer_eol(int n_new, const int *__restrict__ face, const double *__restrict__ bary,
                               const int *__restrict__ tri, const double *__restrict__ x_old,
                               const double *__restrict__ u_old, const double *__restrict__ v_old,
                               const int *__restrict__ eidx, const double *__restrict__ V_old,
                               double *__restrict__ v_new)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_new || face[n] < 0) return;
    const int *t = tri + 3 * face[n];
    if (eidx[t[0]] < 0 && eidx[t[1]] < 0 && eidx[t[2]] < 0) return;
    const int a = t[0], b = t[1], c = t[2];
    const double d00 = u_old[2 * b] - u_old[2 * a], d01 = u_old[2 * c] - u_old[2 * a];
    const double d10 = u_old[2 * b + 1] - u_old[2 * a + 1], d11 = u_old[2 * c + 1] - u_old[2 * a + 1];
    const double det = d00 * d11 - d01 * d10;
    const double i00 = d11 / det, i01 = -d01 / det, i10 = -d10 / det, i11 = d00 / det;
    double F[3][2];
    for (int r = 0; r < 3; ++r) {
        const double e0 = x_old[3 * b + r] - x_old[3 * a + r], e1 = x_old[3 * c + r] - x_old[3 * a + r];
        F[r][0] = e0 * i00 + e1 * i10;
        F[r][1] = e0 * i01 + e1 * i11;
    }
    double vw[3][3];
    for (int k = 0; k < 3; ++k) {
        const int j = t[k];
        for (int r = 0; r < 3; ++r) vw[k][r] = v_old[3 * j + r];
        if (eidx[j] >= 0) {
            const double V0 = V_old[2 * eidx[j]], V1 = V_old[2 * eidx[j] + 1];
            for (int r = 0; r < 3; ++r) vw[k][r] = vw[k][r] + ((-F[r][0]) * V0 + (-F[r][1]) * V1);
        }
    }
    const double b0 = bary[3 * n], b1 = bary[3 * n + 1], b2 = bary[3 * n + 2];
    for (int r = 0; r < 3; ++r) v_new[3 * n + r] = b0 * vw[0][r] + b1 * vw[1][r] + b2 * vw[2][r];
}

inline int blocks_for(int64_t n, int t = 256) { return (int)((n + t - 1) / t); }

// exclusive scan of cnt[0..n) into start[0..n], start[n] = total
hipError_t scan(hipStream_t st, int n, const int *cnt, int *start, int *totals)
{
    const int nb = (n + kScanWG - 1) / kScanWG;
    hipLaunchKernelGGL(k_scan_block, dim3(nb), dim3(kScanWG), 0, st, n, cnt, start, totals);
    hipLaunchKernelGGL(k_scan_totals, dim3(1), dim3(kScanWG), 0, st, nb, totals, start + n);
    hipLaunchKernelGGL(k_scan_add, dim3(blocks_for(n)), dim3(256), 0, st, n, totals, start);
    return hipGetLastError();
}

struct Events {
    hipEvent_t a = nullptr, b = nullptr;
    ~Events()
    {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

template <class T> struct DevBuf {
    T *p = nullptr;
    hipError_t alloc(size_t n) { return hipMalloc((void **)&p, std::max<size_t>(n, 1) * sizeof(T)); }
    ~DevBuf() { if (p) (void)hipFree(p); }
};

}  // namespace

namespace eolc {

// Returns hipSuccess and the number of new vertices with neither an old vertex nor an
// enclosing old face in *n_missing.  Host arrays in and out.
hipError_t velocity_transfer(eolc_ctx *c, int n_old, const double *u_old, const double *v_old, int f_old,
                             const int *tri_old, int n_new, const double *u_new, double *v_new, int *src, int *face,
                             double *bary, int *n_missing, float *ms, const double *x_old, const int *eol_old,
                             int n_eol, const double *V_old)
{
    hipStream_t st = c->stream;
    // grid over the old vertices' bounding box, about 4 vertices per cell
    double lx = u_old[0], hx = u_old[0], ly = u_old[1], hy = u_old[1];
    for (int j = 1; j < n_old; ++j) {
        lx = std::min(lx, u_old[2 * j]); hx = std::max(hx, u_old[2 * j]);
        ly = std::min(ly, u_old[2 * j + 1]); hy = std::max(hy, u_old[2 * j + 1]);
    }
    const double w = std::max(hx - lx, 1e-12), h = std::max(hy - ly, 1e-12);
    double s = 2.0 * std::sqrt(w * h / std::max(n_old, 1));
    s = std::max(s, std::max(w, h) / 2048.0);   // at most 2049 x 2049 cells
    Grid g;
    g.lox = lx; g.loy = ly; g.inv = 1.0 / s;
    g.nx = std::max(1, (int)std::ceil(w / s)) + 1;
    g.ny = std::max(1, (int)std::ceil(h / s)) + 1;
    const int ncell = g.nx * g.ny;
    DevBuf<double> du_old, dv_old, du_new, dv_new, dbary;
    DevBuf<int> dtri, vcnt, vstart, vslot, vlist, fcnt, fstart, fslot, flist, totals, dsrc, dface, dmiss;
    hipError_t e;
#define TRY(x) do { if ((e = (x)) != hipSuccess) return e; } while (0)
    TRY(du_old.alloc(2 * (size_t)n_old)); TRY(dv_old.alloc(3 * (size_t)n_old));
    TRY(du_new.alloc(2 * (size_t)n_new)); TRY(dv_new.alloc(3 * (size_t)n_new)); TRY(dbary.alloc(3 * (size_t)n_new));
    TRY(dtri.alloc(3 * (size_t)f_old));
    TRY(vcnt.alloc(ncell)); TRY(vstart.alloc(ncell + 1)); TRY(vslot.alloc(ncell)); TRY(vlist.alloc(n_old));
    TRY(fcnt.alloc(ncell)); TRY(fstart.alloc(ncell + 1)); TRY(fslot.alloc(ncell));
    TRY(totals.alloc((ncell + kScanWG - 1) / kScanWG));
    TRY(dsrc.alloc(n_new)); TRY(dface.alloc(n_new)); TRY(dmiss.alloc(1));
    TRY(hipMemcpyAsync(du_old.p, u_old, 2 * (size_t)n_old * sizeof(double), hipMemcpyHostToDevice, st));
    TRY(hipMemcpyAsync(dv_old.p, v_old, 3 * (size_t)n_old * sizeof(double), hipMemcpyHostToDevice, st));
    TRY(hipMemcpyAsync(du_new.p, u_new, 2 * (size_t)n_new * sizeof(double), hipMemcpyHostToDevice, st));
    TRY(hipMemcpyAsync(dtri.p, tri_old, 3 * (size_t)f_old * sizeof(int), hipMemcpyHostToDevice, st));
    DevBuf<double> dx_old, dV;
    DevBuf<int> deidx;
    if (eol_old) {
        TRY(dx_old.alloc(3 * (size_t)n_old)); TRY(deidx.alloc(n_old)); TRY(dV.alloc(2 * (size_t)n_eol));
        TRY(hipMemcpyAsync(dx_old.p, x_old, 3 * (size_t)n_old * sizeof(double), hipMemcpyHostToDevice, st));
        TRY(hipMemcpyAsync(deidx.p, eol_old, (size_t)n_old * sizeof(int), hipMemcpyHostToDevice, st));
        if (n_eol) TRY(hipMemcpyAsync(dV.p, V_old, 2 * (size_t)n_eol * sizeof(double), hipMemcpyHostToDevice, st));
    }
    Events ev;
    TRY(hipEventCreate(&ev.a));
    TRY(hipEventCreate(&ev.b));
    TRY(hipEventRecord(ev.a, st));
    TRY(hipMemsetAsync(vcnt.p, 0, ncell * sizeof(int), st));
    TRY(hipMemsetAsync(vslot.p, 0, ncell * sizeof(int), st));
    TRY(hipMemsetAsync(fcnt.p, 0, ncell * sizeof(int), st));
    TRY(hipMemsetAsync(fslot.p, 0, ncell * sizeof(int), st));
    TRY(hipMemsetAsync(dmiss.p, 0, sizeof(int), st));
    hipLaunchKernelGGL(k_vbin, dim3(blocks_for(n_old)), dim3(256), 0, st, n_old, du_old.p, g, vcnt.p, vslot.p, 0,
                       vstart.p, vlist.p);
    TRY(scan(st, ncell, vcnt.p, vstart.p, totals.p));
    hipLaunchKernelGGL(k_vbin, dim3(blocks_for(n_old)), dim3(256), 0, st, n_old, du_old.p, g, vcnt.p, vslot.p, 1,
                       vstart.p, vlist.p);
    hipLaunchKernelGGL(k_fbin, dim3(blocks_for(f_old)), dim3(256), 0, st, f_old, du_old.p, dtri.p, g, fcnt.p, fslot.p,
                       0, fstart.p, flist.p);
    TRY(scan(st, ncell, fcnt.p, fstart.p, totals.p));
    int nfl = 0;
    TRY(hipMemcpyAsync(&nfl, fstart.p + ncell, sizeof(int), hipMemcpyDeviceToHost, st));
    TRY(hipStreamSynchronize(st));
    TRY(flist.alloc(nfl));
    hipLaunchKernelGGL(k_fbin, dim3(blocks_for(f_old)), dim3(256), 0, st, f_old, du_old.p, dtri.p, g, fcnt.p, fslot.p,
                       1, fstart.p, flist.p);
    hipLaunchKernelGGL(k_transfer, dim3(blocks_for(n_new)), dim3(256), 0, st, n_new, du_new.p, du_old.p, dv_old.p,
                       dtri.p, g, vstart.p, vlist.p, fstart.p, flist.p, dv_new.p, dsrc.p, dface.p, dbary.p, dmiss.p);
    TRY(hipGetLastError());
    if (eol_old) {
        hipLaunchKernelGGL(k_transfer_eol, dim3(blocks_for(n_new)), dim3(256), 0, st, n_new, dface.p, dbary.p, dtri.p,
                           dx_old.p, du_old.p, dv_old.p, deidx.p, dV.p, dv_new.p);
        TRY(hipGetLastError());
    }
    TRY(hipEventRecord(ev.b, st));
    TRY(hipMemcpyAsync(v_new, dv_new.p, 3 * (size_t)n_new * sizeof(double), hipMemcpyDeviceToHost, st));
    TRY(hipMemcpyAsync(src, dsrc.p, (size_t)n_new * sizeof(int), hipMemcpyDeviceToHost, st));
    TRY(hipMemcpyAsync(face, dface.p, (size_t)n_new * sizeof(int), hipMemcpyDeviceToHost, st));
    if (bary) TRY(hipMemcpyAsync(bary, dbary.p, 3 * (size_t)n_new * sizeof(double), hipMemcpyDeviceToHost, st));
    TRY(hipMemcpyAsync(n_missing, dmiss.p, sizeof(int), hipMemcpyDeviceToHost, st));
    TRY(hipStreamSynchronize(st));
    float t = 0.f;
    if (ms && hipEventElapsedTime(&t, ev.a, ev.b) == hipSuccess) *ms = t;   // includes the one host read of the bin total
#undef TRY
    return hipSuccess;
}

}  // namespace eolc
