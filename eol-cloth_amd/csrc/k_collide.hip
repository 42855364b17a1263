// Proximity queries on the GPU (gfx950, FP64): btc::pointTriCollision
// (boxTriCollision.cpp:1067-1224) and btc::boxTriCollision (602-1063).
//
// Compiled with -ffp-contract=off: every expression is evaluated in the order the
// reference (through Eigen's 3-vector kernels) evaluates it -- dot products
// (a0 b0 + a1 b1) + a2 b2, cross products component-wise, normalized() as
// v / sqrt(squaredNorm) when squaredNorm > 0 -- so the records are bitwise equal
// to a CPU evaluation of the same algorithm.  The only transcendental, acos, is
// used in threshold comparisons (E2-E1 test); whenever a compared angle lies
// within kAcosBand of its threshold the pair is flagged and the host decides it
// with the C library's acos, so the pair set does not depend on the device acos.
//
// Work decomposition:
//   k_cd_prep       jittered positions and jittered face normals
//   k_cd_aabb       cloth AABB (min/max are exact: order-free)
//   k_grid_*        uniform-grid hash of the cloth faces: each face is entered in
//                   every cell its AABB, grown by a margin that bounds
//                   |x - projection| and the barycentric rounding, overlaps; faces
//                   too ill-shaped for that bound go to an always-tested list.  A
//                   query point tests the faces of its own cell, so it sees every
//                   face the brute-force loop would accept, and the lexicographic
//                   minimum (dist, face index) reproduces the reference's
//                   "strictly closer wins, first wins ties".
//   k_cd_query      one workgroup per query (point obstacle or box corner)
//   k_cd_v2t1       one thread per cloth vertex (box V2-T1)
//   k_cd_ee         one thread per cloth edge, 12 box edges each (box E2-E1)
//   k_cd_eolpts     one thread per cloth vertex (pointTri EOL branch)
// Variable-length outputs are compacted in reference order by count -> scan ->
// write passes (no atomics on record positions).
#include <mutex>

#include "eolc_internal.h"
#include "collide_common.h"

using namespace eolc_cd;

namespace {

constexpr double kAcosBand = 1e-12;

__device__ __forceinline__ void ld(const double *p, double *o) { o[0] = p[0]; o[1] = p[1]; o[2] = p[2]; }

// ---------------------------------------------------------------------------
__global__ void k_cd_prep(int N, int F, double thr, const double *__restrict__ x, const double *__restrict__ unit,
                          const int *__restrict__ tri, double *__restrict__ xj, double *__restrict__ fn)
{
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < 3 * N) xj[t] = x[t] + unit[t] * thr * 1e-3;
    (void)F; (void)tri; (void)fn;
}

__global__ void k_cd_fnormals(int F, const double *__restrict__ xj, const int *__restrict__ tri,
                              double *__restrict__ fn)
{
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    face_normal(xj + 3 * tri[3 * f], xj + 3 * tri[3 * f + 1], xj + 3 * tri[3 * f + 2], fn + 3 * f);
}

// cloth AABB of the jittered positions: per-block min/max, then a one-block pass
__global__ void __launch_bounds__(256) k_cd_aabb(int N, const double *__restrict__ xj, double *__restrict__ part)
{
    __shared__ double s[6][256];
    double lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < N; i += gridDim.x * blockDim.x)
        for (int k = 0; k < 3; ++k) {
            double v = xj[3 * i + k];
            lo[k] = fmin(lo[k], v);
            hi[k] = fmax(hi[k], v);
        }
    for (int k = 0; k < 3; ++k) { s[k][threadIdx.x] = lo[k]; s[3 + k][threadIdx.x] = hi[k]; }
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o)
            for (int k = 0; k < 3; ++k) {
                s[k][threadIdx.x] = fmin(s[k][threadIdx.x], s[k][threadIdx.x + o]);
                s[3 + k][threadIdx.x] = fmax(s[3 + k][threadIdx.x], s[3 + k][threadIdx.x + o]);
            }
        __syncthreads();
    }
    if (threadIdx.x == 0)
        for (int k = 0; k < 6; ++k) part[6 * blockIdx.x + k] = s[k][0];
}

__global__ void k_cd_aabb_final(int nb, const double *__restrict__ part, double *__restrict__ out)
{
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double r[6] = {INFINITY, INFINITY, INFINITY, -INFINITY, -INFINITY, -INFINITY};
    for (int b = 0; b < nb; ++b)
        for (int k = 0; k < 3; ++k) {
            r[k] = fmin(r[k], part[6 * b + k]);
            r[3 + k] = fmax(r[3 + k], part[6 * b + 3 + k]);
        }
    for (int k = 0; k < 6; ++k) out[k] = r[k];
}

// ---------------------------------------------------------------------------
// uniform-grid hash
struct GridParams {
    double org[3];
    double inv_cell;
    int dim[3];
    unsigned mask;      // table size - 1 (power of two)
    double margin;
    double cond_max;    // faces with |v0|^2 |v1|^2 / denom above this are always tested
};

__device__ __forceinline__ unsigned cell_hash(int ix, int iy, int iz, unsigned mask)
{
    unsigned h = (unsigned)ix * 73856093u ^ (unsigned)iy * 19349663u ^ (unsigned)iz * 83492791u;
    return h & mask;
}

__device__ __forceinline__ int clampi(int v, int lo, int hi) { return v < lo ? lo : (v > hi ? hi : v); }

__device__ __forceinline__ void face_cells(const GridParams &g, const double *a, const double *b, const double *c,
                                           int lo[3], int hi[3])
{
    for (int k = 0; k < 3; ++k) {
        double mn = fmin(a[k], fmin(b[k], c[k])) - g.margin;
        double mx = fmax(a[k], fmax(b[k], c[k])) + g.margin;
        lo[k] = clampi((int)floor((mn - g.org[k]) * g.inv_cell), 0, g.dim[k] - 1);
        hi[k] = clampi((int)floor((mx - g.org[k]) * g.inv_cell), 0, g.dim[k] - 1);
    }
}

__device__ __forceinline__ bool ill_shaped(const GridParams &g, const double *a, const double *b, const double *c)
{
    double v0[3], v1[3];
    v_sub(b, a, v0);
    v_sub(c, a, v1);
    double d00 = v_dot(v0, v0), d01 = v_dot(v0, v1), d11 = v_dot(v1, v1);
    double denom = d00 * d11 - d01 * d01;
    return !(denom > 0.0) || d00 * d11 > g.cond_max * denom;
}

// pass 0: count per bucket (and the always-tested list), pass 1: fill
__global__ void k_grid_insert(int F, int pass, GridParams g, const double *__restrict__ xj, const int *__restrict__ tri,
                              unsigned *__restrict__ bucket_cnt, const unsigned *__restrict__ bucket_start,
                              unsigned *__restrict__ cursor, int *__restrict__ entries, unsigned *__restrict__ nalways,
                              int *__restrict__ always)
{
    int f = blockIdx.x * blockDim.x + threadIdx.x;
    if (f >= F) return;
    const double *a = xj + 3 * tri[3 * f], *b = xj + 3 * tri[3 * f + 1], *c = xj + 3 * tri[3 * f + 2];
    if (ill_shaped(g, a, b, c)) {
        if (pass == 1) {
            unsigned p = atomicAdd(nalways, 1u);
            always[p] = f;
        }
        return;
    }
    int lo[3], hi[3];
    face_cells(g, a, b, c, lo, hi);
    for (int iz = lo[2]; iz <= hi[2]; ++iz)
        for (int iy = lo[1]; iy <= hi[1]; ++iy)
            for (int ix = lo[0]; ix <= hi[0]; ++ix) {
                unsigned h = cell_hash(ix, iy, iz, g.mask);
                if (pass == 0) {
                    atomicAdd(&bucket_cnt[h], 1u);
                } else {
                    unsigned p = atomicAdd(&cursor[h], 1u);
                    entries[bucket_start[h] + p] = f;
                }
            }
}

// exclusive scan of the bucket counts (single workgroup, sequential chunks)
__global__ void __launch_bounds__(1024) k_scan_u32(unsigned n, const unsigned *__restrict__ in,
                                                   unsigned *__restrict__ out, unsigned *__restrict__ total)
{
    __shared__ unsigned s[1024];
    __shared__ unsigned carry;
    if (threadIdx.x == 0) carry = 0;
    __syncthreads();
    for (unsigned base = 0; base < n; base += 1024) {
        unsigned i = base + threadIdx.x;
        unsigned v = i < n ? in[i] : 0u;
        s[threadIdx.x] = v;
        __syncthreads();
        for (unsigned o = 1; o < 1024; o <<= 1) {
            unsigned t = threadIdx.x >= o ? s[threadIdx.x - o] : 0u;
            __syncthreads();
            s[threadIdx.x] += t;
            __syncthreads();
        }
        if (i < n) out[i] = carry + s[threadIdx.x] - v;
        __syncthreads();
        if (threadIdx.x == 1023) carry += s[1023];
        __syncthreads();
    }
    if (threadIdx.x == 0) *total = carry;
}

// ---------------------------------------------------------------------------
// One query point against a set of faces (the loop of boxTriCollision.cpp:782-844
// and 1152-1212); the best (dist, face) pair is kept lexicographically.
struct Best {
    double dist;
    int face;
};

__device__ __forceinline__ void consider(const double *x1, const double *nor1, double thr, int j2,
                                         const double *__restrict__ xj, const int *__restrict__ tri,
                                         const double *__restrict__ fn, Best &best)
{
    const int *f2 = tri + 3 * j2;
    double nor2[3], dx[3], x2[3];
    ld(fn + 3 * j2, nor2);
    if (v_dot(nor1, nor2) < 0.0) v_neg(nor2);
    v_sub(x1, xj + 3 * f2[0], dx);
    double proj = v_dot(dx, nor2);
    if (proj < 0.0) return;
    for (int i = 0; i < 3; ++i) x2[i] = x1[i] - proj * nor2[i];
    v_sub(x2, x1, dx);
    double dist = v_norm(dx);
    if (dist > 5.0 * thr) return;
    double u, v;
    barycentric(&u, &v, xj + 3 * f2[0], xj + 3 * f2[1], xj + 3 * f2[2], x1);
    double w = 1.0 - u - v;
    if (u < 0.0 || 1.0 < u || v < 0.0 || 1.0 < v || w < 0.0 || 1.0 < w) return;
    if (best.face < 0 || dist < best.dist || (dist == best.dist && j2 < best.face)) {
        best.dist = dist;
        best.face = j2;
    }
}

__device__ __forceinline__ Best best_min(Best a, Best b)
{
    if (a.face < 0) return b;
    if (b.face < 0) return a;
    if (b.dist < a.dist || (b.dist == a.dist && b.face < a.face)) return b;
    return a;
}

// qpts/qnrm: query points and their normals; out_face[q] = winning face or -1.
__global__ void __launch_bounds__(256)
k_cd_query(int Q, GridParams g, const double *__restrict__ qpts, const double *__restrict__ qnrm, double thr,
           const double *__restrict__ clothaabb, const double *__restrict__ xj, const int *__restrict__ tri,
           const double *__restrict__ fn, const unsigned *__restrict__ bstart, const unsigned *__restrict__ bcnt,
           const int *__restrict__ entries, const unsigned *__restrict__ nalways, const int *__restrict__ always,
           int *__restrict__ out_face)
{
    __shared__ double sd[256];
    __shared__ int sf[256];
    const int q = blockIdx.x;
    if (q >= Q) return;
    double x1[3], n1[3];
    ld(qpts + 3 * q, x1);
    ld(qnrm + 3 * q, n1);
    double pt[6] = {x1[0], x1[1], x1[2], x1[0], x1[1], x1[2]};
    Best best{0.0, -1};
    if (check_aabb(pt, clothaabb)) {
        int ix = clampi((int)floor((x1[0] - g.org[0]) * g.inv_cell), 0, g.dim[0] - 1);
        int iy = clampi((int)floor((x1[1] - g.org[1]) * g.inv_cell), 0, g.dim[1] - 1);
        int iz = clampi((int)floor((x1[2] - g.org[2]) * g.inv_cell), 0, g.dim[2] - 1);
        unsigned h = cell_hash(ix, iy, iz, g.mask);
        const unsigned s0 = bstart[h], n = bcnt[h];
        for (unsigned e = threadIdx.x; e < n; e += blockDim.x) consider(x1, n1, thr, entries[s0 + e], xj, tri, fn, best);
        const unsigned na = *nalways;
        for (unsigned e = threadIdx.x; e < na; e += blockDim.x) consider(x1, n1, thr, always[e], xj, tri, fn, best);
    }
    sd[threadIdx.x] = best.dist;
    sf[threadIdx.x] = best.face;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if (threadIdx.x < o) {
            Best a{sd[threadIdx.x], sf[threadIdx.x]}, b{sd[threadIdx.x + o], sf[threadIdx.x + o]};
            Best m = best_min(a, b);
            sd[threadIdx.x] = m.dist;
            sf[threadIdx.x] = m.face;
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) out_face[q] = sf[0];
}

// Record of a query that found a face (V1-T2 / point-triangle), recomputed from the
// winner exactly as the reference builds it.
__global__ void k_cd_query_records(int Q, double thr, const double *__restrict__ qpts, const double *__restrict__ qnrm,
                                   const int *__restrict__ face, const double *__restrict__ xj,
                                   const int *__restrict__ tri, const double *__restrict__ fn,
                                   eolc_collision *__restrict__ out)
{
    int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= Q) return;
    eolc_collision c;
    rec_init(&c);
    int j2 = face[q];
    if (j2 >= 0) {
        double x1[3], nor1[3], nor2[3], dx[3], x2[3];
        ld(qpts + 3 * q, x1);
        ld(qnrm + 3 * q, nor1);
        const int *f2 = tri + 3 * j2;
        ld(fn + 3 * j2, nor2);
        if (v_dot(nor1, nor2) < 0.0) v_neg(nor2);
        v_sub(x1, xj + 3 * f2[0], dx);
        double proj = v_dot(dx, nor2);
        for (int i = 0; i < 3; ++i) x2[i] = x1[i] - proj * nor2[i];
        v_sub(x2, x1, dx);
        double u, v;
        barycentric(&u, &v, xj + 3 * f2[0], xj + 3 * f2[1], xj + 3 * f2[2], x1);
        c.dist = v_norm(dx);
        v_copy(nor1, c.nor1); v_copy(nor2, c.nor2);
        v_copy(x1, c.pos1); v_copy(x2, c.pos2);
        c.count1 = 1; c.count2 = 3;
        c.verts1[0] = q; c.verts1[1] = -1; c.verts1[2] = -1;
        c.verts2[0] = f2[0]; c.verts2[1] = f2[1]; c.verts2[2] = f2[2];
        c.weights1[0] = 1.0;
        c.weights2[0] = u; c.weights2[1] = v; c.weights2[2] = 1.0 - u - v;
        c.tri2 = j2;
    }
    out[q] = c;
}

// ---------------------------------------------------------------------------
// Box V2-T1 (boxTriCollision.cpp:675-764): one thread per cloth vertex.
__global__ void __launch_bounds__(256)
k_cd_v2t1(int N, int F, double thr, const BoxGeom *__restrict__ boxp, const double *__restrict__ xj, const double *__restrict__ fn,
          int *__restrict__ flag, eolc_collision *__restrict__ rec, int *__restrict__ err)
{
    int i2 = blockIdx.x * blockDim.x + threadIdx.x;
    if (i2 >= N) return;
    flag[i2] = 0;
    const BoxGeom &box = *boxp;
    double x2[3];
    ld(xj + 3 * i2, x2);
    double pt[6] = {x2[0], x2[1], x2[2], x2[0], x2[1], x2[2]};
    if (!check_aabb(pt, box.aabb)) return;
    for (int j1 = 0; j1 < 24; ++j1) {
        double dx[3];
        v_sub(x2, box.v[kBoxFaces[j1][0]], dx);
        if (v_dot(box.fn[j1], dx) > 0.0) return;
    }
    eolc_collision best;
    int found = 0;
    for (int j1 = 0; j1 < 24; ++j1) {
        const double *x1a = box.v[kBoxFaces[j1][0]], *x1b = box.v[kBoxFaces[j1][1]], *x1c = box.v[kBoxFaces[j1][2]];
        const double *nor1 = box.fn[j1];
        double dx[3], x1[3];
        v_sub(x2, x1a, dx);
        double proj = v_dot(nor1, dx);
        if (proj > 0.0) continue;
        for (int i = 0; i < 3; ++i) x1[i] = x2[i] - proj * nor1[i];
        v_sub(x2, x1, dx);
        double dist = v_norm(dx);
        if (dist > 5.0 * thr) continue;
        double u, v;
        barycentric(&u, &v, x1a, x1b, x1c, x1);
        double w = 1.0 - u - v;
        if (u < 0.0 || 1.0 < u || v < 0.0 || 1.0 < v || w < 0.0 || 1.0 < w) continue;
        if (i2 >= F) {   // faceNors2.col(i2) would read past the face normals
            atomicOr(err, 1);
            continue;
        }
        double nor2[3];
        ld(fn + 3 * i2, nor2);
        if (v_dot(nor2, nor1) < 0.0) v_neg(nor2);
        if (!found || dist < best.dist) {
            rec_init(&best);
            best.dist = dist;
            v_copy(nor1, best.nor1); v_copy(nor2, best.nor2);
            v_copy(x1, best.pos1); v_copy(x2, best.pos2);
            best.count1 = 3; best.count2 = 1;
            for (int i = 0; i < 3; ++i) best.verts1[i] = kBoxFaces[j1][i];
            best.verts2[0] = i2; best.verts2[1] = -1; best.verts2[2] = -1;
            best.weights1[0] = u; best.weights1[1] = v; best.weights1[2] = w;
            best.weights2[0] = 1.0;
            best.tri1 = j1;
            found = 1;
        }
    }
    if (found) {
        flag[i2] = 1;
        rec[i2] = best;
    }
}

// pointTri EOL branch (boxTriCollision.cpp:1099-1139): nearest point within thr.
__global__ void __launch_bounds__(256)
k_cd_eolpts(int N, int P, double thr, const double *__restrict__ pts, const double *__restrict__ nrm,
            const double *__restrict__ xj, int *__restrict__ flag, eolc_collision *__restrict__ rec)
{
    int i2 = blockIdx.x * blockDim.x + threadIdx.x;
    if (i2 >= N) return;
    flag[i2] = 0;
    double x2[3];
    ld(xj + 3 * i2, x2);
    int bi = -1;
    double bd = 0.0;
    for (int i1 = 0; i1 < P; ++i1) {
        double dx[3];
        v_sub(x2, pts + 3 * i1, dx);
        double dist = v_norm(dx);
        if (dist < thr && (bi < 0 || dist < bd)) { bi = i1; bd = dist; }
    }
    if (bi < 0) return;
    eolc_collision c;
    rec_init(&c);
    c.dist = bd;
    ld(nrm + 3 * bi, c.nor1); ld(nrm + 3 * bi, c.nor2);
    ld(pts + 3 * bi, c.pos1); v_copy(x2, c.pos2);
    c.count1 = 3; c.count2 = 1;
    c.verts1[0] = bi; c.verts1[1] = -1; c.verts1[2] = -1;
    c.verts2[0] = i2; c.verts2[1] = -1; c.verts2[2] = -1;
    c.weights1[0] = 1.0; c.weights2[0] = 1.0;
    flag[i2] = 1;
    rec[i2] = c;
}

// ---------------------------------------------------------------------------
// Box E2-E1: one thread per cloth edge (ee_edge in collide_common.h); pass 0
// counts the records and flags acos-ambiguous pairs, pass 1 writes the records
// at their prefix-sum offsets.
__global__ void __launch_bounds__(256)
k_cd_ee(int E, int pass, double thr, const BoxGeom *__restrict__ boxp, const int *__restrict__ cedges, const double *__restrict__ x,
        const double *__restrict__ xj, const int *__restrict__ tri, unsigned *__restrict__ counts,
        const unsigned *__restrict__ offsets, int *__restrict__ ambig, eolc_collision *__restrict__ out)
{
    int k2 = blockIdx.x * blockDim.x + threadIdx.x;
    if (k2 >= E) return;
    const BoxGeom &box = *boxp;
    int mask = 0;
    // AABB quick reject against the box (the per-edge test below pads by 1e-3 too)
    const int *e2 = cedges + 4 * k2;
    double aabbE[6];
    for (int i = 0; i < 3; ++i) {
        double a = xj[3 * e2[0] + i], b = xj[3 * e2[1] + i];
        aabbE[i] = a < b ? a : b;
        aabbE[i + 3] = a < b ? b : a;
    }
    if (!check_aabb(box.aabb, aabbE)) {
        if (pass == 0) { counts[k2] = 0; ambig[k2] = 0; }
        return;
    }
    if (pass == 0) {
        counts[k2] = (unsigned)ee_edge(k2, thr, box, cedges, x, xj, tri, nullptr, &mask, kAcosBand);
        ambig[k2] = mask;
    } else {
        ee_edge(k2, thr, box, cedges, x, xj, tri, out + offsets[k2], &mask, kAcosBand);
    }
}

// compaction of per-vertex flagged records in vertex order
__global__ void k_compact(int n, const int *__restrict__ flag, const unsigned *__restrict__ offs,
                          const eolc_collision *__restrict__ rec, eolc_collision *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n || !flag[i]) return;
    out[offs[i]] = rec[i];
}

__global__ void k_flag_to_u32(int n, const int *__restrict__ flag, unsigned *__restrict__ out)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = flag[i] ? 1u : 0u;
}

inline int blocks_for(int64_t n, int t = 256) { return (int)((n + t - 1) / t); }

// Device memory of eolc_collide_noalloc: one array per GPU, held from the loading of
// the code object until the process ends.  Host code never passes `scratch` itself
// to a kernel or a copy (in host code the name is a stand-in, not a device address):
// it uses the address that hipGetSymbolAddress returns for the current device.
__device__ __align__(256) unsigned char scratch[eolc::kCollideScratchBytes];

constexpr int kMaxScratchDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxScratchDevices] = {};
std::mutex g_scratch_call_mu[kMaxScratchDevices];

}  // namespace

namespace eolc {

hipError_t collide_scratch(int device, void **ptr)
{
    *ptr = nullptr;
    if (device < 0 || device >= kMaxScratchDevices) return hipErrorInvalidDevice;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        int cur = -1;
        hipError_t e = hipGetDevice(&cur);
        if (e != hipSuccess) return e;
        if (cur != device && (e = hipSetDevice(device)) != hipSuccess) return e;
        void *p = nullptr;
        e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess) return e;
        if (!p) return hipErrorInvalidSymbol;
        g_scratch_ptr[device] = p;
    }
    *ptr = g_scratch_ptr[device];
    return hipSuccess;
}

std::mutex *collide_scratch_lock(int device)
{
    return (device >= 0 && device < kMaxScratchDevices) ? &g_scratch_call_mu[device] : nullptr;
}

// Launch wrappers used by collide_host.cpp --------------------------------------
hipError_t cd_prep(eolc_ctx *c, double thr, double *xj, double *fn, double *aabb_part, double *aabb)
{
    const int N = c->topo.N, F = c->topo.F;
    hipLaunchKernelGGL(k_cd_prep, dim3(blocks_for(3 * N)), dim3(256), 0, c->stream, N, F, thr, c->x, c->cd_unit,
                       c->tri, xj, fn);
    hipLaunchKernelGGL(k_cd_fnormals, dim3(blocks_for(F)), dim3(256), 0, c->stream, F, xj, c->tri, fn);
    const int nb = 256;
    hipLaunchKernelGGL(k_cd_aabb, dim3(nb), dim3(256), 0, c->stream, N, xj, aabb_part);
    hipLaunchKernelGGL(k_cd_aabb_final, dim3(1), dim3(64), 0, c->stream, nb, aabb_part, aabb);
    return hipGetLastError();
}

hipError_t cd_scan(eolc_ctx *c, unsigned n, const unsigned *in, unsigned *out, unsigned *total)
{
    hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(1024), 0, c->stream, n, in, out, total);
    return hipGetLastError();
}

hipError_t cd_grid_build(eolc_ctx *c, const eolc_cd::CdGrid &gp, const double *xj, unsigned *bcnt, unsigned *bstart,
                         unsigned *cursor, int *entries, unsigned *nalways, int *always, unsigned *total, bool fill)
{
    const int F = c->topo.F;
    GridParams g;
    for (int k = 0; k < 3; ++k) { g.org[k] = gp.org[k]; g.dim[k] = gp.dim[k]; }
    g.inv_cell = gp.inv_cell;
    g.mask = gp.mask;
    g.margin = gp.margin;
    g.cond_max = gp.cond_max;
    if (!fill) {
        hipLaunchKernelGGL(k_grid_insert, dim3(blocks_for(F)), dim3(256), 0, c->stream, F, 0, g, xj, c->tri, bcnt,
                           bstart, cursor, entries, nalways, always);
        hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(1024), 0, c->stream, gp.mask + 1, bcnt, bstart, total);
    } else {
        hipLaunchKernelGGL(k_grid_insert, dim3(blocks_for(F)), dim3(256), 0, c->stream, F, 1, g, xj, c->tri, bcnt,
                           bstart, cursor, entries, nalways, always);
    }
    return hipGetLastError();
}

hipError_t cd_query(eolc_ctx *c, const eolc_cd::CdGrid &gp, int Q, const double *qpts, const double *qnrm, double thr,
                    const double *aabb, const double *xj, const double *fn, const unsigned *bstart, const unsigned *bcnt,
                    const int *entries, const unsigned *nalways, const int *always, int *face, eolc_collision *rec)
{
    if (Q <= 0) return hipSuccess;
    GridParams g;
    for (int k = 0; k < 3; ++k) { g.org[k] = gp.org[k]; g.dim[k] = gp.dim[k]; }
    g.inv_cell = gp.inv_cell;
    g.mask = gp.mask;
    g.margin = gp.margin;
    g.cond_max = gp.cond_max;
    hipLaunchKernelGGL(k_cd_query, dim3(Q), dim3(256), 0, c->stream, Q, g, qpts, qnrm, thr, aabb, xj, c->tri, fn,
                       bstart, bcnt, entries, nalways, always, face);
    hipLaunchKernelGGL(k_cd_query_records, dim3(blocks_for(Q)), dim3(256), 0, c->stream, Q, thr, qpts, qnrm, face, xj,
                       c->tri, fn, rec);
    return hipGetLastError();
}

hipError_t cd_v2t1(eolc_ctx *c, double thr, const BoxGeom *box, const double *xj, const double *fn, int *flag,
                   eolc_collision *rec, int *err)
{
    const int N = c->topo.N;
    hipLaunchKernelGGL(k_cd_v2t1, dim3(blocks_for(N)), dim3(256), 0, c->stream, N, c->topo.F, thr, box, xj, fn, flag,
                       rec, err);
    return hipGetLastError();
}

hipError_t cd_eolpts(eolc_ctx *c, int P, double thr, const double *pts, const double *nrm, const double *xj, int *flag,
                     eolc_collision *rec)
{
    const int N = c->topo.N;
    hipLaunchKernelGGL(k_cd_eolpts, dim3(blocks_for(N)), dim3(256), 0, c->stream, N, P, thr, pts, nrm, xj, flag, rec);
    return hipGetLastError();
}

hipError_t cd_ee(eolc_ctx *c, int pass, double thr, const BoxGeom *box, const double *xj, unsigned *counts,
                 const unsigned *offsets, int *ambig, eolc_collision *out)
{
    const int E = c->n_cedges;
    if (E == 0) return hipSuccess;
    hipLaunchKernelGGL(k_cd_ee, dim3(blocks_for(E)), dim3(256), 0, c->stream, E, pass, thr, box, c->cedges, c->x, xj,
                       c->tri, counts, offsets, ambig, out);
    return hipGetLastError();
}

hipError_t cd_compact(eolc_ctx *c, int n, const int *flag, unsigned *u32, unsigned *offs, unsigned *total,
                      const eolc_collision *rec, eolc_collision *out, bool write)
{
    if (!write) {
        hipLaunchKernelGGL(k_flag_to_u32, dim3(blocks_for(n)), dim3(256), 0, c->stream, n, flag, u32);
        hipLaunchKernelGGL(k_scan_u32, dim3(1), dim3(1024), 0, c->stream, (unsigned)n, u32, offs, total);
    } else {
        hipLaunchKernelGGL(k_compact, dim3(blocks_for(n)), dim3(256), 0, c->stream, n, flag, offs, rec, out);
    }
    return hipGetLastError();
}

}  // namespace eolc
