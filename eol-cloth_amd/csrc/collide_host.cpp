// Host side of the proximity queries: CD2 / CD (Collisions.cpp:11-77) on top of the
// gfx950 kernels in k_collide.hip.
//
//  - the jitter stream is the reference's own: std::mt19937 seeded with 1 feeding
//    std::uniform_real_distribution<double>(-1, 1) (boxTriCollision.cpp:649-659),
//    generated once per topology and kept on the device as unit values;
//  - the cloth edge table is createEdges (boxTriCollision.cpp:141-231): face edges
//    stable-sorted by the 32-bit hash kmin + (3F + 1) kmax -- which wraps modulo 2^32
//    for large meshes, and that wrap-around is reproduced -- twins paired when
//    adjacent with equal hashes.  Topology only; built at set-up;
//  - box geometry (createBox, boxTriCollision.cpp:400-420) is computed here per call;
//  - records are assembled in the reference's order: point records, then per box
//    V2-T1, V1-T2, E2-E1, corner de-duplication and the pos1_ snap.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "collide_common.h"
#include "eolc_internal.h"

using namespace eolc_cd;

namespace eolc {
hipError_t cd_prep(eolc_ctx *c, double thr, double *xj, double *fn, double *aabb_part, double *aabb);
hipError_t cd_grid_build(eolc_ctx *c, const CdGrid &gp, const double *xj, unsigned *bcnt, unsigned *bstart,
                         unsigned *cursor, int *entries, unsigned *nalways, int *always, unsigned *total, bool fill);
hipError_t cd_query(eolc_ctx *c, const CdGrid &gp, int Q, const double *qpts, const double *qnrm, double thr,
                    const double *aabb, const double *xj, const double *fn, const unsigned *bstart, const unsigned *bcnt,
                    const int *entries, const unsigned *nalways, const int *always, int *face, eolc_collision *rec);
hipError_t cd_v2t1(eolc_ctx *c, double thr, const BoxGeom *box, const double *xj, const double *fn, int *flag,
                   eolc_collision *rec, int *err);
hipError_t cd_eolpts(eolc_ctx *c, int P, double thr, const double *pts, const double *nrm, const double *xj, int *flag,
                     eolc_collision *rec);
hipError_t cd_ee(eolc_ctx *c, int pass, double thr, const BoxGeom *box, const double *xj, unsigned *counts,
                 const unsigned *offsets, int *ambig, eolc_collision *out);
hipError_t cd_compact(eolc_ctx *c, int n, const int *flag, unsigned *u32, unsigned *offs, unsigned *total,
                      const eolc_collision *rec, eolc_collision *out, bool write);
hipError_t cd_scan(eolc_ctx *c, unsigned n, const unsigned *in, unsigned *out, unsigned *total);
}  // namespace eolc

namespace {

#define EOLC_HIP(call)                                                                   \
    do {                                                                                 \
        hipError_t e__ = (call);                                                         \
        if (e__ != hipSuccess) {                                                         \
            c->err = std::string("HIP error ") + hipGetErrorString(e__) + " at " #call;  \
            return EOLC_ERR_HIP;                                                         \
        }                                                                                \
    } while (0)

// createBox (boxTriCollision.cpp:400-420) with createFaceNormals/createVertNormals.
void make_box(BoxGeom &B, const double *whd, const double *E1 /* column-major 4x4 */)
{
    double S[16] = {0}, E[16];
    S[0] = 0.5 * whd[0]; S[5] = 0.5 * whd[1]; S[10] = 0.5 * whd[2]; S[15] = 1.0;
    for (int i = 0; i < 4; ++i)
        for (int j = 0; j < 4; ++j) {
            double s = E1[0 * 4 + i] * S[j * 4 + 0];
            for (int k = 1; k < 4; ++k) s = s + E1[k * 4 + i] * S[j * 4 + k];
            E[j * 4 + i] = s;
        }
    for (int c = 0; c < 14; ++c) {
        const double vv[4] = {kBoxUnitVerts[c][0], kBoxUnitVerts[c][1], kBoxUnitVerts[c][2], 1.0};
        for (int i = 0; i < 3; ++i) {
            double s = E[0 * 4 + i] * vv[0];
            for (int k = 1; k < 4; ++k) s = s + E[k * 4 + i] * vv[k];
            B.v[c][i] = s;
        }
    }
    for (int f = 0; f < 24; ++f) face_normal(B.v[kBoxFaces[f][0]], B.v[kBoxFaces[f][1]], B.v[kBoxFaces[f][2]], B.fn[f]);
    double acc[14][3], ang[14];
    std::memset(acc, 0, sizeof(acc));
    std::memset(ang, 0, sizeof(ang));
    for (int f = 0; f < 24; ++f) {
        const double *xa = B.v[kBoxFaces[f][0]], *xb = B.v[kBoxFaces[f][1]], *xc = B.v[kBoxFaces[f][2]];
        double dba[3], dcb[3], dac[3], nd[3], nor[3], t[3];
        v_sub(xb, xa, dba); v_sub(xc, xb, dcb); v_sub(xa, xc, dac);
        v_copy(dac, nd); v_neg(nd);
        v_cross(dba, nd, nor);
        v_normalized(nor, nor); v_normalized(dba, dba); v_normalized(dcb, dcb); v_normalized(dac, dac);
        v_copy(dac, t); v_neg(t);
        const double a1 = std::acos(v_dot(dba, t));
        v_copy(dba, t); v_neg(t);
        const double a2 = std::acos(v_dot(dcb, t));
        v_copy(dcb, t); v_neg(t);
        const double a3 = std::acos(v_dot(dac, t));
        const int i0 = kBoxFaces[f][0], i1 = kBoxFaces[f][1], i2 = kBoxFaces[f][2];
        for (int k = 0; k < 3; ++k) {
            acc[i0][k] = acc[i0][k] + a1 * nor[k];
            acc[i1][k] = acc[i1][k] + a2 * nor[k];
            acc[i2][k] = acc[i2][k] + a3 * nor[k];
        }
        ang[i0] += a1; ang[i1] += a2; ang[i2] += a3;
    }
    for (int c = 0; c < 14; ++c) {
        double nn[3] = {acc[c][0] / ang[c], acc[c][1] / ang[c], acc[c][2] / ang[c]};
        v_normalized(nn, B.vn[c]);
    }
    for (int k = 0; k < 12; ++k) {
        v_copy(B.fn[kBoxEdgeFaces[k][0]], B.en[k][0]);
        v_copy(B.fn[kBoxEdgeFaces[k][1]], B.en[k][1]);
        B.eangle[k] = std::acos(v_dot(B.en[k][0], B.en[k][1]));
    }
    for (int i = 0; i < 3; ++i) {
        double lo = B.v[0][i], hi = B.v[0][i];
        for (int c = 1; c < 14; ++c) { lo = std::min(lo, B.v[c][i]); hi = std::max(hi, B.v[c][i]); }
        B.aabb[i] = lo; B.aabb[i + 3] = hi;
    }
    for (int f = 0; f < 24; ++f) {
        const double *xa = B.v[kBoxFaces[f][0]], *xb = B.v[kBoxFaces[f][1]], *xc = B.v[kBoxFaces[f][2]];
        for (int i = 0; i < 3; ++i) {
            double lo = xa[i], hi = xa[i];
            lo = std::min(xb[i], lo); lo = std::min(xc[i], lo);
            hi = std::max(xb[i], hi); hi = std::max(xc[i], hi);
            B.faabb[f][i] = lo; B.faabb[f][i + 3] = hi;
        }
    }
}

// Device memory of one collide call.  Heap mode (base == nullptr): every buffer is
// a hipMalloc, freed when the call ends or at release().  Arena mode: buffers are
// carved in order, 256-byte aligned, from a fixed region (the scratch array of
// eolc_collide_noalloc); release() rewinds to a mark, and no allocation is made.
// Reuse after a rewind is safe because every kernel runs on the context's one
// stream and every host read of a released buffer has been synchronised.
struct Workspace {
    char *base = nullptr;
    size_t cap = 0, used = 0, peak = 0;
    bool exhausted = false;
    std::vector<void *> owned;
    Workspace() = default;
    Workspace(const Workspace &) = delete;
    Workspace &operator=(const Workspace &) = delete;
    ~Workspace()
    {
        for (void *p : owned) (void)hipFree(p);
    }
    template <class T> hipError_t get(T **p, size_t n)
    {
        const size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
        if (!base) {
            void *q = nullptr;
            const hipError_t e = hipMalloc(&q, bytes);
            if (e == hipSuccess) {
                owned.push_back(q);
                *p = static_cast<T *>(q);
            }
            return e;
        }
        const size_t off = (used + 255) & ~(size_t)255;
        if (off > cap || bytes > cap - off) {
            exhausted = true;
            return hipErrorOutOfMemory;
        }
        *p = reinterpret_cast<T *>(base + off);
        used = off + bytes;
        peak = std::max(peak, used);
        return hipSuccess;
    }
    size_t mark() const { return base ? used : owned.size(); }
    void release(size_t m)
    {
        if (base) {
            used = m;
            return;
        }
        while (owned.size() > m) {
            (void)hipFree(owned.back());
            owned.pop_back();
        }
    }
};

}  // namespace

namespace eolc {

// createEdges order and pairing (topology only).  cedges: v0, v1, face0, face1|-1.
// The face-edges are ordered by their (signed, wrapping) 32-bit hash with ties in
// creation order, which is what the reference's stable sort gives; here a stable
// two-pass radix sort over the hash as an order-preserving unsigned key.
void build_collision_edges(eolc_ctx *c, int F, const int *tri)
{
    const int n = 3 * F;
    // items: key in the high word, creation index in the low word
    std::vector<uint64_t> item((size_t)n), tmp((size_t)n);
    for (int k = 0; k < F; ++k)
        for (int i = 0; i < 3; ++i) {
            const int a = tri[3 * k + i], b = tri[3 * k + (i + 1) % 3];
            const int kmin = std::min(a, b) + 1, kmax = std::max(a, b) + 1;
            const int hash = (int)((uint32_t)kmin + (uint32_t)(n + 1) * (uint32_t)kmax);   // int overflow wraps
            const uint32_t key = (uint32_t)hash ^ 0x80000000u;   // signed order as unsigned order
            item[3 * (size_t)k + i] = ((uint64_t)key << 32) | (uint32_t)(3 * k + i);
        }
    for (int shift = 32; shift < 64; shift += 16) {
        std::vector<int> cnt(65537, 0);
        for (int i = 0; i < n; ++i) cnt[((item[i] >> shift) & 0xffffu) + 1]++;
        for (int d = 0; d < 65536; ++d) cnt[d + 1] += cnt[d];
        for (int i = 0; i < n; ++i) tmp[cnt[(item[i] >> shift) & 0xffffu]++] = item[i];
        item.swap(tmp);
    }
    std::vector<int> &E = c->h_cedges;
    E.clear();
    E.reserve(4 * (size_t)n);
    int k = 0;
    while (k < n) {
        const int s = (int)(uint32_t)item[k], f0 = s / 3, i0 = s % 3;
        const int v0 = tri[3 * f0 + i0], v1 = tri[3 * f0 + (i0 + 1) % 3];
        if (k < n - 1 && (item[k + 1] >> 32) == (item[k] >> 32)) {
            E.insert(E.end(), {v0, v1, f0, (int)(uint32_t)item[k + 1] / 3});
            k += 2;
        } else {
            E.insert(E.end(), {v0, v1, f0, -1});
            k += 1;
        }
    }
    c->n_cedges = (int)(E.size() / 4);
}

// unit jitter values: the reference's random stream
void build_jitter(std::vector<double> &u, int N)
{
    std::mt19937 gen;
    std::uniform_real_distribution<> dis(-1.0, 1.0);
    gen.seed(1);
    u.resize(3 * (size_t)N);
    for (size_t i = 0; i < u.size(); ++i) u[i] = dis(gen);
}

}  // namespace eolc

namespace {

// Grid-hash table size for F faces: the smallest power of two >= max(2F, 64).
unsigned grid_table_size(int F)
{
    unsigned tbl = 1;
    while (tbl < (unsigned)std::max(2 * F, 64)) tbl <<= 1;
    return tbl;
}

// Device bytes a collide call needs before the data-dependent buffers (the grid's
// entry list and the record outputs), alignment included.
size_t collide_fixed_bytes(const eolc_ctx *c, int P, int B)
{
    const size_t N = (size_t)c->topo.N, F = (size_t)c->topo.F, E = (size_t)c->n_cedges;
    const size_t Q = std::max<size_t>((size_t)P + 8 * (size_t)B, 1), tbl = grid_table_size(c->topo.F);
    const size_t bytes = 8 * (3 * N + 3 * F + 6 * 256 + 6) + 4 * (3 * tbl + 2 + F) + 8 * 6 * Q +
                         (4 + sizeof(eolc_collision)) * Q + (4 + sizeof(eolc_collision) + 8) * N + 4 +
                         8 * 6 * (size_t)P + sizeof(BoxGeom) + 12 * E;
    return bytes + 256 * 32;
}

int collide_impl(eolc_ctx *c, Workspace &ws, double thr, int P, const double *pts, const double *nrm, int B,
                 const double *whd, const double *Eb, int eol_points, int box_npts, int box_nedges,
                 std::vector<eolc_collision> &out)
{
    const int N = c->topo.N, F = c->topo.F;
    double *xj = nullptr, *fn = nullptr, *aabb_part = nullptr, *aabb = nullptr;
    EOLC_HIP(ws.get(&xj, 3 * (size_t)N));
    EOLC_HIP(ws.get(&fn, 3 * (size_t)F));
    EOLC_HIP(ws.get(&aabb_part, 6 * 256));
    EOLC_HIP(ws.get(&aabb, 6));
    EOLC_HIP(eolc::cd_prep(c, thr, xj, fn, aabb_part, aabb));
    double hab[6];
    EOLC_HIP(hipMemcpyAsync(hab, aabb, sizeof(hab), hipMemcpyDeviceToHost, c->stream));
    EOLC_HIP(hipStreamSynchronize(c->stream));
    // boxes
    std::vector<BoxGeom> boxes((size_t)B);
    for (int b = 0; b < B; ++b) make_box(boxes[b], whd + 3 * b, Eb + 16 * b);
    // ---- uniform grid over the cloth faces
    CdGrid g;
    const double L = std::max({hab[3] - hab[0], hab[4] - hab[1], hab[5] - hab[2], 1e-300});
    g.margin = 5.0 * thr * (1.0 + 1e-6) + 1e-7 * L + 1e-12;
    const double area = std::max((hab[3] - hab[0]) * (hab[4] - hab[1]), 1e-300);
    double cell = std::max(2.0 * g.margin, std::sqrt(area / std::max(F, 1)) * 4.0);
    int dims[3];
    for (;;) {
        double tot = 1.0;
        for (int k = 0; k < 3; ++k) {
            dims[k] = std::max(1, (int)std::ceil((hab[3 + k] - hab[k] + 2 * g.margin) / cell) + 1);
            tot *= dims[k];
        }
        if (tot <= 1 << 24) break;
        cell *= 1.5;
    }
    for (int k = 0; k < 3; ++k) { g.org[k] = hab[k] - g.margin; g.dim[k] = dims[k]; }
    g.inv_cell = 1.0 / cell;
    const unsigned tbl = grid_table_size(F);
    g.mask = tbl - 1;
    g.cond_max = 1e6;
    unsigned *bcnt = nullptr, *bstart = nullptr, *cursor = nullptr, *total = nullptr, *nalways = nullptr;
    int *always = nullptr, *entries = nullptr;
    EOLC_HIP(ws.get(&bcnt, tbl));
    EOLC_HIP(ws.get(&bstart, tbl));
    EOLC_HIP(ws.get(&cursor, tbl));
    EOLC_HIP(ws.get(&total, 1));
    EOLC_HIP(ws.get(&nalways, 1));
    EOLC_HIP(ws.get(&always, (size_t)F));
    EOLC_HIP(hipMemsetAsync(bcnt, 0, tbl * sizeof(unsigned), c->stream));
    EOLC_HIP(hipMemsetAsync(cursor, 0, tbl * sizeof(unsigned), c->stream));
    EOLC_HIP(hipMemsetAsync(nalways, 0, sizeof(unsigned), c->stream));
    EOLC_HIP(eolc::cd_grid_build(c, g, xj, bcnt, bstart, cursor, nullptr, nalways, always, total, false));
    unsigned nentries = 0;
    EOLC_HIP(hipMemcpyAsync(&nentries, total, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
    EOLC_HIP(hipStreamSynchronize(c->stream));
    EOLC_HIP(ws.get(&entries, nentries));
    EOLC_HIP(eolc::cd_grid_build(c, g, xj, bcnt, bstart, cursor, entries, nalways, always, total, true));
    // ---- point / corner queries against the faces
    const int Q = P + 8 * B;
    std::vector<double> hq(3 * (size_t)std::max(Q, 1)), hn(3 * (size_t)std::max(Q, 1));
    for (int i = 0; i < 3 * P; ++i) { hq[i] = pts[i]; hn[i] = nrm[i]; }
    for (int b = 0; b < B; ++b)
        for (int i1 = 0; i1 < 8; ++i1)
            for (int k = 0; k < 3; ++k) {
                hq[3 * (P + 8 * b + i1) + k] = boxes[b].v[i1][k];
                hn[3 * (P + 8 * b + i1) + k] = boxes[b].vn[i1][k];
            }
    double *qpts = nullptr, *qnrm = nullptr;
    int *qface = nullptr;
    eolc_collision *qrec = nullptr;
    EOLC_HIP(ws.get(&qpts, hq.size()));
    EOLC_HIP(ws.get(&qnrm, hn.size()));
    EOLC_HIP(ws.get(&qface, std::max(Q, 1)));
    EOLC_HIP(ws.get(&qrec, std::max(Q, 1)));
    std::vector<eolc_collision> qres((size_t)std::max(Q, 1));
    if (Q > 0) {
        EOLC_HIP(hipMemcpyAsync(qpts, hq.data(), hq.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
        EOLC_HIP(hipMemcpyAsync(qnrm, hn.data(), hn.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
        EOLC_HIP(eolc::cd_query(c, g, Q, qpts, qnrm, thr, aabb, xj, fn, bstart, bcnt, entries, nalways, always, qface,
                                qrec));
        EOLC_HIP(hipMemcpyAsync(qres.data(), qrec, sizeof(eolc_collision) * Q, hipMemcpyDeviceToHost, c->stream));
    }
    // ---- per-vertex passes (EOL points and V2-T1) share flag/record buffers
    int *flag = nullptr, *err = nullptr;
    eolc_collision *vrec = nullptr;
    unsigned *u32 = nullptr, *offs = nullptr;
    EOLC_HIP(ws.get(&flag, (size_t)N));
    EOLC_HIP(ws.get(&vrec, (size_t)N));
    EOLC_HIP(ws.get(&u32, (size_t)N));
    EOLC_HIP(ws.get(&offs, (size_t)N));
    EOLC_HIP(ws.get(&err, 1));
    EOLC_HIP(hipMemsetAsync(err, 0, sizeof(int), c->stream));
    auto per_vertex_records = [&](std::vector<eolc_collision> &dst) -> int {
        EOLC_HIP(eolc::cd_compact(c, N, flag, u32, offs, total, nullptr, nullptr, false));
        unsigned cnt = 0;
        EOLC_HIP(hipMemcpyAsync(&cnt, total, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
        EOLC_HIP(hipStreamSynchronize(c->stream));
        dst.resize(cnt);
        if (cnt) {
            const size_t m = ws.mark();
            eolc_collision *o = nullptr;
            EOLC_HIP(ws.get(&o, cnt));
            EOLC_HIP(eolc::cd_compact(c, N, flag, u32, offs, total, vrec, o, true));
            EOLC_HIP(hipMemcpyAsync(dst.data(), o, sizeof(eolc_collision) * cnt, hipMemcpyDeviceToHost, c->stream));
            EOLC_HIP(hipStreamSynchronize(c->stream));
            ws.release(m);
        }
        return EOLC_OK;
    };
    const double snap = 0.1 * thr;
    auto snap_all = [&](std::vector<eolc_collision> &v, size_t from) {
        for (size_t k = from; k < v.size(); ++k)
            for (int i = 0; i < 3; ++i) v[k].pos1_[i] = v[k].pos1[i] - snap * v[k].nor1[i];
    };
    // ---- pointTriCollision
    if (eol_points && P > 0) {
        double *dpts = nullptr, *dnrm = nullptr;
        EOLC_HIP(ws.get(&dpts, 3 * (size_t)P));
        EOLC_HIP(ws.get(&dnrm, 3 * (size_t)P));
        EOLC_HIP(hipMemcpyAsync(dpts, pts, 3 * (size_t)P * sizeof(double), hipMemcpyHostToDevice, c->stream));
        EOLC_HIP(hipMemcpyAsync(dnrm, nrm, 3 * (size_t)P * sizeof(double), hipMemcpyHostToDevice, c->stream));
        EOLC_HIP(eolc::cd_eolpts(c, P, thr, dpts, dnrm, xj, flag, vrec));
        std::vector<eolc_collision> ev;
        int rc = per_vertex_records(ev);
        if (rc) return rc;
        out.insert(out.end(), ev.begin(), ev.end());
    } else {
        EOLC_HIP(hipStreamSynchronize(c->stream));
    }
    for (int i1 = 0; i1 < P; ++i1) {
        eolc_collision r = qres[i1];
        if (r.count1 == 0) continue;
        r.verts1[0] = i1;
        out.push_back(r);
    }
    snap_all(out, 0);
    // ---- boxTriCollision per box
    BoxGeom *dbox = nullptr;
    unsigned *ecnt = nullptr, *eoff = nullptr;
    int *eamb = nullptr;
    EOLC_HIP(ws.get(&dbox, 1));
    const int E = c->n_cedges;
    EOLC_HIP(ws.get(&ecnt, (size_t)E));
    EOLC_HIP(ws.get(&eoff, (size_t)E));
    EOLC_HIP(ws.get(&eamb, (size_t)E));
    std::vector<double> hx, hxj;   // host copies, only if an acos decision must be re-made
    const size_t base_after_points = out.size();
    for (int b = 0; b < B; ++b) {
        std::vector<eolc_collision> clst;
        EOLC_HIP(hipMemcpyAsync(dbox, &boxes[b], sizeof(BoxGeom), hipMemcpyHostToDevice, c->stream));
        // V2-T1
        EOLC_HIP(eolc::cd_v2t1(c, thr, dbox, xj, fn, flag, vrec, err));
        {
            std::vector<eolc_collision> vv;
            int rc = per_vertex_records(vv);
            if (rc) return rc;
            clst.insert(clst.end(), vv.begin(), vv.end());
        }
        // V1-T2
        for (int i1 = 0; i1 < 8; ++i1) {
            eolc_collision r = qres[P + 8 * b + i1];
            if (r.count1 == 0) continue;
            r.verts1[0] = i1;
            r.nedge1 = 3;
            for (int i = 0; i < 3; ++i) r.edge1[i] = kBoxVertEdges[i1][i];
            clst.push_back(r);
        }
        // E2-E1
        if (E > 0) {
            EOLC_HIP(eolc::cd_ee(c, 0, thr, dbox, xj, ecnt, nullptr, eamb, nullptr));
            EOLC_HIP(eolc::cd_scan(c, (unsigned)E, ecnt, eoff, total));
            unsigned ne = 0;
            EOLC_HIP(hipMemcpyAsync(&ne, total, sizeof(unsigned), hipMemcpyDeviceToHost, c->stream));
            std::vector<int> amb((size_t)E);
            EOLC_HIP(hipMemcpyAsync(amb.data(), eamb, sizeof(int) * E, hipMemcpyDeviceToHost, c->stream));
            EOLC_HIP(hipStreamSynchronize(c->stream));
            std::vector<eolc_collision> ee(ne);
            if (ne) {
                const size_t m = ws.mark();
                eolc_collision *o = nullptr;
                EOLC_HIP(ws.get(&o, ne));
                EOLC_HIP(eolc::cd_ee(c, 1, thr, dbox, xj, ecnt, eoff, eamb, o));
                EOLC_HIP(hipMemcpyAsync(ee.data(), o, sizeof(eolc_collision) * ne, hipMemcpyDeviceToHost, c->stream));
                EOLC_HIP(hipStreamSynchronize(c->stream));
                ws.release(m);
            }
            // re-decide edges whose angle tests were within the band, with the C library acos
            bool any = std::any_of(amb.begin(), amb.end(), [](int m) { return m != 0; });
            if (any) {
                if (hx.empty()) {
                    hx.resize(3 * (size_t)N);
                    hxj.resize(3 * (size_t)N);
                    EOLC_HIP(hipMemcpy(hx.data(), c->x, hx.size() * sizeof(double), hipMemcpyDeviceToHost));
                    EOLC_HIP(hipMemcpy(hxj.data(), xj, hxj.size() * sizeof(double), hipMemcpyDeviceToHost));
                }
                std::vector<eolc_collision> fixed;
                size_t pos = 0;
                for (int k2 = 0; k2 < E; ++k2) {
                    // records of edge k2 are contiguous and ordered by k2
                    size_t end = pos;
                    while (end < ee.size() && ee[end].edge2 == k2) ++end;
                    if (amb[k2]) {
                        eolc_collision tmp[12];
                        int mask = 0;
                        int n = ee_edge(k2, thr, boxes[b], c->h_cedges.data(), hx.data(), hxj.data(),
                                        c->topo_tri.data(), tmp, &mask, 0.0);
                        fixed.insert(fixed.end(), tmp, tmp + n);
                    } else {
                        fixed.insert(fixed.end(), ee.begin() + pos, ee.begin() + end);
                    }
                    pos = end;
                }
                ee.swap(fixed);
                c->cd_host_decisions += 1;
            }
            clst.insert(clst.end(), ee.begin(), ee.end());
        }
        // corner de-duplication (boxTriCollision.cpp:1020-1052)
        for (int i1 = 0; i1 < 8; ++i1) {
            int kmin = -1;
            double dmin = 1e9;
            for (size_t k = 0; k < clst.size(); ++k)
                if (clst[k].count1 == 1 && clst[k].verts1[0] == i1) {
                    double dx[3];
                    v_sub(clst[k].pos2, boxes[b].v[i1], dx);
                    double d = v_dot(dx, dx);
                    if (d < dmin) { kmin = (int)k; dmin = d; }
                }
            if (kmin != -1) {
                std::vector<size_t> dl;
                for (size_t k = 0; k < clst.size(); ++k)
                    if (clst[k].count1 == 1 && clst[k].verts1[0] == i1 && (int)k != kmin) dl.push_back(k);
                for (size_t kd : dl) {
                    clst[kd] = clst.back();
                    clst.pop_back();
                }
            }
        }
        snap_all(clst, 0);
        out.insert(out.end(), clst.begin(), clst.end());
        if (eol_points) {
            // CD: box feature indices shifted per obstacle (Collisions.cpp:38-49)
            const size_t start = out.size() - clst.size();
            for (size_t k = start; k < out.size(); ++k) {
                eolc_collision &r = out[k];
                const int base = P + b * box_npts + b * box_nedges;
                if (r.count1 == 1 && r.count2 == 3) r.verts1[0] = base + r.verts1[0];
                for (int e = 0; e < r.nedge1; ++e) r.edge1[e] = base + (box_npts + r.edge1[e]);
            }
        }
    }
    (void)base_after_points;
    int herr = 0;
    EOLC_HIP(hipMemcpy(&herr, err, sizeof(int), hipMemcpyDeviceToHost));
    if (herr) {
        c->err = "collide: a cloth vertex with index >= face count reached the face-normal lookup of V2-T1 "
                 "(boxTriCollision.cpp:731 indexes face normals by vertex index)";
        return EOLC_ERR_INVALID;
    }
    return EOLC_OK;
}

}  // namespace

extern "C" int eolc_collide(eolc_ctx *c, double threshold, int P, const double *pts, const double *nrm, int B,
                            const double *whd, const double *E, int eol_points, int box_npts, int box_nedges,
                            eolc_collision *out, int cap, int *count)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (!count) { c->err = "collide: null count"; return EOLC_ERR_INVALID; }
    *count = 0;
    if (!c->have_topo || !c->have_state) { c->err = "collide: topology and state must be set first"; return EOLC_ERR_INVALID; }
    if (!(threshold > 0.0) || P < 0 || B < 0 || (P && (!pts || !nrm)) || (B && (!whd || !E)) || cap < 0) {
        c->err = "collide: bad arguments";
        return EOLC_ERR_INVALID;
    }
    EOLC_HIP(hipSetDevice(c->device));
    std::vector<eolc_collision> recs;
    Workspace ws;   // heap mode: device buffers allocated for this call
    int rc = collide_impl(c, ws, threshold, P, pts, nrm, B, whd, E, eol_points, box_npts, box_nedges, recs);
    if (rc) return rc;
    *count = (int)recs.size();
    if ((int)recs.size() > cap || (!out && !recs.empty())) {
        c->err = "collide: output capacity too small";
        return EOLC_ERR_OVERFLOW;
    }
    if (!recs.empty()) std::memcpy(out, recs.data(), recs.size() * sizeof(eolc_collision));
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "collide");
}

// The same query with no device allocation: every buffer of the call is carved from
// this GPU's scratch array (eolc::kCollideScratchBytes), whose address was looked up
// once per device.  The call holds a per-device lock for its whole duration, so calls
// from several host threads on one GPU run one after the other; the buffers that
// must start at zero are cleared by each call (the array keeps the last call's data).
extern "C" int eolc_collide_noalloc(eolc_ctx *c, double threshold, int P, const double *pts, const double *nrm,
                                    int B, const double *whd, const double *E, int eol_points, int box_npts,
                                    int box_nedges, eolc_collision *out, int cap, int *count)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (!count) { c->err = "collide_noalloc: null count"; return EOLC_ERR_INVALID; }
    *count = 0;
    if (!c->have_topo || !c->have_state) {
        c->err = "collide_noalloc: topology and state must be set first";
        return EOLC_ERR_INVALID;
    }
    if (!(threshold > 0.0) || P < 0 || B < 0 || (P && (!pts || !nrm)) || (B && (!whd || !E)) || cap < 0) {
        c->err = "collide_noalloc: bad arguments";
        return EOLC_ERR_INVALID;
    }
    const size_t need = collide_fixed_bytes(c, P, B);
    if (need > eolc::kCollideScratchBytes) {
        c->err = "collide_noalloc: the query needs at least " + std::to_string(need >> 20) + " MiB of device memory (" +
                 std::to_string(c->topo.N) + " vertices, " + std::to_string(c->topo.F) + " faces), more than the " +
                 std::to_string(eolc::kCollideScratchBytes >> 20) + " MiB scratch; use eolc_collide";
        return EOLC_ERR_INVALID;
    }
    EOLC_HIP(hipSetDevice(c->device));
    void *base = nullptr;
    EOLC_HIP(eolc::collide_scratch(c->device, &base));
    std::lock_guard<std::mutex> lk(*eolc::collide_scratch_lock(c->device));
    Workspace ws;
    ws.base = static_cast<char *>(base);
    ws.cap = eolc::kCollideScratchBytes;
    std::vector<eolc_collision> recs;
    int rc = collide_impl(c, ws, threshold, P, pts, nrm, B, whd, E, eol_points, box_npts, box_nedges, recs);
    if (ws.exhausted) {
        (void)hipStreamSynchronize(c->stream);
        c->err = "collide_noalloc: the query's grid or record buffers need more device memory than the " +
                 std::to_string(eolc::kCollideScratchBytes >> 20) + " MiB scratch; use eolc_collide";
        return EOLC_ERR_INVALID;
    }
    if (rc) return rc;
    *count = (int)recs.size();
    if ((int)recs.size() > cap || (!out && !recs.empty())) {
        c->err = "collide_noalloc: output capacity too small";
        return EOLC_ERR_OVERFLOW;
    }
    if (!recs.empty()) std::memcpy(out, recs.data(), recs.size() * sizeof(eolc_collision));
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "collide_noalloc");
}
