// C ABI of the per-step cloth hot path: context, topology set-up, Forces::fill,
// velocity solve (solver "none") and the step driver.  Host C++; every device
// computation is a hand-written gfx950 kernel (k_*.hip).
//
// Reference interfaces replaced (see include/eolc.h):
//   Forces::fill                      Forces.cpp:912-930
//   Cloth::solve                      Cloth.cpp:343-357
//   GeneralizedSolver::velocitySolve  GeneralizedSolver.cpp:107-172
//   Cloth::step                       Cloth.cpp:359-413
#include <algorithm>
#include <atomic>
#include <climits>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <new>
#include <thread>

#include "eolc_internal.h"

using eolc::Topology;

namespace {

#define EOLC_HIP(call)                                                                            \
    do {                                                                                          \
        hipError_t e__ = (call);                                                                  \
        if (e__ != hipSuccess) {                                                                  \
            c->err = std::string("HIP error ") + hipGetErrorString(e__) + " at " #call;           \
            return EOLC_ERR_HIP;                                                                  \
        }                                                                                         \
    } while (0)

int fail(eolc_ctx *c, int code, const std::string &msg)
{
    c->err = msg;
    return code;
}

template <class T> hipError_t dalloc(eolc_ctx *c, T **p, size_t n)
{
    if (*p) {
        hipError_t e = hipFree(*p);
        if (e != hipSuccess) return e;
        *p = nullptr;
    }
    if (n == 0) n = 1;
    hipError_t e = hipMalloc((void **)p, n * sizeof(T));
    if (e == hipSuccess) c->alloc_bytes += n * sizeof(T);
    return e;
}

template <class T> void dfree(T *&p)
{
    if (p) (void)hipFree(p);
    p = nullptr;
}

void free_eol(eolc_ctx *c)
{
    eolc::EolState &S = c->eol;
    dfree(S.d_elem); dfree(S.d_eidx); dfree(S.d_soff); dfree(S.d_foff); dfree(S.d_cptr); dfree(S.d_clist);
    dfree(S.d_fptr); dfree(S.d_flist); dfree(S.d_xv); dfree(S.d_xvptr); dfree(S.d_xvslot);
    dfree(S.d_stK); dfree(S.d_stM); dfree(S.d_stF); dfree(S.d_A); dfree(S.d_M); dfree(S.d_fX);
    S = eolc::EolState{};
}

void free_topology_buffers(eolc_ctx *c)
{
    dfree(c->X); dfree(c->tri); dfree(c->hinge);
    dfree(c->slice_off); dfree(c->srow); dfree(c->slen); dfree(c->scol); dfree(c->diagpos);
    dfree(c->cptr); dfree(c->clist); dfree(c->fptr); dfree(c->flist);
    dfree(c->stageK); dfree(c->stageF); dfree(c->stageM);
    dfree(c->x); dfree(c->v); dfree(c->f); dfree(c->rhs);
    dfree(c->A64); dfree(c->S32); dfree(c->M32); dfree(c->entry_is_diag); dfree(c->Mval); dfree(c->D64); dfree(c->D32);
    dfree(c->supper); dfree(c->uoff); dfree(c->loff); dfree(c->lrec); dfree(c->e2u); dfree(c->S32u); dfree(c->S32a);
    dfree(c->wwin);
    dfree(c->S64u); dfree(c->M64);
    dfree(c->pt_row); dfree(c->pt_cnt); dfree(c->pt_roff); dfree(c->pt_soff); dfree(c->pt_loff); dfree(c->pt_lo);
    dfree(c->pt_e2r); dfree(c->pt_e2l); dfree(c->pt_woff); dfree(c->pt_win); dfree(c->P32); dfree(c->P64);
    dfree(c->stiff); dfree(c->smask); dfree(c->sellbase); dfree(c->A64s);
    dfree(c->fixmask); dfree(c->vfix); dfree(c->fix_dofs); dfree(c->fix_vals);
    dfree(c->b64); dfree(c->xs64); dfree(c->r64); dfree(c->z64); dfree(c->q64);
    dfree(c->p64[0]); dfree(c->p64[1]);
    dfree(c->r32); dfree(c->z32); dfree(c->q32); dfree(c->xl32); dfree(c->xl64);
    dfree(c->p32[0]); dfree(c->p32[1]);
    dfree(c->cedges); dfree(c->cd_unit);
    eolc::LearnedState &L = c->lp;
    dfree(L.tpos); dfree(L.w); dfree(L.wT); dfree(L.X);
    for (auto &p : L.h) dfree(p);
    for (auto &p : L.a) dfree(p);
    for (auto &p : L.dpre) dfree(p);
    dfree(L.o); dfree(L.dO); dfree(L.dinp); dfree(L.gpart); dfree(L.dpart); dfree(L.scal);
    dfree(L.e); dfree(L.r); dfree(L.z); dfree(L.Az);
    L.N = 0;
    L.nbs = 0;
    eolc::BatchState &S = c->batch;
    dfree(S.d_inst); dfree(S.d_order); dfree(S.d_user_fixed); dfree(S.d_scripted); dfree(S.slice_part); dfree(S.stats); dfree(S.dstats);
    dfree(S.d_all_fixed);
    S.n_all_fixed = 0;
    dfree(S.d_ml_lo);
    S.ml = S.graph_ml = false;
    if (S.graph) (void)hipGraphExecDestroy(S.graph);
    S.graph = nullptr;
    S.B = 0;
    S.inst.clear(); S.order.clear(); S.fixed_nodes.clear(); S.user_fixed.clear();
    dfree(c->e_face); dfree(c->beta_hinge);
    free_eol(c);
    eolc::MLState &ml = c->ml;
    dfree(ml.hord); dfree(ml.ptr1); dfree(ml.agg1); dfree(ml.gpos); dfree(ml.cptr); dfree(ml.par); dfree(ml.lam);
    dfree(ml.V); dfree(ml.Dinv); dfree(ml.R); dfree(ml.C); dfree(ml.anc3); dfree(ml.kptr2);
    dfree(ml.geo); dfree(ml.gchunk); dfree(ml.gcp); dfree(ml.gpart); dfree(ml.D9); dfree(ml.feat2); dfree(ml.feat3);
    dfree(ml.R9p); dfree(ml.C9);
    ml.lin_ok = ml.lin_on = ml.lin_built = false;
    ml.hord_h.clear();
    ml.nchunks = 0;
    ml.built = false;
    ml.lv = eolc::MLLevels{};
    ml.counts.clear();
    ml.maps.clear();
    c->bj_iters_hint = c->bj_iters_hint_ml = 0;
    c->fp64_sticky = 0;
    c->alloc_bytes = 0;
}

void free_learned_weights(eolc_ctx *c)
{
    eolc::LearnedState &L = c->lp;
    dfree(L.theta); dfree(L.grad); dfree(L.adam_m); dfree(L.adam_v); dfree(L.wq);
    L.have_weights = false;
}

// Local blocks of a face (fillxMI / fillxxMI order) and of a hinge (fillxB / fillxxB).
const int kFacePairs[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
const int kHingePairs[10][2] = {{0, 0}, {1, 1}, {2, 2}, {3, 3}, {0, 1}, {0, 2}, {0, 3}, {1, 2}, {1, 3}, {2, 3}};

int find_block(const Topology &t, int i, int j)
{
    const int *b = t.col.data() + t.rowptr[i];
    const int *e = t.col.data() + t.rowptr[i + 1];
    const int *p = std::lower_bound(b, e, j);
    return (p != e && *p == j) ? (int)(p - t.col.data()) : -1;
}

// Runs fn(begin, end) over contiguous chunks of [0, n) on up to 16 host threads
// (one chunk, on the calling thread, below `grain` items per thread).  An exception
// in a worker is rethrown on the calling thread after all workers joined.
template <class Fn> void host_parallel(int64_t n, int64_t grain, Fn &&fn)
{
    const int64_t hw = std::max(1u, std::min(std::thread::hardware_concurrency(), 16u));
    const int T = (int)std::max<int64_t>(1, std::min(hw, n / std::max<int64_t>(grain, 1)));
    if (T <= 1) {
        fn((int64_t)0, n);
        return;
    }
    std::vector<std::exception_ptr> err(T);
    std::vector<std::thread> th;
    th.reserve(T);
    for (int k = 0; k < T; ++k)
        th.emplace_back([&fn, &err, n, k, T] {
            try {
                fn(n * k / T, n * (k + 1) / T);
            } catch (...) {
                err[k] = std::current_exception();
            }
        });
    for (auto &x : th) x.join();
    for (auto &e : err)
        if (e) std::rethrow_exception(e);
}

// Exclusive prefix sum in place over v[0..n) into v[0..n] (v has n + 1 entries,
// v[i + 1] holds the count of item i on entry); false on 32-bit overflow.
bool prefix_counts(std::vector<int> &v, int64_t n)
{
    const int T = (int)std::max<int64_t>(1, std::min<int64_t>(16, n / (1 << 16)));
    std::vector<int64_t> part(T + 1, 0);
    host_parallel(T, 1, [&](int64_t b, int64_t e) {
        for (int64_t k = b; k < e; ++k) {
            int64_t s = 0;
            for (int64_t i = n * k / T; i < n * (k + 1) / T; ++i) s += v[i + 1];
            part[k + 1] = s;
        }
    });
    for (int k = 0; k < T; ++k) part[k + 1] += part[k];
    if (part[T] > (int64_t)INT32_MAX) return false;
    host_parallel(T, 1, [&](int64_t b, int64_t e) {
        for (int64_t k = b; k < e; ++k) {
            int64_t s = part[k];
            for (int64_t i = n * k / T; i < n * (k + 1) / T; ++i) {
                s += v[i + 1];
                v[i + 1] = (int)s;
            }
        }
    });
    v[0] = 0;
    return true;
}

// Element incident to a vertex through force slot s (3f + v for faces, then
// 3F + 4h + v for hinges): its vertices, their count, the local index v and the
// element's first block slot (6f for faces, 6F + 10h for hinges).
struct Incidence {
    const int *verts;
    int nv, local;
    int64_t slot0;
};
inline Incidence incidence(int s, int F, const int *tri, const int *hinge)
{
    if (s < 3 * F) return {tri + 3 * (s / 3), 3, s % 3, 6 * (int64_t)(s / 3)};
    const int h = (s - 3 * F) / 4;
    return {hinge + 4 * h, 4, (s - 3 * F) % 4, 6 * (int64_t)F + 10 * (int64_t)h};
}

// Calls fn(column, code) for every block contribution of row `row`, elements in
// slot order: an element pair (a, b) puts slot << 1 on block (x_a, x_b) and, for
// x_a != x_b, (slot << 1) | 1 (the transpose) on block (x_b, x_a).
template <class Fn> void row_contributions(const Topology &t, const int *tri, const int *hinge, int row, Fn &&fn)
{
    for (int k = t.fptr[row]; k < t.fptr[row + 1]; ++k) {
        const Incidence in = incidence(t.flist[k], t.F, tri, hinge);
        const int(*pairs)[2] = in.nv == 3 ? kFacePairs : kHingePairs;
        const int np = in.nv == 3 ? 6 : 10;
        for (int p = 0; p < np; ++p) {
            const int a = pairs[p][0], b = pairs[p][1];
            if (a == in.local) fn(in.verts[b], (int)((in.slot0 + p) << 1));
            else if (b == in.local && in.verts[a] != in.verts[b]) fn(in.verts[a], (int)(((in.slot0 + p) << 1) | 1));
        }
    }
}

// Builds the BSR pattern and the gather maps.  Returns an error message or "".
// seg: optional segment starts (ascending, multiples of 64, first 0); rows are
// never reordered across a segment boundary (batched instances).  Work is split
// over host threads by rows or slices; every output equals the sequential
// element-order construction (blocks' contributor lists in slot order).
std::string build_topology(Topology &t, int N, int F, const int *tri, int H, const int *hinge,
                           const std::vector<int> *seg = nullptr)
{
    t.N = N; t.F = F; t.H = H;
    for (int i = 0; i < 3 * F; ++i)
        if (tri[i] < 0 || tri[i] >= N) return "triangle vertex index out of range";
    for (int i = 0; i < 4 * H; ++i)
        if (hinge[i] < 0 || hinge[i] >= N) return "hinge vertex index out of range";
    if (2 * (6 * (int64_t)F + 10 * (int64_t)H) + 1 > (int64_t)INT32_MAX)
        return "too many element slots for 32-bit indices";
    // vertex force slots: faces (3f+v), then hinges (3F + 4h + v); also the
    // vertex -> element incidence, in slot order
    t.fptr.assign(N + 1, 0);
    for (int i = 0; i < 3 * F; ++i) t.fptr[tri[i] + 1]++;
    for (int i = 0; i < 4 * H; ++i) t.fptr[hinge[i] + 1]++;
    for (int i = 0; i < N; ++i) t.fptr[i + 1] += t.fptr[i];
    t.flist.assign(t.fptr[N], 0);
    {
        std::vector<int> ff(t.fptr.begin(), t.fptr.end() - 1);
        for (int i = 0; i < 3 * F; ++i) t.flist[ff[tri[i]]++] = i;
        for (int i = 0; i < 4 * H; ++i) t.flist[ff[hinge[i]]++] = 3 * F + i;
    }
    // row patterns: the vertex itself and every vertex of its elements, sorted
    std::vector<int> rowlen(N + 1, 0);
    {
        const int T = (int)std::max<int64_t>(1, std::min<int64_t>(16, N / 4096));
        std::vector<std::vector<int>> local(T);
        host_parallel(T, 1, [&](int64_t k0, int64_t k1) {
            for (int64_t k = k0; k < k1; ++k) {
                std::vector<int> &out = local[k];
                std::vector<int> buf;
                for (int64_t i = (int64_t)N * k / T; i < (int64_t)N * (k + 1) / T; ++i) {
                    buf.assign(1, (int)i);
                    for (int s = t.fptr[i]; s < t.fptr[i + 1]; ++s) {
                        const Incidence in = incidence(t.flist[s], F, tri, hinge);
                        buf.insert(buf.end(), in.verts, in.verts + in.nv);
                    }
                    std::sort(buf.begin(), buf.end());
                    const auto e = std::unique(buf.begin(), buf.end());
                    out.insert(out.end(), buf.begin(), e);
                    rowlen[i + 1] = (int)(e - buf.begin());
                }
            }
        });
        int64_t total = 0;
        for (auto &l : local) total += (int64_t)l.size();
        if (total > (int64_t)INT32_MAX) return "too many blocks for 32-bit indices";
        t.rowptr.assign(N + 1, 0);
        for (int i = 0; i < N; ++i) t.rowptr[i + 1] = t.rowptr[i] + rowlen[i + 1];
        t.col.resize((size_t)total);
        host_parallel(T, 1, [&](int64_t k0, int64_t k1) {
            for (int64_t k = k0; k < k1; ++k)
                std::copy(local[k].begin(), local[k].end(), t.col.begin() + t.rowptr[(int64_t)N * k / T]);
        });
    }
    t.nb = (int64_t)t.col.size();
    // SELL-C-sigma layout: rows sorted by decreasing length inside windows of
    // sigma = kSellSigma rows (never across a segment boundary), then cut into
    // slices of C = 64.
    const int C = eolc::kSlice;
    t.nslices = (N + C - 1) / C;
    t.slice_off.assign(t.nslices + 1, 0);
    t.srow.assign(N, 0);
    t.slen.assign(N, 0);
    {
        std::vector<int> bounds = seg ? *seg : std::vector<int>{0};
        bounds.push_back(N);
        std::vector<std::pair<int, int>> win;
        for (size_t si = 0; si + 1 < bounds.size(); ++si)
            for (int w0 = bounds[si]; w0 < bounds[si + 1]; w0 += eolc::kSellSigma)
                win.emplace_back(w0, std::min(bounds[si + 1], w0 + eolc::kSellSigma));
        host_parallel((int64_t)win.size(), 256, [&](int64_t b, int64_t e) {
            std::vector<int> rows;
            for (int64_t w = b; w < e; ++w) {
                const int w0 = win[w].first, w1 = win[w].second;
                rows.clear();
                for (int r = w0; r < w1; ++r) rows.push_back(r);
                std::stable_sort(rows.begin(), rows.end(), [&](int x, int y) {
                    return (t.rowptr[x + 1] - t.rowptr[x]) > (t.rowptr[y + 1] - t.rowptr[y]);
                });
                for (int l = 0; l < (int)rows.size(); ++l) {
                    t.srow[w0 + l] = rows[l];
                    t.slen[w0 + l] = t.rowptr[rows[l] + 1] - t.rowptr[rows[l]];
                }
            }
        });
    }
    for (int sl = 0; sl < t.nslices; ++sl) {
        const int r0 = sl * C, r1 = std::min(N, r0 + C);
        int W = 0;
        for (int g = r0; g < r1; ++g) W = std::max(W, t.slen[g]);
        if ((int64_t)t.slice_off[sl] + (int64_t)W * C > (int64_t)INT32_MAX) return "too many SELL entries";
        t.slice_off[sl + 1] = t.slice_off[sl] + W * C;
    }
    t.nbs = t.slice_off[t.nslices];
    t.scol.assign(t.nbs, 0);
    t.sellpos.assign(t.nb, -1);
    t.sdiag.assign(N, -1);
    std::vector<int> entry_block(t.nbs, -1);
    // Entry order within a row: the upper blocks (column > row) in increasing column
    // order, then the lower blocks (column < row) in increasing column order, then
    // the diagonal block last.  The FP32 difference-form product stops one entry
    // early and never reads the diagonal, and the symmetric FP32 layout below sums
    // in exactly this order.
    t.supper.assign(N, 0);
    std::atomic<bool> bad{false};
    host_parallel(t.nslices, 64, [&](int64_t s0, int64_t s1) {
        for (int sl = (int)s0; sl < (int)s1; ++sl) {
            const int r0 = sl * C, r1 = std::min(N, r0 + C);
            for (int l = 0; l < r1 - r0; ++l) {
                const int row = t.srow[r0 + l];
                int k = 0, bdiag = -1;
                for (int pass = 0; pass < 2; ++pass)
                    for (int b = t.rowptr[row]; b < t.rowptr[row + 1]; ++b) {
                        if (t.col[b] == row) {
                            bdiag = b;
                            continue;
                        }
                        if ((pass == 0) != (t.col[b] > row)) continue;
                        const int e = t.slice_off[sl] + k * C + l;
                        t.scol[e] = t.col[b];
                        t.sellpos[b] = e;
                        entry_block[e] = b;
                        ++k;
                        if (pass == 0) t.supper[r0 + l] = k;
                    }
                if (bdiag < 0 || k != t.slen[r0 + l] - 1) {
                    bad = true;
                    return;
                }
                const int e = t.slice_off[sl] + k * C + l;
                t.scol[e] = row;
                t.sellpos[bdiag] = e;
                entry_block[e] = bdiag;
                t.sdiag[row] = e;
                // padding entries point at the row itself (never read as values)
                for (int k = t.slen[r0 + l]; k < (t.slice_off[sl + 1] - t.slice_off[sl]) / C; ++k)
                    t.scol[t.slice_off[sl] + k * C + l] = row;
            }
        }
    });
    if (bad) return "row without its diagonal block";
    // contributor lists in SELL entry order (padding entries: empty), each in slot
    // order; has_tri marks the blocks of the mass pattern
    auto entry_of = [&](int row, int j) {
        const int *b = t.col.data() + t.rowptr[row], *e = t.col.data() + t.rowptr[row + 1];
        return t.sellpos[std::lower_bound(b, e, j) - t.col.data()];
    };
    t.cptr.assign(t.nbs + 1, 0);
    host_parallel(N, 4096, [&](int64_t r0, int64_t r1) {
        for (int row = (int)r0; row < (int)r1; ++row)
            row_contributions(t, tri, hinge, row, [&](int j, int) { t.cptr[entry_of(row, j) + 1]++; });
    });
    if (!prefix_counts(t.cptr, t.nbs)) return "too many element contributions for 32-bit indices";
    t.clist.assign(t.cptr[t.nbs], 0);
    t.has_tri.assign(t.nb, 0);
    {
        std::vector<int> fill(t.cptr.begin(), t.cptr.end() - 1);
        host_parallel(N, 4096, [&](int64_t r0, int64_t r1) {
            for (int row = (int)r0; row < (int)r1; ++row) {
                row_contributions(t, tri, hinge, row, [&](int j, int code) {
                    const int e = entry_of(row, j);
                    t.clist[fill[e]++] = code;
                });
                for (int b = t.rowptr[row]; b < t.rowptr[row + 1]; ++b) {
                    const int e = t.sellpos[b];
                    int *L = t.clist.data() + t.cptr[e], n = t.cptr[e + 1] - t.cptr[e];
                    for (int a = 1; a < n; ++a)   // insertion sort: lists are short and almost sorted
                        for (int c = a; c > 0 && L[c - 1] > L[c]; --c) std::swap(L[c - 1], L[c]);
                    for (int a = 0; a < n; ++a)
                        if ((L[a] >> 1) < 6 * F) { t.has_tri[b] = 1; break; }
                }
            }
        });
    }
    // Symmetric FP32 layout: each off-diagonal block stored once, with its upper
    // row.  Per slice, upper records (column + 9 values, 40 B) at
    // uoff[s] + 64 k + lane for k < supper, and lower records (column, upper slot of
    // the transposed block) at loff[s] + 64 k + lane for k < slen - 1 - supper.
    t.uoff.assign(t.nslices + 1, 0);
    t.loff.assign(t.nslices + 1, 0);
    for (int sl = 0; sl < t.nslices; ++sl) {
        const int r0 = sl * C, r1 = std::min(N, r0 + C);
        int Wu = 0, Wl = 0;
        for (int g = r0; g < r1; ++g) {
            Wu = std::max(Wu, t.supper[g]);
            Wl = std::max(Wl, t.slen[g] - 1 - t.supper[g]);
        }
        t.uoff[sl + 1] = t.uoff[sl] + Wu * C;
        t.loff[sl + 1] = t.loff[sl] + Wl * C;
    }
    t.nup = t.uoff[t.nslices];
    t.nlo = t.loff[t.nslices];
    // vertex window of every PCG workgroup (16 slices each): its rows and the columns
    // of their entries (a row's padding entries point at the row itself; the lanes of
    // the last slice beyond N hold no row)
    {
        const int G = eolc::pcg_grid(N), spw = eolc::kPcgWG / C;
        t.wwin.assign(2 * (size_t)G, 0);
        for (int lw = 0; lw < G; ++lw) {
            int lo = INT32_MAX, hi = -1;
            for (int sl = lw * spw; sl < std::min(t.nslices, (lw + 1) * spw); ++sl) {
                const int rows = std::min(N - sl * C, C), width = (t.slice_off[sl + 1] - t.slice_off[sl]) / C;
                for (int k = 0; k < width; ++k)
                    for (int l = 0; l < rows; ++l) {
                        const int j = t.scol[t.slice_off[sl] + k * C + l];
                        lo = std::min(lo, j);
                        hi = std::max(hi, j);
                    }
            }
            if (hi < 0) lo = hi = 0;
            else ++hi;
            t.wwin[2 * (size_t)lw] = lo;
            t.wwin[2 * (size_t)lw + 1] = hi;
            t.wmax = std::max(t.wmax, hi - lo);
        }
    }
    t.e2u.assign(t.nbs, -1);
    t.ucol.assign(t.nup, 0);
    t.lrec.assign(2 * (size_t)t.nlo, 0);
    std::vector<int> uposb(t.nb, -1);
    host_parallel(t.nslices, 64, [&](int64_t s0, int64_t s1) {
        for (int sl = (int)s0; sl < (int)s1; ++sl) {
            const int r0 = sl * C, r1 = std::min(N, r0 + C);
            for (int l = 0; l < r1 - r0; ++l) {
                const int row = t.srow[r0 + l];
                const int wu = (t.uoff[sl + 1] - t.uoff[sl]) / C;
                for (int k = 0; k < wu; ++k) {
                    const int u = t.uoff[sl] + k * C + l;
                    if (k < t.supper[r0 + l]) {
                        const int e = t.slice_off[sl] + k * C + l;
                        t.e2u[e] = u;
                        t.ucol[u] = t.scol[e];
                        uposb[entry_block[e]] = u;
                    } else {
                        t.ucol[u] = row;   // padding
                    }
                }
            }
        }
    });
    host_parallel(t.nslices, 64, [&](int64_t s0, int64_t s1) {
        for (int sl = (int)s0; sl < (int)s1; ++sl) {
            const int r0 = sl * C, r1 = std::min(N, r0 + C);
            for (int l = 0; l < r1 - r0; ++l) {
                const int row = t.srow[r0 + l];
                const int nl = t.slen[r0 + l] - 1 - t.supper[r0 + l];
                const int wl = (t.loff[sl + 1] - t.loff[sl]) / C;
                for (int k = 0; k < wl; ++k) {
                    const int lo = t.loff[sl] + k * C + l;
                    if (k < nl) {
                        const int e = t.slice_off[sl] + (t.supper[r0 + l] + k) * C + l;
                        const int j = t.scol[e];
                        const int bt = find_block(t, j, row);
                        if (bt < 0 || uposb[bt] < 0) {
                            bad = true;
                            return;
                        }
                        t.lrec[2 * (size_t)lo] = j;
                        t.lrec[2 * (size_t)lo + 1] = uposb[bt];
                    } else {
                        t.lrec[2 * (size_t)lo] = row;   // padding: never read
                        t.lrec[2 * (size_t)lo + 1] = 0;
                    }
                }
            }
        }
    });
    if (bad) return "block pattern is not symmetric";
    return "";
}

// Patch layout of SpMV layout 5 (Topology::pt_*).  The vertices are cut into
// G = pcg_grid(N) patches by recursive coordinate bisection of the material
// coordinates X: a range of `parts` patches is split at the proportional rank along
// its wider axis (ties by vertex index), so every patch holds floor or ceil of N / G
// vertices (at most kPcgWG) and is compact in the material plane, which keeps most
// stencil pairs inside one patch.  Any mesh gets a valid layout: locality only
// decides how many pairs are cross blocks.  Inside a patch the rows are ordered by
// decreasing upper count (vertex index within equal counts).  Returns "" or why the
// layout does not apply (the patch variants then run their symmetric-record twin).
std::string build_patch_layout(Topology &t, const double *X)
{
    const int N = t.N, C = eolc::kSlice, S = eolc::kPcgWG / C;
    const int G = eolc::pcg_grid(N);
    if ((int64_t)G * eolc::kPcgWG < N) return "more vertices than PCG workgroup rows";
    std::vector<int> ord(N);
    for (int i = 0; i < N; ++i) ord[i] = i;
    std::vector<int> poff(G + 1, 0);
    // explicit stack; left halves are finished first, so patches come out in order
    struct Rng { int lo, hi, parts, first; };
    std::vector<Rng> st{{0, N, G, 0}};
    while (!st.empty()) {
        const Rng r = st.back();
        st.pop_back();
        if (r.parts == 1) {
            poff[r.first] = r.lo;
            poff[r.first + 1] = r.hi;
            continue;
        }
        const int a = r.parts / 2;
        const int cut = r.lo + (int)((int64_t)(r.hi - r.lo) * a / r.parts);
        double lo0 = INFINITY, hi0 = -INFINITY, lo1 = INFINITY, hi1 = -INFINITY;
        for (int k = r.lo; k < r.hi; ++k) {
            const double u = X[2 * (size_t)ord[k]], w = X[2 * (size_t)ord[k] + 1];
            lo0 = std::min(lo0, u); hi0 = std::max(hi0, u);
            lo1 = std::min(lo1, w); hi1 = std::max(hi1, w);
        }
        const int ax = (hi1 - lo1 > hi0 - lo0) ? 1 : 0;
        if (cut > r.lo && cut < r.hi)
            std::nth_element(ord.begin() + r.lo, ord.begin() + cut, ord.begin() + r.hi, [&](int i, int j) {
                const double ui = X[2 * (size_t)i + ax], uj = X[2 * (size_t)j + ax];
                return ui < uj || (ui == uj && i < j);
            });
        st.push_back({cut, r.hi, r.parts - a, r.first + a});
        st.push_back({r.lo, cut, a, r.first});
    }
    std::vector<int> patch(N, -1);
    for (int p = 0; p < G; ++p) {
        if (poff[p + 1] - poff[p] > eolc::kPcgWG) return "a patch holds more than one workgroup's rows";
        for (int k = poff[p]; k < poff[p + 1]; ++k) patch[ord[k]] = p;
    }
    // per-row counts: upper blocks, cross lower blocks, lower blocks
    std::vector<int> nup(N, 0), ncl(N, 0), nlo(N, 0);
    host_parallel(N, 4096, [&](int64_t r0, int64_t r1) {
        for (int i = (int)r0; i < (int)r1; ++i)
            for (int b = t.rowptr[i]; b < t.rowptr[i + 1]; ++b) {
                const int j = t.col[b];
                if (j > i) ++nup[i];
                else if (j < i) {
                    ++nlo[i];
                    if (patch[j] != patch[i]) ++ncl[i];
                }
            }
    });
    // (a row's stiff blocks are flagged in a 64-bit mask)
    for (int i = 0; i < N; ++i)
        if (nup[i] + nlo[i] > 64) return "a row has more than 64 off-diagonal blocks";
    // storage positions: within a patch, decreasing upper count, then vertex index
    t.pt_row.assign((size_t)G * eolc::kPcgWG, -1);
    t.pt_cnt.assign((size_t)G * eolc::kPcgWG, 0);
    std::vector<int> pos(N, -1);
    host_parallel(G, 8, [&](int64_t p0, int64_t p1) {
        std::vector<int> v;
        for (int p = (int)p0; p < (int)p1; ++p) {
            v.assign(ord.begin() + poff[p], ord.begin() + poff[p + 1]);
            std::sort(v.begin(), v.end(), [&](int i, int j) { return nup[i] > nup[j] || (nup[i] == nup[j] && i < j); });
            for (int k = 0; k < (int)v.size(); ++k) {
                const int g = p * eolc::kPcgWG + k;
                t.pt_row[g] = v[k];
                t.pt_cnt[g] = nup[v[k]] | (nlo[v[k]] << 16);
                pos[v[k]] = g;
            }
        }
    });
    // slice widths -> record, lower-entry and LDS-slot offsets
    const int ns = G * S;
    t.pt_roff.assign(ns + 1, 0);
    t.pt_loff.assign(ns + 1, 0);
    t.pt_soff.assign(ns, 0);
    int slots = 0, slots_max = 0;
    for (int sl = 0; sl < ns; ++sl) {
        int wr = 0, wu = 0, wl = 0;
        for (int l = 0; l < C; ++l) {
            const int i = t.pt_row[(size_t)sl * C + l];
            if (i < 0) continue;
            wr = std::max(wr, nup[i] + ncl[i]);
            wu = std::max(wu, nup[i]);
            wl = std::max(wl, nlo[i]);
        }
        if (sl % S == 0) slots = 0;
        t.pt_soff[sl] = slots;
        slots += wu * C;
        slots_max = std::max(slots_max, slots);
        if ((int64_t)t.pt_roff[sl] + (int64_t)wr * C > INT32_MAX || (int64_t)t.pt_loff[sl] + (int64_t)wl * C > INT32_MAX)
            return "too many patch records for 32-bit offsets";
        t.pt_roff[sl + 1] = t.pt_roff[sl] + wr * C;
        t.pt_loff[sl + 1] = t.pt_loff[sl] + wl * C;
    }
    // windows: the patch's rows in storage order (window index = position in the
    // patch), then the vertices of other patches its rows reference, in increasing
    // index
    std::vector<std::vector<int>> halo(G);
    host_parallel(G, 8, [&](int64_t p0, int64_t p1) {
        for (int p = (int)p0; p < (int)p1; ++p) {
            std::vector<int> &h = halo[p];
            for (int k = 0; k < poff[p + 1] - poff[p]; ++k) {
                const int i = t.pt_row[(size_t)p * eolc::kPcgWG + k];
                for (int b = t.rowptr[i]; b < t.rowptr[i + 1]; ++b)
                    if (patch[t.col[b]] != p) h.push_back(t.col[b]);
            }
            std::sort(h.begin(), h.end());
            h.erase(std::unique(h.begin(), h.end()), h.end());
        }
    });
    t.pt_woff.assign(G + 1, 0);
    int wmax = 0;
    for (int p = 0; p < G; ++p) {
        const int w = poff[p + 1] - poff[p] + (int)halo[p].size();
        wmax = std::max(wmax, w);
        t.pt_woff[p + 1] = t.pt_woff[p] + w;
    }
    if (wmax > eolc::kPatchWin)
        return "a patch window holds " + std::to_string(wmax) + " vertices (limit " +
               std::to_string(eolc::kPatchWin) + ")";
    if (slots_max + wmax > eolc::kPatchPool)
        return "a patch needs " + std::to_string(slots_max) + " LDS slots and " + std::to_string(wmax) +
               " window vertices (limit " + std::to_string(eolc::kPatchPool) + " together)";
    t.pt_win.assign(std::max(t.pt_woff[G], 1), 0);
    for (int p = 0; p < G; ++p) {
        const int n = poff[p + 1] - poff[p];
        for (int k = 0; k < n; ++k) t.pt_win[t.pt_woff[p] + k] = t.pt_row[(size_t)p * eolc::kPcgWG + k];
        std::copy(halo[p].begin(), halo[p].end(), t.pt_win.begin() + t.pt_woff[p] + n);
    }
    t.pt_wmax = wmax;
    auto local = [&](int p, int j) {   // window index of vertex j in patch p's window
        if (patch[j] == p) return pos[j] - p * eolc::kPcgWG;
        return poff[p + 1] - poff[p] + (int)(std::lower_bound(halo[p].begin(), halo[p].end(), j) - halo[p].begin());
    };
    t.pt_slots = slots_max;
    t.pt_nrec = t.pt_roff[ns];
    t.pt_nlo = t.pt_loff[ns];
    // records, and the LDS slot each in-patch upper block leaves for its lower row
    // (indexed by the CSR block of that lower entry)
    t.pt_e2r.assign(t.nbs, -1);
    t.pt_e2l.assign(t.nbs, 0);
    t.pt_lo.assign(std::max<int64_t>(t.pt_nlo, 1), eolc::kPatchCross);
    std::vector<int> tslot(t.nb, -1);
    std::atomic<bool> bad{false};
    host_parallel(N, 4096, [&](int64_t r0, int64_t r1) {
        for (int i = (int)r0; i < (int)r1; ++i) {
            const int g = pos[i], sl = g / C, l = g % C;
            int k = 0;
            for (int b = t.rowptr[i]; b < t.rowptr[i + 1]; ++b) {
                const int j = t.col[b];
                if (j <= i) continue;
                t.pt_e2r[t.sellpos[b]] = t.pt_roff[sl] + k * C + l;
                t.pt_e2l[t.sellpos[b]] = local(patch[i], j);
                if (patch[j] == patch[i]) {
                    const int bt = find_block(t, j, i);
                    if (bt < 0) {
                        bad = true;
                        return;
                    }
                    tslot[bt] = t.pt_soff[sl] + k * C + l;
                }
                ++k;
            }
            int c = 0;
            for (int b = t.rowptr[i]; b < t.rowptr[i + 1]; ++b) {
                const int j = t.col[b];
                if (j < i && patch[j] != patch[i]) {
                    t.pt_e2r[t.sellpos[b]] = t.pt_roff[sl] + (nup[i] + c++) * C + l;
                    t.pt_e2l[t.sellpos[b]] = local(patch[i], j);
                }
            }
        }
    });
    if (bad) return "block pattern is not symmetric";
    host_parallel(N, 4096, [&](int64_t r0, int64_t r1) {
        for (int i = (int)r0; i < (int)r1; ++i) {
            const int g = pos[i], sl = g / C, l = g % C;
            int k = 0;
            for (int b = t.rowptr[i]; b < t.rowptr[i + 1]; ++b) {
                const int j = t.col[b];
                if (j >= i) continue;
                const size_t at = (size_t)t.pt_loff[sl] + (size_t)k * C + l;
                if (patch[j] == patch[i]) {
                    if (tslot[b] < 0 || tslot[b] >= (int)eolc::kPatchCross) {
                        bad = true;
                        return;
                    }
                    t.pt_lo[at] = (unsigned short)tslot[b];
                } else {
                    t.pt_lo[at] = eolc::kPatchCross;
                }
                ++k;
            }
        }
    });
    if (bad) return "patch lower entry without its upper block";
    return "";
}

template <class T> hipError_t upload(eolc_ctx *c, T **dst, const std::vector<T> &src)
{
    hipError_t e = dalloc(c, dst, src.size());
    if (e != hipSuccess) return e;
    if (src.empty()) return hipSuccess;
    return hipMemcpy(*dst, src.data(), src.size() * sizeof(T), hipMemcpyHostToDevice);
}

double elapsed_ms(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.f;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return (double)ms;
}

// Per-topology buffers of the learned preconditioner, (re)allocated when the
// topology changed since the last use.  tpos maps every SELL entry (i, j) to the
// SELL entry of (j, i) (the pattern is symmetric); padding maps to -1.
int learned_buffers(eolc_ctx *c)
{
    eolc::LearnedState &L = c->lp;
    const Topology &t = c->topo;
    if (L.N == t.N && L.nbs == t.nbs && L.tpos) return EOLC_OK;
    const size_t N = (size_t)t.N;
    std::vector<int> tpos((size_t)t.nbs, -1);
    for (int i = 0; i < t.N; ++i)
        for (int k = t.rowptr[i]; k < t.rowptr[i + 1]; ++k) {
            const int k2 = find_block(t, t.col[k], i);
            if (k2 < 0) return fail(c, EOLC_ERR_INVALID, "learned: block pattern is not symmetric");
            tpos[(size_t)t.sellpos[k]] = t.sellpos[k2];
        }
    EOLC_HIP(upload(c, &L.tpos, tpos));
    EOLC_HIP(dalloc(c, &L.w, (size_t)t.nbs));
    EOLC_HIP(dalloc(c, &L.wT, (size_t)t.nbs));
    EOLC_HIP(dalloc(c, &L.X, N * eolc::kLpFeat));
    for (auto &p : L.h) EOLC_HIP(dalloc(c, &p, N * eolc::kLpHidden));
    for (auto &p : L.a) EOLC_HIP(dalloc(c, &p, N * eolc::kLpHidden));
    for (auto &p : L.dpre) EOLC_HIP(dalloc(c, &p, N * eolc::kLpHidden));
    EOLC_HIP(dalloc(c, &L.o, N * eolc::kLpOut));
    EOLC_HIP(dalloc(c, &L.dO, N * eolc::kLpOut));
    EOLC_HIP(dalloc(c, &L.dinp, N * 2 * eolc::kLpHidden));
    const size_t nblk = (N + 255) / 256;
    int Gw = 0, rows = 0;
    eolc::learned_wgrad_shape(t.N, &Gw, &rows);
    L.gpart_n = std::max(nblk * (eolc::kLpHidden * eolc::kLpOut + eolc::kLpOut),
                         (size_t)Gw * eolc::kLpLayerSize);
    EOLC_HIP(dalloc(c, &L.gpart, L.gpart_n));
    EOLC_HIP(dalloc(c, &L.dpart, nblk));
    EOLC_HIP(dalloc(c, &L.scal, 8));
    EOLC_HIP(hipMemset(L.scal, 0, 8 * sizeof(double)));
    for (double **p : {&L.e, &L.r, &L.z, &L.Az}) EOLC_HIP(dalloc(c, p, 3 * N));
    L.N = t.N;
    L.nbs = t.nbs;
    return EOLC_OK;
}

// Geometry and set-up chunks of the linear coarse fields (levels >= kMLLinFrom): the
// centroid, RMS radius and inverse radius of every aggregate in the material
// coordinates, and per aggregate of those levels its hord range cut into chunks of
// at most kGal9Rows positions.  A hierarchy whose aggregates are not contiguous in
// hord keeps the translations alone (lin_ok false).
hipError_t ml_prepare_linear(eolc_ctx *c, const std::vector<int> &hord)
{
    eolc::MLState &m = c->ml;
    const eolc::MLLevels &lv = m.lv;
    const int N = c->topo.N, L = lv.L;
    std::vector<double> X((size_t)2 * N);
    hipError_t e = hipMemcpy(X.data(), c->X, sizeof(double) * 2 * N, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return e;
    const size_t na = (size_t)lv.off[L + 1];
    std::vector<double> geo(4 * na, 0.0), cnt(na, 0.0);
    for (int l = 1; l <= L; ++l)
        for (int i = 0; i < N; ++i) {
            const size_t g = (size_t)lv.off[l] + m.maps[l][i];
            geo[4 * g] += X[2 * (size_t)i];
            geo[4 * g + 1] += X[2 * (size_t)i + 1];
            cnt[g] += 1.0;
        }
    for (size_t g = 0; g < na; ++g) {
        geo[4 * g] /= cnt[g];
        geo[4 * g + 1] /= cnt[g];
    }
    for (int l = 1; l <= L; ++l)
        for (int i = 0; i < N; ++i) {
            const size_t g = (size_t)lv.off[l] + m.maps[l][i];
            const double du = X[2 * (size_t)i] - geo[4 * g], dv = X[2 * (size_t)i + 1] - geo[4 * g + 1];
            geo[4 * g + 2] += du * du + dv * dv;
        }
    for (size_t g = 0; g < na; ++g) {
        double s = std::sqrt(geo[4 * g + 2] / cnt[g]);
        if (!(s > 0.0)) s = 1.0;
        geo[4 * g + 2] = s;
        geo[4 * g + 3] = 1.0 / s;
    }
    // hord ranges of the aggregates of levels >= kMLLinFrom, and their chunks
    const int l0 = eolc::kMLLinFrom;
    const int n_hi = lv.off[L + 1] - lv.off[l0];
    std::vector<int> k0(n_hi, INT_MAX), k1(n_hi, -1), seen(n_hi, 0);
    bool contiguous = true;
    for (int l = l0; l <= L && contiguous; ++l) {
        int prev = -1;
        for (int k = 0; k < N; ++k) {
            const int A = m.maps[l][hord[k]];
            const int h = lv.off[l] - lv.off[l0] + A;
            if (A != prev) {
                if (seen[h]) { contiguous = false; break; }
                seen[h] = 1;
                k0[h] = k;
                prev = A;
            }
            k1[h] = k + 1;
        }
    }
    std::vector<int> chunk, gcp(n_hi + 1, 0);
    if (contiguous) {
        for (int l = l0; l <= L; ++l)
            for (int A = 0; A < lv.n[l]; ++A) {
                const int h = lv.off[l] - lv.off[l0] + A;
                gcp[h] = (int)chunk.size() / 4;
                for (int k = k0[h]; k < k1[h]; k += eolc::kGal9Rows) {
                    chunk.push_back(l);
                    chunk.push_back(lv.off[l] + A);
                    chunk.push_back(k);
                    chunk.push_back(std::min(k + eolc::kGal9Rows, k1[h]));
                }
            }
        gcp[n_hi] = (int)chunk.size() / 4;
    }
    if ((e = upload(c, &m.geo, geo)) != hipSuccess) return e;
    // the coarse kernel keeps one aggregate of these levels per thread (about 1.4 M
    // vertices on a grid); larger hierarchies keep the translations
    m.lin_ok = contiguous && n_hi > 0 && n_hi <= eolc::kCoarse9WG;
    if (!m.lin_ok) return hipSuccess;
    m.nchunks = (int)chunk.size() / 4;
    if ((e = upload(c, &m.gchunk, chunk)) != hipSuccess) return e;
    if ((e = upload(c, &m.gcp, gcp)) != hipSuccess) return e;
    if ((e = dalloc(c, &m.gpart, (size_t)45 * m.nchunks)) != hipSuccess) return e;
    if ((e = dalloc(c, &m.D9, (size_t)54 * n_hi)) != hipSuccess) return e;
    // per-vertex features of levels 2 (in hord order, for the restriction) and 3 (for
    // the prolongation), formed as the kernels would form them from X and geo
    std::vector<double> f2((size_t)2 * N), f3((size_t)2 * N);
    for (int k = 0; k < N; ++k) {
        const int i = hord[k];
        const size_t g = (size_t)lv.off[2] + m.maps[2][i];
        f2[2 * (size_t)k] = (X[2 * (size_t)i] - geo[4 * g]) * geo[4 * g + 3];
        f2[2 * (size_t)k + 1] = (X[2 * (size_t)i + 1] - geo[4 * g + 1]) * geo[4 * g + 3];
    }
    for (int i = 0; i < N; ++i) {
        const size_t g = (size_t)lv.off[3] + m.maps[3][i];
        f3[2 * (size_t)i] = (X[2 * (size_t)i] - geo[4 * g]) * geo[4 * g + 3];
        f3[2 * (size_t)i + 1] = (X[2 * (size_t)i + 1] - geo[4 * g + 1]) * geo[4 * g + 3];
    }
    if ((e = upload(c, &m.feat2, f2)) != hipSuccess) return e;
    if ((e = upload(c, &m.feat3, f3)) != hipSuccess) return e;
    if ((e = dalloc(c, &m.R9p, (size_t)9 * eolc::kRestrict3Parts * lv.n[3])) != hipSuccess) return e;
    return dalloc(c, &m.C9, (size_t)9 * n_hi);
}

// Aggregation hierarchy of the multilevel preconditioner for the current topology
// (built once, on first use).
int ml_prepare(eolc_ctx *c)
{
    eolc::MLState &m = c->ml;
    if (m.built) return EOLC_OK;
    const Topology &t = c->topo;
    std::vector<int> hord, ptr1, agg1, cptr, par;
    std::vector<unsigned char> lam;
    eolc::MLLevels lv{};
    const std::string msg = eolc::ml_build(t, m.maps, hord, ptr1, agg1, cptr, par, lam, lv);
    if (!msg.empty()) return fail(c, EOLC_ERR_INVALID, msg);
    m.lv = lv;
    m.counts.assign(lv.n, lv.n + lv.L + 1);
    if (lv.L > 0) {
        std::vector<int> gpos((size_t)t.N);
        for (int g = 0; g < t.N; ++g) gpos[t.srow[g]] = g;
        EOLC_HIP(upload(c, &m.hord, hord));
        EOLC_HIP(upload(c, &m.ptr1, ptr1));
        EOLC_HIP(upload(c, &m.agg1, agg1));
        EOLC_HIP(upload(c, &m.gpos, gpos));
        EOLC_HIP(upload(c, &m.cptr, cptr));
        EOLC_HIP(upload(c, &m.par, par));
        EOLC_HIP(upload(c, &m.lam, lam));
        const size_t na = (size_t)lv.off[lv.L + 1];
        EOLC_HIP(dalloc(c, &m.V, (size_t)9 * lv.L * na));
        EOLC_HIP(dalloc(c, &m.Dinv, 6 * na));
        EOLC_HIP(dalloc(c, &m.R, 3 * na));
        EOLC_HIP(dalloc(c, &m.C, 3 * na));
        if (lv.L >= 2) {
            std::vector<int> kptr2((size_t)lv.n[2] + 1);
            for (int w = 0; w <= lv.n[2]; ++w) kptr2[w] = ptr1[cptr[(size_t)lv.coff[2] + w]];
            EOLC_HIP(upload(c, &m.kptr2, kptr2));
        }
        if (lv.L >= 3) {
            std::vector<int> anc3((size_t)t.N);
            for (int i = 0; i < t.N; ++i) anc3[i] = lv.off[3] + par[(size_t)lv.off[2] + par[agg1[i]]];
            EOLC_HIP(upload(c, &m.anc3, anc3));
            m.hord_h = hord;   // for the linear coarse fields, built on first use
        }
    }
    m.built = true;
    return EOLC_OK;
}

// Whether the multilevel levels >= 3 carry linear fields for iterations in this
// precision (eolc_set_ml_linear).
static bool ml_lin_for(const eolc_ctx *c, bool mixed) { return c->ml_linear == 2 || (c->ml_linear == 1 && mixed); }

// The linear coarse fields' geometry, features and set-up chunks, built on the first
// call that uses them for this hierarchy.
static int ml_lin_prepare(eolc_ctx *c, bool want)
{
    eolc::MLState &m = c->ml;
    if (!want || m.lin_built || !m.built || m.lv.L < eolc::kMLLinFrom) return EOLC_OK;
    EOLC_HIP(ml_prepare_linear(c, m.hord_h));
    m.lin_built = true;
    return EOLC_OK;
}

// The multilevel flag is on for exactly one call.
struct MlScope {
    eolc_ctx *c;
    MlScope(eolc_ctx *ctx, bool on) : c(ctx) { c->ml_on = on && c->ml.lv.L > 0; }
    ~MlScope() { c->ml_on = false; }
};

bool valid_mode(int mode)
{
    const int base = mode & ~EOLC_SOLVE_MULTILEVEL;
    return base >= EOLC_SOLVE_FP64_BJ && base <= EOLC_SOLVE_FP64_JACOBI;
}

bool mixed_mode(int mode)
{
    const int base = mode & ~EOLC_SOLVE_MULTILEVEL;
    return base == EOLC_SOLVE_MIXED_BJ || base == EOLC_SOLVE_MIXED_LEARNED;
}

// Fixed DOFs -> device mask (and values) for the duration of one call.
// Marks one call's fixed DOFs in fixmask / vfix from the context's own device lists
// (capacity 3N, allocated with the topology: no allocation per call) and clears
// them again on every return path.  vals = nullptr: zero values; scripted: vals are
// scripted velocities and the fixed values become (1 - 0.01) v_old + vals, formed
// on the device.  sync = false leaves the host-to-device copies in flight: the
// caller keeps dofs and vals alive until its next stream synchronisation.
struct FixedScope {
    eolc_ctx *c = nullptr;
    int n = 0;
    int set(eolc_ctx *ctx, int nf, const int *dofs, const double *vals, bool scripted = false, bool sync = true)
    {
        const int N = ctx->topo.N;
        if (nf > 3 * N) return fail(ctx, EOLC_ERR_INVALID, "more fixed DOFs than the mesh has DOFs");
        for (int i = 0; i < nf; ++i)
            if (dofs[i] < 0 || dofs[i] >= 3 * N || (vals && !std::isfinite(vals[i])))
                return fail(ctx, EOLC_ERR_INVALID, "fixed DOF index out of range or value non-finite");
        if (nf <= 0) return EOLC_OK;
        c = ctx;
        n = nf;   // from here on the destructor clears the mask
        EOLC_HIP(hipMemcpyAsync(c->fix_dofs, dofs, sizeof(int) * nf, hipMemcpyHostToDevice, c->stream));
        if (vals)
            EOLC_HIP(hipMemcpyAsync(c->fix_vals, vals, sizeof(double) * nf, hipMemcpyHostToDevice, c->stream));
        else
            EOLC_HIP(hipMemsetAsync(c->fix_vals, 0, sizeof(double) * nf, c->stream));
        if (scripted) EOLC_HIP(eolc::fixed_from_vold(c, nf, c->fix_dofs, c->fix_vals));
        EOLC_HIP(eolc::set_fixed(c, nf, c->fix_dofs, c->fix_vals, 1));
        if (sync) EOLC_HIP(hipStreamSynchronize(c->stream));
        return EOLC_OK;
    }
    // stream-ordered: later work on the context's stream sees the cleared mask; a
    // batched context gets its own fixed nodes (corners and padding) back
    ~FixedScope()
    {
        if (c && n > 0) {
            (void)eolc::set_fixed(c, n, c->fix_dofs, c->fix_vals, 0);
            if (c->batch.B > 0) (void)eolc::batch_set_fixmask(c, c->batch.d_all_fixed, c->batch.n_all_fixed);
        }
    }
};

// hipEvents destroyed on every return path
struct EventSet {
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    ~EventSet()
    {
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    }
};

void drop_batch_graph(eolc_ctx *c)
{
    if (c->batch.graph) (void)hipGraphExecDestroy(c->batch.graph);
    c->batch.graph = nullptr;
}

}  // namespace

extern "C" {

int eolc_create(int device, eolc_ctx **out)
try {
    if (!out) return EOLC_ERR_INVALID;
    *out = nullptr;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return EOLC_ERR_NO_DEVICE;
    if (device < 0 || device >= ndev) return EOLC_ERR_NO_DEVICE;
    eolc_ctx *c = new (std::nothrow) eolc_ctx();
    if (!c) return EOLC_ERR_INVALID;
    c->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return EOLC_ERR_HIP;
    }
    if (hipMalloc((void **)&c->scal, sizeof(eolc::SolverScalars)) != hipSuccess ||
        hipMalloc((void **)&c->partials, sizeof(double) * 6 * eolc::kMaxPartials) != hipSuccess ||
        hipMalloc((void **)&c->counter, sizeof(unsigned) * 4) != hipSuccess ||
        hipHostMalloc((void **)&c->hscal, sizeof(eolc::SolverScalars), hipHostMallocDefault) != hipSuccess ||
        hipMemset(c->counter, 0, sizeof(unsigned) * 4) != hipSuccess ||
        hipMemset(c->scal, 0, sizeof(eolc::SolverScalars)) != hipSuccess ||
        hipMemset(c->partials, 0, sizeof(double) * 6 * eolc::kMaxPartials) != hipSuccess) {
        eolc_destroy(c);
        return EOLC_ERR_HIP;
    }
    // the no-allocation query's scratch array: its address on this device, looked up
    // once per device with the device current
    void *scratch_ptr = nullptr;
    if (eolc::collide_scratch(device, &scratch_ptr) != hipSuccess) {
        eolc_destroy(c);
        return EOLC_ERR_HIP;
    }
    *out = c;
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(nullptr, "create");
}

void eolc_destroy(eolc_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    free_topology_buffers(c);
    free_learned_weights(c);
    dfree(c->scal); dfree(c->partials); dfree(c->counter);
    if (c->hscal) (void)hipHostFree(c->hscal);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *eolc_last_error(const eolc_ctx *c) { return c ? c->err.c_str() : "null context"; }

static int set_topology_impl(eolc_ctx *c, int N, const double *X, int F, const int *tri, int H, const int *hinge,
                             const std::vector<int> *seg);

int eolc_set_topology(eolc_ctx *c, int N, const double *X, int F, const int *tri, int H, const int *hinge)
try {
    return set_topology_impl(c, N, X, F, tri, H, hinge, nullptr);
} catch (...) {
    return eolc::report_exception(c, "set_topology");
}

static int set_topology_impl(eolc_ctx *c, int N, const double *X, int F, const int *tri, int H, const int *hinge,
                             const std::vector<int> *seg)
{
    if (!c) return EOLC_ERR_INVALID;
    if (N <= 0 || F < 0 || H < 0 || !X || (F && !tri) || (H && !hinge))
        return fail(c, EOLC_ERR_INVALID, "set_topology: bad sizes or null arrays");
    EOLC_HIP(hipSetDevice(c->device));
    EOLC_HIP(hipStreamSynchronize(c->stream));
    for (int f = 0; f < F; ++f) {
        const int a = tri[3 * f], b = tri[3 * f + 1], cc = tri[3 * f + 2];
        if (a < 0 || a >= N || b < 0 || b >= N || cc < 0 || cc >= N)
            return fail(c, EOLC_ERR_INVALID, "set_topology: triangle index out of range");
        double A2 = (X[2 * b] - X[2 * a]) * (X[2 * cc + 1] - X[2 * a + 1]) -
                    (X[2 * b + 1] - X[2 * a + 1]) * (X[2 * cc] - X[2 * a]);
        if (!(A2 > 0.0))
            return fail(c, EOLC_ERR_DEGENERATE,
                        "set_topology: face " + std::to_string(f) + " has non-positive material area (must be CCW)");
    }
    if (eolc::pcg_grid(N) > eolc::kMaxPartials)
        return fail(c, EOLC_ERR_INVALID, "set_topology: mesh too large for the PCG reduction buffers");
    Topology t;
    std::string msg = build_topology(t, N, F, tri, H, hinge, seg);
    if (!msg.empty()) return fail(c, EOLC_ERR_INVALID, "set_topology: " + msg);
    t.pt_why = build_patch_layout(t, X);
    t.pt_ok = t.pt_why.empty();
    if (!t.pt_ok) {   // the patch variants fall back to their symmetric-record twin
        t.pt_row.clear(); t.pt_cnt.clear(); t.pt_roff.clear(); t.pt_soff.clear(); t.pt_loff.clear();
        t.pt_lo.clear(); t.pt_e2r.clear(); t.pt_e2l.clear(); t.pt_woff.clear(); t.pt_win.clear();
        t.pt_nrec = t.pt_nlo = 0;
    }
    free_topology_buffers(c);
    c->have_topo = c->have_state = c->have_forces = false;
    c->topo = std::move(t);
    // an LDS-staged SpMV variant chosen for an earlier mesh: its unstaged twin when
    // this mesh's windows are too wide
    for (int &v : c->spmv_variant)
        if (!eolc::spmv_usable(c->topo, v)) v = eolc::spmv_fallback(v);
    const Topology &T = c->topo;
    std::vector<double> Xv(X, X + 2 * (size_t)N);
    std::vector<int> triv(tri, tri + 3 * (size_t)F), hv(hinge, hinge + 4 * (size_t)H);
    EOLC_HIP(upload(c, &c->X, Xv));
    EOLC_HIP(upload(c, &c->tri, triv));
    EOLC_HIP(upload(c, &c->hinge, hv));
    EOLC_HIP(upload(c, &c->slice_off, T.slice_off));
    EOLC_HIP(upload(c, &c->srow, T.srow));
    EOLC_HIP(upload(c, &c->slen, T.slen));
    EOLC_HIP(upload(c, &c->scol, T.scol));
    EOLC_HIP(upload(c, &c->diagpos, T.sdiag));
    EOLC_HIP(upload(c, &c->supper, T.supper));
    EOLC_HIP(upload(c, &c->uoff, T.uoff));
    EOLC_HIP(upload(c, &c->loff, T.loff));
    EOLC_HIP(upload(c, &c->lrec, T.lrec));
    EOLC_HIP(upload(c, &c->e2u, T.e2u));
    EOLC_HIP(upload(c, &c->wwin, T.wwin));
    if (T.pt_ok) {
        EOLC_HIP(upload(c, &c->pt_row, T.pt_row));
        EOLC_HIP(upload(c, &c->pt_cnt, T.pt_cnt));
        EOLC_HIP(upload(c, &c->pt_roff, T.pt_roff));
        EOLC_HIP(upload(c, &c->pt_soff, T.pt_soff));
        EOLC_HIP(upload(c, &c->pt_loff, T.pt_loff));
        EOLC_HIP(upload(c, &c->pt_lo, T.pt_lo));
        EOLC_HIP(upload(c, &c->pt_e2r, T.pt_e2r));
        EOLC_HIP(upload(c, &c->pt_e2l, T.pt_e2l));
        EOLC_HIP(upload(c, &c->pt_woff, T.pt_woff));
        EOLC_HIP(upload(c, &c->pt_win, T.pt_win));
        EOLC_HIP(dalloc(c, &c->P32, 10 * (size_t)std::max<int64_t>(T.pt_nrec, 1)));
        EOLC_HIP(dalloc(c, &c->stiff, (size_t)T.nbs));
        EOLC_HIP(dalloc(c, &c->A64s, (size_t)9 * T.nbs));   // written for the stiff entries only
        EOLC_HIP(dalloc(c, &c->smask, (size_t)N));
        std::vector<int> sb(N);
        for (int g = 0; g < N; ++g) sb[T.srow[g]] = T.slice_off[g / eolc::kSlice] + g % eolc::kSlice;
        EOLC_HIP(upload(c, &c->sellbase, sb));
        EOLC_HIP(dalloc(c, &c->P64, 10 * (size_t)std::max<int64_t>(T.pt_nrec, 1)));
    }
    EOLC_HIP(upload(c, &c->cptr, T.cptr));
    EOLC_HIP(upload(c, &c->clist, T.clist));
    EOLC_HIP(upload(c, &c->fptr, T.fptr));
    EOLC_HIP(upload(c, &c->flist, T.flist));
    const size_t nslots = 6 * (size_t)F + 10 * (size_t)H;
    const size_t n3 = 3 * (size_t)N;
    EOLC_HIP(dalloc(c, &c->stageK, 9 * nslots));
    EOLC_HIP(dalloc(c, &c->stageF, 3 * (3 * (size_t)F + 4 * (size_t)H)));
    EOLC_HIP(dalloc(c, &c->stageM, 6 * (size_t)F));
    EOLC_HIP(dalloc(c, &c->x, n3));
    EOLC_HIP(dalloc(c, &c->v, n3));
    EOLC_HIP(dalloc(c, &c->f, n3));
    EOLC_HIP(dalloc(c, &c->rhs, n3));
    EOLC_HIP(dalloc(c, &c->A64, 9 * (size_t)T.nbs));
    EOLC_HIP(dalloc(c, &c->S32, 9 * (size_t)T.nbs));
    EOLC_HIP(dalloc(c, &c->M32, (size_t)N));
    EOLC_HIP(dalloc(c, &c->S32u, 10 * (size_t)std::max<int64_t>(T.nup, 1)));
    EOLC_HIP(dalloc(c, &c->S32a, 10 * (size_t)T.nbs));
    EOLC_HIP(dalloc(c, &c->S64u, 10 * (size_t)std::max<int64_t>(T.nup, 1)));
    EOLC_HIP(dalloc(c, &c->M64, (size_t)N));
    {
        std::vector<int> isd((size_t)T.nbs, 0);
        for (int i = 0; i < N; ++i) isd[T.sdiag[i]] = 1;
        EOLC_HIP(upload(c, &c->entry_is_diag, isd));
    }
    EOLC_HIP(dalloc(c, &c->Mval, (size_t)T.nbs));
    EOLC_HIP(dalloc(c, &c->D64, eolc::kDsym * (size_t)N));
    EOLC_HIP(dalloc(c, &c->D32, eolc::kDsym * (size_t)N));
    EOLC_HIP(dalloc(c, &c->fixmask, ((size_t)N + 3) & ~(size_t)3));
    EOLC_HIP(dalloc(c, &c->vfix, n3));
    EOLC_HIP(dalloc(c, &c->fix_dofs, n3));
    EOLC_HIP(dalloc(c, &c->fix_vals, n3));
    EOLC_HIP(hipMemset(c->fixmask, 0, ((size_t)N + 3) & ~(size_t)3));
    EOLC_HIP(hipMemset(c->vfix, 0, n3 * sizeof(double)));
    for (double **p : {&c->b64, &c->xs64, &c->r64, &c->z64, &c->q64, &c->p64[0], &c->p64[1], &c->xl64})
        EOLC_HIP(dalloc(c, p, n3));
    for (float **p : {&c->r32, &c->z32, &c->q32, &c->xl32, &c->p32[0], &c->p32[1]}) EOLC_HIP(dalloc(c, p, n3));
    // proximity-query tables (createEdges order, jitter stream)
    c->topo_tri.assign(tri, tri + 3 * (size_t)F);
    c->topo_hinge.assign(hinge, hinge + 4 * (size_t)H);
    eolc::build_collision_edges(c, F, tri);
    EOLC_HIP(upload(c, &c->cedges, c->h_cedges));
    std::vector<double> unit;
    eolc::build_jitter(unit, N);
    EOLC_HIP(upload(c, &c->cd_unit, unit));
    c->have_topo = true;
    return EOLC_OK;
}

int eolc_set_material(eolc_ctx *c, double density, double e, double nu, double beta, double dampingA,
                      double dampingB)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (!(std::isfinite(density) && std::isfinite(e) && std::isfinite(nu) && std::isfinite(beta) &&
          std::isfinite(dampingB)) || nu >= 0.5 || nu <= -1.0)
        return fail(c, EOLC_ERR_INVALID, "set_material: non-finite value or Poisson ratio outside (-1, 0.5)");
    c->rho = density; c->E = e; c->nu = nu; c->beta = beta; c->dampA = dampingA; c->dampB = dampingB;
    c->have_mat = true;
    c->have_forces = false;
    drop_batch_graph(c);   // the captured step holds the old scalars
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "set_material");
}

int eolc_set_options(eolc_ctx *c, int flags)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (flags & ~(EOLC_OPT_INCLUDE_BENDING_FORCE | EOLC_OPT_SPD_PROJECT_BENDING))
        return fail(c, EOLC_ERR_INVALID, "set_options: unknown flag");
    c->flags = flags;
    c->have_forces = false;
    drop_batch_graph(c);   // the captured step chose its kernels by the old flags
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "set_options");
}

int eolc_upload_state(eolc_ctx *c, const double *x, const double *v)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (!c->have_topo) return fail(c, EOLC_ERR_INVALID, "upload_state: no topology");
    if (!x || !v) return fail(c, EOLC_ERR_INVALID, "upload_state: null array");
    EOLC_HIP(hipSetDevice(c->device));
    const size_t n = 3 * (size_t)c->topo.N;
    for (size_t i = 0; i < n; ++i)
        if (!std::isfinite(x[i]) || !std::isfinite(v[i]))
            return fail(c, EOLC_ERR_NONFINITE, "upload_state: non-finite entry at DOF " + std::to_string(i));
    EOLC_HIP(hipMemcpyAsync(c->x, x, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    EOLC_HIP(hipMemcpyAsync(c->v, v, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    EOLC_HIP(hipStreamSynchronize(c->stream));
    c->have_state = true;
    c->have_forces = false;
    c->fp64_sticky = 0;   // a new state: mixed solves try FP32 again
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "upload_state");
}

int eolc_download_state(eolc_ctx *c, double *x, double *v)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (!c->have_state) return fail(c, EOLC_ERR_INVALID, "download_state: no state");
    EOLC_HIP(hipSetDevice(c->device));
    const size_t n = 3 * (size_t)c->topo.N;
    if (x) EOLC_HIP(hipMemcpyAsync(x, c->x, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (v) EOLC_HIP(hipMemcpyAsync(v, c->v, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    EOLC_HIP(hipStreamSynchronize(c->stream));
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "download_state");
}

static int forces_fill_async(eolc_ctx *c, const double grav[3], double h, bool want32)
{
    EOLC_HIP(hipMemsetAsync(c->counter + 1, 0, 2 * sizeof(unsigned), c->stream));
    EOLC_HIP(eolc::launch_face_eval(c, h, grav));
    EOLC_HIP(eolc::launch_hinge_eval(c, h));
    c->h_last = h;   // (the assembly's stiff threshold scales with h^2)
    EOLC_HIP(eolc::launch_assemble(c, (want32 ? eolc::fp32_layout(c->spmv_variant[0]) : 0) |
                                          eolc::fp64_layout(c->spmv_variant[1])));
    c->have_forces = true;
    return EOLC_OK;
}

int eolc_forces_fill(eolc_ctx *c, const double grav[3], double h)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (!c->have_topo || !c->have_mat || !c->have_state)
        return fail(c, EOLC_ERR_INVALID, "forces_fill: topology, material and state must be set first");
    if (!grav || !std::isfinite(h) || h <= 0.0) return fail(c, EOLC_ERR_INVALID, "forces_fill: bad gravity or h");
    EOLC_HIP(hipSetDevice(c->device));
    int rc = forces_fill_async(c, grav, h, true);
    if (rc) return rc;
    EOLC_HIP(eolc::launch_rhs(c, h));
    EOLC_HIP(eolc::eol_fill(c));
    EOLC_HIP(hipStreamSynchronize(c->stream));
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "forces_fill");
}

int eolc_set_eol(eolc_ctx *c, int n, const int *nodes)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (!c->have_topo) return fail(c, EOLC_ERR_INVALID, "set_eol: no topology");
    if (c->batch.B > 0) return fail(c, EOLC_ERR_INVALID, "set_eol: not for a batch topology");
    if (n < 0 || (n && !nodes)) return fail(c, EOLC_ERR_INVALID, "set_eol: bad arguments");
    const Topology &T = c->topo;
    const int N = T.N, F = T.F, H = T.H;
    std::vector<int> eidx((size_t)N, -1);
    for (int k = 0; k < n; ++k) {
        if (nodes[k] < 0 || nodes[k] >= N || eidx[nodes[k]] >= 0)
            return fail(c, EOLC_ERR_INVALID, "set_eol: node out of range or repeated");
        eidx[nodes[k]] = k;
    }
    if (n == 0) {
        EOLC_HIP(hipSetDevice(c->device));
        EOLC_HIP(hipStreamSynchronize(c->stream));
        free_eol(c);
        c->have_forces = false;
        return EOLC_OK;
    }
    // host maps first (the current EOL state stays as it is if this fails)
    eolc::EolState S;
    S.E = n;
    S.nodes.assign(nodes, nodes + n);
    const int *tri = c->topo_tri.data(), *hinge = c->topo_hinge.data();
    // EOL elements in the reference's order (faces, then hinges) and their staged
    // entries: (X row, column, stage index), forces (X row, stage index)
    struct Trip { int row, col, sidx; char face; };
    std::vector<Trip> trips;
    std::vector<std::pair<int, int>> ftrips;
    std::vector<int> elem;
    std::vector<int64_t> soff{0}, foff{0};
    std::vector<std::pair<int, int>> xslots;   // (vertex, hinge force slot) of EOL hinges
    for (int id = 0; id < F + H; ++id) {
        const bool face = id < F;
        const int nv = face ? 3 : 4;
        const int *vt = face ? tri + 3 * (size_t)id : hinge + 4 * (size_t)(id - F);
        int loc[4], nE = 0;
        for (int v = 0; v < nv; ++v)
            if (eidx[vt[v]] >= 0) loc[nE++] = v;
        if (nE == 0) continue;
        const int64_t base = soff.back(), fbase = foff.back();
        const int rec = 4 * nE + 6 * nv;
        if (base + (int64_t)nE * rec > (int64_t)INT32_MAX) return fail(c, EOLC_ERR_INVALID, "set_eol: too many EOL entries");
        for (int ev = 0; ev < nE; ++ev) {
            const int row0 = 2 * eidx[vt[loc[ev]]];
            const int64_t r0 = base + (int64_t)ev * rec;
            for (int ew = 0; ew < nE; ++ew)
                for (int j = 0; j < 2; ++j)
                    for (int k = 0; k < 2; ++k)
                        trips.push_back({row0 + j, 3 * N + 2 * eidx[vt[loc[ew]]] + k, (int)(r0 + 4 * ew + 2 * j + k),
                                         (char)face});
            for (int u = 0; u < nv; ++u)
                for (int j = 0; j < 2; ++j)
                    for (int k = 0; k < 3; ++k)
                        trips.push_back({row0 + j, 3 * vt[u] + k, (int)(r0 + 4 * nE + 6 * u + 3 * j + k), (char)face});
            for (int j = 0; j < 2; ++j) ftrips.push_back({row0 + j, (int)(fbase + 2 * ev + j)});
        }
        if (!face)
            for (int v = 0; v < 4; ++v) xslots.push_back({vt[v], 3 * F + 4 * (id - F) + v});
        elem.push_back(id);
        soff.push_back(base + (int64_t)nE * rec);
        foff.push_back(fbase + 2 * nE);
    }
    S.ne = (int)elem.size();
    std::stable_sort(trips.begin(), trips.end(),
                     [](const Trip &a, const Trip &b) { return a.row != b.row ? a.row < b.row : a.col < b.col; });
    S.rowptr.assign(2 * (size_t)n + 1, 0);
    std::vector<int64_t> cptr{0};
    std::vector<int> clist;
    clist.reserve(trips.size());
    for (size_t k = 0; k < trips.size(); ++k) {
        const bool first = k == 0 || trips[k].row != trips[k - 1].row || trips[k].col != trips[k - 1].col;
        if (first) {
            if (k) cptr.push_back((int64_t)clist.size());
            S.col.push_back(trips[k].col);
            S.has_face.push_back(0);
            S.rowptr[trips[k].row + 1]++;
        }
        clist.push_back(trips[k].sidx);
        if (trips[k].face) S.has_face.back() = 1;
    }
    if (!trips.empty()) cptr.push_back((int64_t)clist.size());
    for (int r = 0; r < 2 * n; ++r) S.rowptr[r + 1] += S.rowptr[r];
    std::stable_sort(ftrips.begin(), ftrips.end(),
                     [](const std::pair<int, int> &a, const std::pair<int, int> &b) { return a.first < b.first; });
    std::vector<int> fptr(2 * (size_t)n + 1, 0), flist;
    for (auto &t : ftrips) {
        fptr[t.first + 1]++;
        flist.push_back(t.second);
    }
    for (int r = 0; r < 2 * n; ++r) fptr[r + 1] += fptr[r];
    std::stable_sort(xslots.begin(), xslots.end(),
                     [](const std::pair<int, int> &a, const std::pair<int, int> &b) { return a.first < b.first; });
    std::vector<int> xv, xvptr{0}, xvslot;
    for (size_t k = 0; k < xslots.size(); ++k) {
        if (k == 0 || xslots[k].first != xslots[k - 1].first) {
            if (k) xvptr.push_back((int)xvslot.size());
            xv.push_back(xslots[k].first);
        }
        xvslot.push_back(xslots[k].second);
    }
    if (!xslots.empty()) xvptr.push_back((int)xvslot.size());
    S.nxv = (int)xv.size();
    const size_t nst = (size_t)soff.back(), nsf = (size_t)foff.back(), nnz = S.col.size();
    // then the device: the old state is released, the new one uploaded; a failure
    // leaves no EOL nodes set
    EOLC_HIP(hipSetDevice(c->device));
    EOLC_HIP(hipStreamSynchronize(c->stream));
    free_eol(c);
    c->have_forces = false;
    c->eol = std::move(S);
    eolc::EolState &D = c->eol;
    hipError_t e = hipSuccess;
    auto step = [&](hipError_t r) { if (e == hipSuccess) e = r; };
    step(upload(c, &D.d_elem, elem));
    step(upload(c, &D.d_eidx, eidx));
    step(upload(c, &D.d_soff, soff));
    step(upload(c, &D.d_foff, foff));
    step(upload(c, &D.d_cptr, cptr));
    step(upload(c, &D.d_clist, clist));
    step(upload(c, &D.d_fptr, fptr));
    step(upload(c, &D.d_flist, flist));
    step(upload(c, &D.d_xv, xv));
    step(upload(c, &D.d_xvptr, xvptr));
    step(upload(c, &D.d_xvslot, xvslot));
    step(dalloc(c, &D.d_stK, nst));
    step(dalloc(c, &D.d_stM, nst));
    step(dalloc(c, &D.d_stF, nsf));
    step(dalloc(c, &D.d_A, nnz));
    step(dalloc(c, &D.d_M, nnz));
    step(dalloc(c, &D.d_fX, 2 * (size_t)n));
    if (e != hipSuccess) {
        free_eol(c);
        return fail(c, EOLC_ERR_HIP, std::string("set_eol: HIP error ") + hipGetErrorString(e));
    }
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "set_eol");
}

int eolc_download_forces(eolc_ctx *c, double *f)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (!c->have_forces || !f) return fail(c, EOLC_ERR_INVALID, "download_forces: no forces");
    EOLC_HIP(hipSetDevice(c->device));
    EOLC_HIP(hipMemcpyAsync(f, c->f, 3 * (size_t)c->topo.N * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (c->eol.E > 0)
        EOLC_HIP(hipMemcpyAsync(f + 3 * (size_t)c->topo.N, c->eol.d_fX, 2 * (size_t)c->eol.E * sizeof(double),
                                hipMemcpyDeviceToHost, c->stream));
    EOLC_HIP(hipStreamSynchronize(c->stream));
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "download_forces");
}

int eolc_download_csr(eolc_ctx *c, int which, int64_t *rowptr, int *col, double *val, int64_t cap, int64_t *nnz)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (!c->have_forces) return fail(c, EOLC_ERR_INVALID, "download_csr: call forces_fill first");
    if (which != EOLC_MAT_M && which != EOLC_MAT_MDK) return fail(c, EOLC_ERR_INVALID, "download_csr: bad matrix id");
    if (!nnz) return fail(c, EOLC_ERR_INVALID, "download_csr: null nnz");
    EOLC_HIP(hipSetDevice(c->device));
    const Topology &T = c->topo;
    const int N = T.N;
    // blocks present: MDK = all; M = blocks with at least one triangle contribution
    const eolc::EolState &S = c->eol;
    const bool wantM = which == EOLC_MAT_M;
    int64_t need = 0;
    for (int64_t b = 0; b < T.nb; ++b) need += (!wantM || T.has_tri[b]) ? 9 : 0;
    // EOL rows, and their transposes in the x rows (the reference pushes both)
    std::vector<std::vector<std::pair<int, int64_t>>> mirror;
    if (S.E > 0) {
        mirror.resize(3 * (size_t)N);
        for (int r = 0; r < 2 * S.E; ++r)
            for (int64_t k = S.rowptr[r]; k < S.rowptr[r + 1]; ++k) {
                if (wantM && !S.has_face[k]) continue;
                ++need;
                if (S.col[k] < 3 * N) {
                    mirror[S.col[k]].push_back({3 * N + r, k});
                    ++need;
                }
            }
    }
    *nnz = need;
    if (!rowptr || !col || !val || cap < need) return fail(c, EOLC_ERR_OVERFLOW, "download_csr: capacity too small");
    std::vector<double> A((size_t)9 * T.nbs), m((size_t)T.nbs), xa;
    EOLC_HIP(hipMemcpyAsync(A.data(), c->A64, A.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    EOLC_HIP(hipMemcpyAsync(m.data(), c->Mval, m.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (S.E > 0) {
        xa.resize(S.col.size());
        if (!xa.empty())
            EOLC_HIP(hipMemcpyAsync(xa.data(), wantM ? S.d_M : S.d_A, xa.size() * sizeof(double),
                                    hipMemcpyDeviceToHost, c->stream));
    }
    EOLC_HIP(hipStreamSynchronize(c->stream));
    int64_t p = 0;
    for (int i = 0; i < N; ++i)
        for (int a = 0; a < 3; ++a) {
            rowptr[3 * i + a] = p;
            for (int k = T.rowptr[i]; k < T.rowptr[i + 1]; ++k) {
                if (wantM && !T.has_tri[k]) continue;
                const int j = T.col[k];
                const size_t e = (size_t)T.sellpos[k];
                for (int cc = 0; cc < 3; ++cc) {
                    col[p] = 3 * j + cc;
                    val[p] = !wantM ? A[(size_t)(3 * a + cc) * T.nbs + e] : (a == cc ? m[e] : 0.0);
                    ++p;
                }
            }
            if (S.E > 0)
                for (const auto &q : mirror[3 * i + a]) {
                    col[p] = q.first;
                    val[p] = xa[q.second];
                    ++p;
                }
        }
    for (int r = 0; r < 2 * S.E; ++r) {
        rowptr[3 * N + r] = p;
        for (int64_t k = S.rowptr[r]; k < S.rowptr[r + 1]; ++k) {
            if (wantM && !S.has_face[k]) continue;
            col[p] = S.col[k];
            val[p] = xa[k];
            ++p;
        }
    }
    rowptr[3 * N + 2 * S.E] = p;
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "download_csr");
}

// ---------------------------------------------------------------------------
// scripted != 0: fixed_val holds scripted velocities and the fixed values are
// (1 - 0.01) v_old + fixed_val, formed on the device (no host read of v_old).
// A solve ends with EOLC_OK only when the FP64 true residual of its answer meets
// the tolerance: when the recursive residual of either precision meets it, the
// true residual replaces it and the iteration goes on from there if it does not.
// keep_v_on_fail: a failed solve leaves the resident velocity as it was (the step
// driver); otherwise v receives the last iterate whatever the status.
static int solve_impl(eolc_ctx *c, int n_fixed, const int *fixed_dof, const double *fixed_val, double tol,
                      int maxit, int mode, eolc_solve_stats *st, int scripted = 0, bool keep_v_on_fail = false)
{
    if (c->eol.E > 0)
        return fail(c, EOLC_ERR_INVALID,
                    "velocity_solve: EOL nodes are set; the EOL system goes to the host QP with its EOL rows");
    const bool multilevel = (mode & EOLC_SOLVE_MULTILEVEL) != 0;
    mode &= ~EOLC_SOLVE_MULTILEVEL;
    if (multilevel) {
        const int rc = ml_prepare(c);
        if (rc) return rc;
    }
    MlScope mls(c, multilevel);
    int &bj_hint = multilevel ? c->bj_iters_hint_ml : c->bj_iters_hint;
    bool mixed = (mode == EOLC_SOLVE_MIXED_BJ || mode == EOLC_SOLVE_MIXED_LEARNED);
    int used_fallback = 0;
    // After an FP32 stall (below) the next kStickyFp64 mixed solves on this context
    // run in FP64 from the start, then FP32 is tried again.
    if (mixed && c->fp64_sticky > 0) {
        --c->fp64_sticky;
        mixed = false;
        mode = EOLC_SOLVE_FP64_BJ;
        used_fallback = 2;
    }
    if (mode == EOLC_SOLVE_MIXED_LEARNED) {
        if (!c->lp.have_weights) return fail(c, EOLC_ERR_INVALID, "velocity_solve: learned mode without weights");
        const int rc = learned_buffers(c);
        if (rc) return rc;
    }
    // fixed DOFs -> device mask and values (the host arrays outlive the first sync below)
    FixedScope fx;
    {
        const int rc = fx.set(c, n_fixed, fixed_dof, fixed_val, scripted != 0, false);
        if (rc) return rc;
    }
    const int precond = (mode == EOLC_SOLVE_FP64_JACOBI) ? 1 : 0;
    EOLC_HIP(eolc::ensure_fp64_layout(c, c->spmv_variant[1]));   // FP64 iterations, or a stall's continuation
    if (mixed) EOLC_HIP(eolc::ensure_fp32_layout(c, c->spmv_variant[0]));
    if (multilevel) {
        const int rc = ml_lin_prepare(c, ml_lin_for(c, mixed));
        if (rc) return rc;
    }
    EOLC_HIP(eolc::pcg_setup(c, mixed, precond, ml_lin_for(c, mixed)));
    // Learned mode: P replaces the FP32 block-Jacobi blocks; after twice the
    // iterations of the last converged block-Jacobi solve on this context (maxit
    // when there was none) the solve falls back to block-Jacobi from its current
    // iterate.
    bool learned = (mode == EOLC_SOLVE_MIXED_LEARNED);
    int cap = maxit;
    if (learned) {
        EOLC_HIP(eolc::learned_precond(c));
        // a zero output layer gives exactly the block-Jacobi blocks: no cap, no fallback
        if (bj_hint > 0 && !c->lp.zero_output) cap = std::min(maxit, 2 * bj_hint);
    }
    const double delta = mixed ? (c->ml_on ? c->mixed_delta_ml : c->mixed_delta) : 0.0;
    EOLC_HIP(eolc::scal_set(c, tol * tol, delta * delta, 1));
    EOLC_HIP(eolc::pcg_init(c, mixed));
    int k = 0, refinements = 0, status = EOLC_ERR_MAXIT, launched = 0, syncs = 0;
    // Iterations are enqueued in chunks; the host synchronises once per chunk.  A
    // chunk that overruns the stop point costs one no-op launch pair per extra
    // iteration, so the next chunk is sized at 60% of the iterations that the mean
    // convergence rate since the start of the current inner cycle predicts to the
    // current threshold (8 .. 128).  The CG residual is not monotone and drops in
    // bursts, so the rate is measured over the whole cycle and the chunk is capped.
    int chunk = 16;
    int k_cycle = 0;
    double rr_cycle = -1.0;
    double relres = NAN;
    // FP32 stall detection: an inner cycle far longer than the completed ones, or the
    // true residual hardly falling over consecutive reliable updates, means the FP32
    // operator no longer resolves the system (extreme stiffness ratios after the sheet
    // deforms); the solve then continues in FP64 from its current iterate (from its
    // initial guess when the FP32 iterations used linear coarse fields), with the
    // translations alone on the multilevel levels (eolc_set_ml_linear).
    double last_true2 = -1.0;
    int slow_updates = 0;
    auto fp64_continue = [&]() -> int {
        // with the translations alone the FP64 iteration continues from the stalled
        // iterate (x64 += the FP32 correction; r = b - A x64); after FP32 iterations
        // with linear coarse fields it restarts from the initial guess
        if (c->ml_on && c->ml.lin_on)
            EOLC_HIP(eolc::pcg_restart_from_guess(c));
        else
            EOLC_HIP(eolc::pcg_reliable_update(c));
        mixed = false;
        learned = false;
        used_fallback |= 2;
        c->fp64_sticky = c->stall_sticky;
        const size_t n3 = 3 * (size_t)c->topo.N;
        EOLC_HIP(hipMemsetAsync(c->p64[0], 0, n3 * sizeof(double), c->stream));
        EOLC_HIP(hipMemsetAsync(c->p64[1], 0, n3 * sizeof(double), c->stream));
        EOLC_HIP(eolc::pcg_setup(c, false, 0, ml_lin_for(c, false)));
        EOLC_HIP(eolc::scal_delta(c, 0.0));
        EOLC_HIP(eolc::pcg_restart_inner(c, false, k));
        EOLC_HIP(eolc::scal_restart(c, k));
        rr_cycle = -1.0;
        k_cycle = k;
        chunk = 16;
        return EOLC_OK;
    };
    for (;;) {
        if (learned && k >= cap) {
            // fallback: block-Jacobi blocks, true residual of the current iterate, beta = 0
            learned = false;
            used_fallback = 1;
            EOLC_HIP(eolc::pcg_setup(c, mixed, 0, ml_lin_for(c, mixed)));
            EOLC_HIP(eolc::pcg_reliable_update(c));
            EOLC_HIP(eolc::pcg_restart_inner(c, mixed, k));
            EOLC_HIP(eolc::scal_restart(c, k));
            rr_cycle = -1.0;
            k_cycle = k;
        }
        const int todo = std::min(chunk, (learned ? cap : maxit) - k);
        if (todo <= 0) { status = EOLC_ERR_MAXIT; break; }
        EOLC_HIP(eolc::pcg_chunk(c, mixed, k, todo));
        launched += todo;
        EOLC_HIP(hipMemcpyAsync(c->hscal, c->scal, sizeof(eolc::SolverScalars), hipMemcpyDeviceToHost, c->stream));
        EOLC_HIP(hipStreamSynchronize(c->stream));
        ++syncs;
        const eolc::SolverScalars s = *c->hscal;
        k = s.iters;
        relres = s.bnorm2 > 0 ? std::sqrt(s.rr / s.bnorm2) : 0.0;
        if (rr_cycle < 0.0) rr_cycle = s.rmax2;   // the (true) residual the cycle started from
        {
            const double thr = std::max(s.tol2 * s.bnorm2, s.delta2 * s.rmax2);
            int next = 32;
            if (rr_cycle > 0.0 && k > k_cycle && s.rr > thr && s.rr < rr_cycle) {
                const double rate = std::log(s.rr / rr_cycle) / (k - k_cycle);   // mean log reduction per iteration
                next = (int)std::min(128.0, std::max(8.0, 0.6 * std::log(thr / s.rr) / rate));
            }
            chunk = next;
        }
        if (!s.done) {
            const int stall = refinements > 0 ? std::max(c->stall_min, 10 * (k_cycle / refinements)) : c->stall_first;
            if (mixed && k - k_cycle > stall) {
                const int rc = fp64_continue();
                if (rc) return rc;
            }
            continue;
        }
        rr_cycle = -1.0;   // the residual is replaced below: a new cycle starts at k
        k_cycle = k;
        chunk = 16;
        if (s.reason == 2 && learned) {
            // breakdown under the learned blocks (P is SPD, so only from an
            // indefinite MDK): the block-Jacobi fallback reports it
            cap = k;
            continue;
        }
        if (s.reason == 2) { status = EOLC_ERR_NOT_SPD; break; }
        if (s.reason == 4) { status = EOLC_ERR_NONFINITE; break; }
        // Replace the recursive residual by the true FP64 residual of the current
        // iterate: mixed, after folding the FP32 correction in (reliable update);
        // FP64, whose recursion drifts from the true residual over long solves.
        if (mixed) EOLC_HIP(eolc::pcg_reliable_update(c));
        else EOLC_HIP(eolc::pcg_replace_residual64(c));
        EOLC_HIP(hipMemcpyAsync(c->hscal, c->scal, sizeof(eolc::SolverScalars), hipMemcpyDeviceToHost, c->stream));
        EOLC_HIP(hipStreamSynchronize(c->stream));
        ++syncs;
        ++refinements;
        const eolc::SolverScalars s2 = *c->hscal;
        relres = s2.bnorm2 > 0 ? std::sqrt(s2.rmax2 / s2.bnorm2) : 0.0;
        if (!std::isfinite(s2.rmax2)) { status = EOLC_ERR_NONFINITE; break; }
        if (s2.rmax2 <= tol * tol * s2.bnorm2) { status = EOLC_OK; break; }
        if (k >= maxit || refinements > eolc::kMaxRefinements) { status = EOLC_ERR_MAXIT; break; }
        if (mixed) {
            // a cycle aims at a true-residual drop of delta; less than half that, thrice
            const double slow = std::min(0.64, std::max(0.16, 4.0 * delta * delta));
            slow_updates = (last_true2 > 0.0 && s2.rmax2 > slow * last_true2) ? slow_updates + 1 : 0;
            last_true2 = s2.rmax2;
            if (slow_updates >= 3) {
                const int rc = fp64_continue();
                if (rc) return rc;
                continue;
            }
        }
        if (learned && k >= cap) continue;   // the fallback restarts from here
        EOLC_HIP(eolc::pcg_restart_inner(c, mixed, k));
        EOLC_HIP(eolc::scal_set(c, 0, 0, 0));
    }
    if (status == EOLC_OK && !learned && used_fallback == 0 &&
        (mode == EOLC_SOLVE_FP64_BJ || mode == EOLC_SOLVE_MIXED_BJ))
        bj_hint = k;
    if (mixed && status != EOLC_OK && status != EOLC_ERR_NOT_SPD) {
        // fold the last FP32 correction in before reporting
        EOLC_HIP(eolc::pcg_reliable_update(c));
    }
    double *out = (status == EOLC_OK || !keep_v_on_fail) ? c->v : c->xs64;
    EOLC_HIP(eolc::pcg_finish(c, out));
    EOLC_HIP(eolc::spmv_true_residual(c, out));
    EOLC_HIP(hipMemcpyAsync(c->hscal, c->scal, sizeof(eolc::SolverScalars), hipMemcpyDeviceToHost, c->stream));
    EOLC_HIP(hipStreamSynchronize(c->stream));
    if (st) {
        st->iters = c->hscal->iters;
        st->launched = launched;
        st->host_syncs = syncs;
        st->refinements = refinements;
        st->status = status;
        st->used_fallback = used_fallback;
        st->relres = relres;
        st->true_relres = c->hscal->bnorm2 > 0 ? std::sqrt(c->hscal->true_rr / c->hscal->bnorm2) : 0.0;
    }
    if (status == EOLC_ERR_NOT_SPD) c->err = "velocity_solve: p^T A p <= 0 (MDK is not positive definite)";
    else if (status == EOLC_ERR_MAXIT) c->err = "velocity_solve: iteration cap reached";
    else if (status == EOLC_ERR_NONFINITE) c->err = "velocity_solve: non-finite value in the solve";
    return status;
}

int eolc_velocity_solve(eolc_ctx *c, int n_fixed, const int *fixed_dof, const double *fixed_val, double tol,
                        int maxit, int mode, eolc_solve_stats *st)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (!c->have_forces) return fail(c, EOLC_ERR_INVALID, "velocity_solve: call forces_fill first");
    if (n_fixed < 0 || (n_fixed && (!fixed_dof || !fixed_val)) || !(tol > 0.0) || maxit <= 0)
        return fail(c, EOLC_ERR_INVALID, "velocity_solve: bad arguments");
    if (!valid_mode(mode)) return fail(c, EOLC_ERR_INVALID, "velocity_solve: unknown mode");
    EOLC_HIP(hipSetDevice(c->device));
    if (st) std::memset(st, 0, sizeof(*st));
    return solve_impl(c, n_fixed, fixed_dof, fixed_val, tol, maxit, mode, st);
} catch (...) {
    return eolc::report_exception(c, "velocity_solve");
}

static int transfer_impl(eolc_ctx *c, int n_old, const double *u_old, const double *v_old, int f_old,
                         const int *tri_old, int n_new, const double *u_new, double *v_new, int *src, int *face,
                         double *bary, double *ms, const double *x_old, const int *eol_old, int n_eol,
                         const double *V_old);

int eolc_velocity_transfer(eolc_ctx *c, int n_old, const double *u_old, const double *v_old, int f_old,
                           const int *tri_old, int n_new, const double *u_new, double *v_new, int *src, int *face,
                           double *bary, double *ms)
try {
    return transfer_impl(c, n_old, u_old, v_old, f_old, tri_old, n_new, u_new, v_new, src, face, bary, ms, nullptr,
                         nullptr, 0, nullptr);
} catch (...) {
    return eolc::report_exception(c, "velocity_transfer");
}

int eolc_velocity_transfer_eol(eolc_ctx *c, int n_old, const double *u_old, const double *x_old, const double *v_old,
                               const int *eol_old, int n_eol, const double *V_old, int f_old, const int *tri_old,
                               int n_new, const double *u_new, double *v_new, int *src, int *face, double *bary,
                               double *ms)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (n_old <= 0 || n_eol < 0 || !x_old || !eol_old || (n_eol && !V_old))
        return fail(c, EOLC_ERR_INVALID, "velocity_transfer_eol: bad arguments");
    for (size_t k = 0; k < 3 * (size_t)n_old; ++k)
        if (!std::isfinite(x_old[k])) return fail(c, EOLC_ERR_NONFINITE, "velocity_transfer_eol: non-finite old x");
    for (size_t k = 0; k < 2 * (size_t)n_eol; ++k)
        if (!std::isfinite(V_old[k])) return fail(c, EOLC_ERR_NONFINITE, "velocity_transfer_eol: non-finite old V");
    for (int j = 0; j < n_old; ++j)
        if (eol_old[j] < -1 || eol_old[j] >= n_eol)
            return fail(c, EOLC_ERR_INVALID, "velocity_transfer_eol: EoL index out of range");
    return transfer_impl(c, n_old, u_old, v_old, f_old, tri_old, n_new, u_new, v_new, src, face, bary, ms, x_old,
                         eol_old, n_eol, V_old);
} catch (...) {
    return eolc::report_exception(c, "velocity_transfer_eol");
}

static int transfer_impl(eolc_ctx *c, int n_old, const double *u_old, const double *v_old, int f_old,
                         const int *tri_old, int n_new, const double *u_new, double *v_new, int *src, int *face,
                         double *bary, double *ms, const double *x_old, const int *eol_old, int n_eol,
                         const double *V_old)
{
    if (!c) return EOLC_ERR_INVALID;
    if (n_old <= 0 || f_old < 0 || n_new < 0 || !u_old || !v_old || (f_old && !tri_old) ||
        (n_new && (!u_new || !v_new || !src || !face)))
        return fail(c, EOLC_ERR_INVALID, "velocity_transfer: bad arguments");
    for (size_t k = 0; k < 2 * (size_t)n_old; ++k)
        if (!std::isfinite(u_old[k])) return fail(c, EOLC_ERR_NONFINITE, "velocity_transfer: non-finite old position");
    for (size_t k = 0; k < 3 * (size_t)n_old; ++k)
        if (!std::isfinite(v_old[k])) return fail(c, EOLC_ERR_NONFINITE, "velocity_transfer: non-finite old velocity");
    for (size_t k = 0; k < 2 * (size_t)n_new; ++k)
        if (!std::isfinite(u_new[k])) return fail(c, EOLC_ERR_NONFINITE, "velocity_transfer: non-finite new position");
    for (size_t k = 0; k < 3 * (size_t)f_old; ++k)
        if (tri_old[k] < 0 || tri_old[k] >= n_old) return fail(c, EOLC_ERR_INVALID, "velocity_transfer: triangle index out of range");
    if (ms) *ms = 0.0;
    if (n_new == 0) return EOLC_OK;
    EOLC_HIP(hipSetDevice(c->device));
    int missing = 0;
    float t = 0.f;
    EOLC_HIP(eolc::velocity_transfer(c, n_old, u_old, v_old, f_old, tri_old, n_new, u_new, v_new, src, face, bary,
                                     &missing, &t, x_old, eol_old, n_eol, V_old));
    if (ms) *ms = t;
    if (missing)
        return fail(c, EOLC_ERR_INVALID, "velocity_transfer: " + std::to_string(missing) +
                                             " new vertices have neither an old vertex within 1e-6 nor an enclosing old face");
    return EOLC_OK;
}

int eolc_ml_info(eolc_ctx *c, int *levels, int *counts)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (!c->have_topo) return fail(c, EOLC_ERR_INVALID, "ml_info: no topology");
    if (!levels) return fail(c, EOLC_ERR_INVALID, "ml_info: null output");
    EOLC_HIP(hipSetDevice(c->device));
    const int rc = ml_prepare(c);
    if (rc) return rc;
    *levels = c->ml.lv.L;
    if (counts)
        for (int l = 0; l <= c->ml.lv.L; ++l) counts[l] = c->ml.counts[l];
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "ml_info");
}

int eolc_ml_aggregates(eolc_ctx *c, int level, int *agg)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (!c->have_topo) return fail(c, EOLC_ERR_INVALID, "ml_aggregates: no topology");
    EOLC_HIP(hipSetDevice(c->device));
    const int rc = ml_prepare(c);
    if (rc) return rc;
    if (!agg || level < 1 || level > c->ml.lv.L) return fail(c, EOLC_ERR_INVALID, "ml_aggregates: bad level or null output");
    std::copy(c->ml.maps[level].begin(), c->ml.maps[level].end(), agg);
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "ml_aggregates");
}

int eolc_precond_apply(eolc_ctx *c, int n_fixed, const int *fixed_dof, int mode, const double *r, double *z)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (!c->have_forces) return fail(c, EOLC_ERR_INVALID, "precond_apply: call forces_fill first");
    if (!r || !z || n_fixed < 0 || (n_fixed && !fixed_dof)) return fail(c, EOLC_ERR_INVALID, "precond_apply: bad arguments");
    if (!valid_mode(mode)) return fail(c, EOLC_ERR_INVALID, "precond_apply: unknown mode");
    EOLC_HIP(hipSetDevice(c->device));
    const int base = mode & ~EOLC_SOLVE_MULTILEVEL;
    const bool multilevel = (mode & EOLC_SOLVE_MULTILEVEL) != 0;
    const bool mixed = mixed_mode(mode);
    const size_t n3 = 3 * (size_t)c->topo.N;
    for (size_t k = 0; k < n3; ++k)
        if (!std::isfinite(r[k])) return fail(c, EOLC_ERR_NONFINITE, "precond_apply: non-finite r");
    int rc = EOLC_OK;
    if (base == EOLC_SOLVE_MIXED_LEARNED) {
        if (!c->lp.have_weights) return fail(c, EOLC_ERR_INVALID, "precond_apply: learned mode without weights");
        if ((rc = learned_buffers(c))) return rc;
    }
    if (multilevel && (rc = ml_prepare(c))) return rc;
    FixedScope fs;
    if ((rc = fs.set(c, n_fixed, fixed_dof, nullptr))) return rc;
    MlScope mls(c, multilevel);
    std::vector<double> rm(r, r + n3);
    for (int i = 0; i < n_fixed; ++i) rm[(size_t)fixed_dof[i]] = 0.0;
    if (multilevel && (rc = ml_lin_prepare(c, ml_lin_for(c, mixed)))) return rc;
    EOLC_HIP(eolc::pcg_setup(c, mixed, base == EOLC_SOLVE_FP64_JACOBI ? 1 : 0, ml_lin_for(c, mixed)));
    if (base == EOLC_SOLVE_MIXED_LEARNED) EOLC_HIP(eolc::learned_precond(c));
    EOLC_HIP(hipMemcpyAsync(c->r64, rm.data(), n3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    EOLC_HIP(eolc::pcg_precondition(c, mixed));
    if (mixed) {
        std::vector<float> zf(n3);
        EOLC_HIP(hipMemcpyAsync(zf.data(), c->z32, n3 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        EOLC_HIP(hipStreamSynchronize(c->stream));
        for (size_t k = 0; k < n3; ++k) z[k] = zf[k];
    } else {
        EOLC_HIP(hipMemcpyAsync(z, c->z64, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        EOLC_HIP(hipStreamSynchronize(c->stream));
    }
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "precond_apply");
}

// One step with single-DOF fixed rows: v[dof] = (1 - 0.01) v_old[dof] + scripted.
static int step_impl(eolc_ctx *c, const double grav[3], double h, int n_dofs, const int *dofs,
                     const double *scripted, double tol, int maxit, int mode, eolc_solve_stats *st)
{
    if (!c->have_topo || !c->have_mat || !c->have_state)
        return fail(c, EOLC_ERR_INVALID, "step: topology, material and state must be set first");
    if (!grav || !(h > 0.0) || n_dofs < 0 || (n_dofs && !dofs) || !(tol > 0.0) || maxit <= 0)
        return fail(c, EOLC_ERR_INVALID, "step: bad arguments");
    if (!valid_mode(mode)) return fail(c, EOLC_ERR_INVALID, "step: unknown mode");
    if (c->eol.E > 0)
        return fail(c, EOLC_ERR_INVALID, "step: EOL nodes are set; the EOL system goes to the host QP with its EOL rows");
    EOLC_HIP(hipSetDevice(c->device));
    const int N = c->topo.N;
    for (int i = 0; i < n_dofs; ++i)
        if (dofs[i] < 0 || dofs[i] >= 3 * N) return fail(c, EOLC_ERR_INVALID, "step: fixed node or DOF out of range");
    EventSet es;
    hipEvent_t *ev = es.ev;
    for (int i = 0; i < 5; ++i) EOLC_HIP(hipEventCreate(&ev[i]));
    EOLC_HIP(hipEventRecord(ev[0], c->stream));
    const bool mixed = mixed_mode(mode);
    int rc = forces_fill_async(c, grav, h, mixed);
    if (rc) return rc;
    EOLC_HIP(hipEventRecord(ev[1], c->stream));
    EOLC_HIP(eolc::launch_rhs(c, h));
    EOLC_HIP(hipEventRecord(ev[2], c->stream));
    // v_fixed = (1 - 0.01) v_old + scripted (Constraints.cpp:114-119), formed on the
    // device from the resident v_old
    std::vector<double> vals((size_t)n_dofs, 0.0);
    if (scripted)
        for (int i = 0; i < n_dofs; ++i) vals[i] = scripted[i];
    eolc_solve_stats local;
    std::memset(&local, 0, sizeof(local));
    rc = solve_impl(c, n_dofs, dofs, vals.data(), tol, maxit, mode, &local, 1, true);
    EOLC_HIP(hipEventRecord(ev[3], c->stream));
    if (rc == EOLC_OK) EOLC_HIP(eolc::launch_update_x(c, h));
    EOLC_HIP(hipEventRecord(ev[4], c->stream));
    EOLC_HIP(hipStreamSynchronize(c->stream));
    local.ms_assembly = elapsed_ms(ev[0], ev[1]);
    local.ms_rhs = elapsed_ms(ev[1], ev[2]);
    local.ms_solve = elapsed_ms(ev[2], ev[3]);
    local.ms_update = elapsed_ms(ev[3], ev[4]);
    if (st) *st = local;
    return rc;
}

int eolc_step(eolc_ctx *c, const double grav[3], double h, int n_fixed_nodes, const int *fixed_nodes,
              const double *scripted_vel, double tol, int maxit, int mode, eolc_solve_stats *st)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (n_fixed_nodes < 0 || (n_fixed_nodes && !fixed_nodes)) return fail(c, EOLC_ERR_INVALID, "step: bad arguments");
    std::vector<int> dofs(3 * (size_t)n_fixed_nodes);
    for (int i = 0; i < n_fixed_nodes; ++i) {
        if (fixed_nodes[i] < 0 || fixed_nodes[i] > INT_MAX / 3 - 1)
            return fail(c, EOLC_ERR_INVALID, "step: fixed node out of range");
        for (int k = 0; k < 3; ++k) dofs[3 * i + k] = 3 * fixed_nodes[i] + k;
    }
    return step_impl(c, grav, h, 3 * n_fixed_nodes, dofs.data(), scripted_vel, tol, maxit, mode, st);
} catch (...) {
    return eolc::report_exception(c, "step");
}

int eolc_step_dofs(eolc_ctx *c, const double grav[3], double h, int n_fixed_dofs, const int *fixed_dofs,
                   const double *scripted_vel, double tol, int maxit, int mode, eolc_solve_stats *st)
try {
    if (!c) return EOLC_ERR_INVALID;
    return step_impl(c, grav, h, n_fixed_dofs, fixed_dofs, scripted_vel, tol, maxit, mode, st);
} catch (...) {
    return eolc::report_exception(c, "step_dofs");
}

// ---- learned block preconditioner ------------------------------------------
int eolc_learned_set_weights(eolc_ctx *c, const float *theta, int n)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (!theta || n != eolc::kLpParams) return fail(c, EOLC_ERR_INVALID, "learned_set_weights: need EOLC_LEARNED_NPARAMS values");
    for (int k = 0; k < n; ++k)
        if (!std::isfinite(theta[k])) return fail(c, EOLC_ERR_NONFINITE, "learned_set_weights: non-finite parameter");
    EOLC_HIP(hipSetDevice(c->device));
    eolc::LearnedState &L = c->lp;
    if (!L.theta) {
        EOLC_HIP(dalloc(c, &L.theta, (size_t)eolc::kLpParams));
        EOLC_HIP(dalloc(c, &L.grad, (size_t)eolc::kLpParams));
        EOLC_HIP(dalloc(c, &L.adam_m, (size_t)eolc::kLpParams));
        EOLC_HIP(dalloc(c, &L.adam_v, (size_t)eolc::kLpParams));
        EOLC_HIP(dalloc(c, &L.wq, (size_t)eolc::kLqSize));
    }
    EOLC_HIP(hipMemcpyAsync(L.theta, theta, sizeof(float) * n, hipMemcpyHostToDevice, c->stream));
    EOLC_HIP(hipMemsetAsync(L.adam_m, 0, sizeof(float) * n, c->stream));
    EOLC_HIP(hipMemsetAsync(L.adam_v, 0, sizeof(float) * n, c->stream));
    EOLC_HIP(hipMemsetAsync(L.grad, 0, sizeof(float) * n, c->stream));
    EOLC_HIP(eolc::learned_quantize(c));
    EOLC_HIP(hipStreamSynchronize(c->stream));
    L.adam_t = 0;
    L.have_weights = true;
    L.zero_output = true;
    for (int k = eolc::kLpOffOut; k < n; ++k) L.zero_output = L.zero_output && theta[k] == 0.0f;
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "learned_set_weights");
}

int eolc_learned_get_weights(eolc_ctx *c, float *theta, int n)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (!theta || n != eolc::kLpParams) return fail(c, EOLC_ERR_INVALID, "learned_get_weights: need EOLC_LEARNED_NPARAMS values");
    if (!c->lp.have_weights) return fail(c, EOLC_ERR_INVALID, "learned_get_weights: no weights set");
    EOLC_HIP(hipSetDevice(c->device));
    EOLC_HIP(hipMemcpyAsync(theta, c->lp.theta, sizeof(float) * n, hipMemcpyDeviceToHost, c->stream));
    EOLC_HIP(hipStreamSynchronize(c->stream));
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "learned_get_weights");
}

static int learned_ready(eolc_ctx *c, const char *who, int n_fixed, const int *fixed_dof)
{
    if (!c->have_forces) return fail(c, EOLC_ERR_INVALID, std::string(who) + ": call forces_fill first");
    if (!c->lp.have_weights) return fail(c, EOLC_ERR_INVALID, std::string(who) + ": no weights set");
    if (n_fixed < 0 || (n_fixed && !fixed_dof)) return fail(c, EOLC_ERR_INVALID, std::string(who) + ": bad fixed DOFs");
    EOLC_HIP(hipSetDevice(c->device));
    return learned_buffers(c);
}

int eolc_learned_infer(eolc_ctx *c, int n_fixed, const int *fixed_dof, float *out, float *P)
try {
    if (!c) return EOLC_ERR_INVALID;
    int rc = learned_ready(c, "learned_infer", n_fixed, fixed_dof);
    if (rc) return rc;
    FixedScope fs;
    if ((rc = fs.set(c, n_fixed, fixed_dof, nullptr))) return rc;
    const size_t N = (size_t)c->topo.N;
    EOLC_HIP(eolc::learned_precond(c));
    std::vector<float> soa(P ? eolc::kDsym * N : 0);
    if (out) EOLC_HIP(hipMemcpyAsync(out, c->lp.o, sizeof(float) * 6 * N, hipMemcpyDeviceToHost, c->stream));
    if (P)
        EOLC_HIP(hipMemcpyAsync(soa.data(), c->D32, sizeof(float) * eolc::kDsym * N, hipMemcpyDeviceToHost, c->stream));
    EOLC_HIP(hipStreamSynchronize(c->stream));
    if (P)
        for (size_t i = 0; i < N; ++i)
            for (int k = 0; k < 9; ++k) P[9 * i + k] = soa[(size_t)eolc::kSymIdx[k] * N + i];
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "learned_infer");
}

static int learned_upload_error(eolc_ctx *c, const double *e)
{
    const size_t N = (size_t)c->topo.N;
    for (size_t k = 0; k < 3 * N; ++k)
        if (!std::isfinite(e[k])) return fail(c, EOLC_ERR_NONFINITE, "learned: non-finite error vector");
    EOLC_HIP(hipMemcpyAsync(c->lp.e, e, sizeof(double) * 3 * N, hipMemcpyHostToDevice, c->stream));
    EOLC_HIP(eolc::learned_mask_error(c));
    return EOLC_OK;
}

int eolc_learned_set_multilevel(eolc_ctx *c, int on)
try {
    if (!c) return EOLC_ERR_INVALID;
    c->lp.with_ml = on != 0;
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "learned_set_multilevel");
}

// The training loss inside the multilevel preconditioner needs the hierarchy, the
// FP64 fine blocks (the restriction divides by them nowhere, but ml_setup reads the
// current fixed set) and the Galerkin blocks of every level for the current MDK.
static int learned_ml_setup(eolc_ctx *c)
{
    const int rc = ml_lin_prepare(c, ml_lin_for(c, true));
    if (rc) return rc;
    EOLC_HIP(eolc::pcg_setup(c, false, 0, ml_lin_for(c, true)));   // the coarse levels of the learned mixed solve
    return EOLC_OK;
}

int eolc_learned_loss_grad(eolc_ctx *c, int n_fixed, const int *fixed_dof, const double *e, double *loss, float *grad)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (!e) return fail(c, EOLC_ERR_INVALID, "learned_loss_grad: null error vector");
    int rc = learned_ready(c, "learned_loss_grad", n_fixed, fixed_dof);
    if (rc) return rc;
    if (c->lp.with_ml && (rc = ml_prepare(c))) return rc;
    FixedScope fs;
    if ((rc = fs.set(c, n_fixed, fixed_dof, nullptr))) return rc;
    MlScope mls(c, c->lp.with_ml);
    if (c->lp.with_ml && (rc = learned_ml_setup(c))) return rc;
    if ((rc = learned_upload_error(c, e))) return rc;
    EOLC_HIP(eolc::learned_loss_grad(c));
    double sc[8];
    EOLC_HIP(hipMemcpyAsync(sc, c->lp.scal, sizeof(sc), hipMemcpyDeviceToHost, c->stream));
    if (grad) EOLC_HIP(hipMemcpyAsync(grad, c->lp.grad, sizeof(float) * eolc::kLpParams, hipMemcpyDeviceToHost, c->stream));
    EOLC_HIP(hipStreamSynchronize(c->stream));
    if (loss) *loss = sc[4];
    if (!std::isfinite(sc[4]) || !(sc[1] > 0.0)) return fail(c, EOLC_ERR_NONFINITE, "learned_loss_grad: r.e <= 0 or non-finite loss");
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "learned_loss_grad");
}

int eolc_learned_train(eolc_ctx *c, int n_fixed, const int *fixed_dof, const double *e, int steps, double lr,
                       double *losses)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (!e || steps <= 0 || !(lr > 0.0)) return fail(c, EOLC_ERR_INVALID, "learned_train: bad arguments");
    int rc = learned_ready(c, "learned_train", n_fixed, fixed_dof);
    if (rc) return rc;
    if (c->lp.with_ml && (rc = ml_prepare(c))) return rc;
    FixedScope fs;
    if ((rc = fs.set(c, n_fixed, fixed_dof, nullptr))) return rc;
    MlScope mls(c, c->lp.with_ml);
    if (c->lp.with_ml && (rc = learned_ml_setup(c))) return rc;
    if ((rc = learned_upload_error(c, e))) return rc;
    for (int s = 0; s < steps; ++s) {
        EOLC_HIP(eolc::learned_loss_grad(c));
        if (losses) EOLC_HIP(hipMemcpyAsync(&losses[s], c->lp.scal + 4, sizeof(double), hipMemcpyDeviceToHost, c->stream));
        EOLC_HIP(eolc::learned_adam(c, (float)lr));
        c->lp.zero_output = false;
    }
    EOLC_HIP(hipStreamSynchronize(c->stream));
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "learned_train");
}

// ---- batched cloths -------------------------------------------------------------
static void batch_invalidate_graph(eolc_ctx *c) { drop_batch_graph(c); }

int eolc_set_batch_topology(eolc_ctx *c, int B, const int *Nb, const double *X, const int *Fb, const int *tri,
                            const int *Hb, const int *hinge, const int *nfixb, const int *fixed, int *vert_off)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (B <= 0 || !Nb || !X || !Fb || !Hb || !nfixb)
        return fail(c, EOLC_ERR_INVALID, "set_batch_topology: bad sizes or null arrays");
    std::vector<int> off(B + 1, 0);
    int64_t Ftot = 0, Htot = 0, Ptot = 0, Ntot = 0;
    for (int b = 0; b < B; ++b) {
        if (Nb[b] <= 0 || Fb[b] < 0 || Hb[b] < 0 || nfixb[b] < 0)
            return fail(c, EOLC_ERR_INVALID, "set_batch_topology: negative or zero instance size");
        const int64_t padded = ((int64_t)Nb[b] + eolc::kSlice - 1) / eolc::kSlice * eolc::kSlice;
        Ntot += padded;
        if (Ntot > INT32_MAX / 3) return fail(c, EOLC_ERR_INVALID, "set_batch_topology: too many vertices");
        off[b + 1] = (int)Ntot;
        Ftot += Fb[b];
        Htot += Hb[b];
        Ptot += nfixb[b];
    }
    if ((Ftot && !tri) || (Htot && !hinge) || (Ptot && !fixed))
        return fail(c, EOLC_ERR_INVALID, "set_batch_topology: null element or fixed arrays");
    if (Ftot > INT32_MAX / 3 || Htot > INT32_MAX / 4)
        return fail(c, EOLC_ERR_INVALID, "set_batch_topology: too many elements");
    std::vector<double> Xc(2 * (size_t)Ntot, 0.0);
    std::vector<int> tc(3 * (size_t)Ftot), hc(4 * (size_t)Htot);
    std::vector<int> user_fixed, all_fixed;
    int64_t xr = 0, fr = 0, hr = 0, pr = 0;
    for (int b = 0; b < B; ++b) {
        for (int64_t k = 0; k < 2 * (int64_t)Nb[b]; ++k) Xc[2 * (size_t)off[b] + k] = X[2 * xr + k];
        for (int64_t k = 0; k < 3 * (int64_t)Fb[b]; ++k) {
            const int v = tri[3 * fr + k];
            if (v < 0 || v >= Nb[b]) return fail(c, EOLC_ERR_INVALID, "set_batch_topology: triangle index out of range");
            tc[3 * fr + k] = v + off[b];
        }
        for (int64_t k = 0; k < 4 * (int64_t)Hb[b]; ++k) {
            const int v = hinge[4 * hr + k];
            if (v < 0 || v >= Nb[b]) return fail(c, EOLC_ERR_INVALID, "set_batch_topology: hinge index out of range");
            hc[4 * hr + k] = v + off[b];
        }
        for (int k = 0; k < nfixb[b]; ++k) {
            const int v = fixed[pr + k];
            if (v < 0 || v >= Nb[b]) return fail(c, EOLC_ERR_INVALID, "set_batch_topology: fixed node out of range");
            user_fixed.push_back(v + off[b]);
        }
        for (int v = off[b] + Nb[b]; v < off[b + 1]; ++v) all_fixed.push_back(v);   // padding
        xr += Nb[b];
        fr += Fb[b];
        hr += Hb[b];
        pr += nfixb[b];
    }
    all_fixed.insert(all_fixed.end(), user_fixed.begin(), user_fixed.end());
    std::vector<int> seg(off.begin(), off.end() - 1);
    int rc = set_topology_impl(c, (int)Ntot, Xc.data(), (int)Ftot, tc.data(), (int)Htot, hc.data(), &seg);
    if (rc) return rc;
    eolc::BatchState &S = c->batch;
    S.B = B;
    S.inst.assign(B, eolc::BatchInst{});
    for (int b = 0; b < B; ++b) {
        eolc::BatchInst &I = S.inst[b];
        I.v0 = off[b];
        I.v1 = off[b + 1];
        I.s0 = off[b] / eolc::kSlice;
        I.s1 = off[b + 1] / eolc::kSlice;
        I.n_real = Nb[b];
    }
    S.order.resize(B);
    for (int b = 0; b < B; ++b) S.order[b] = b;
    std::stable_sort(S.order.begin(), S.order.end(), [&](int a, int b2) {
        return (S.inst[a].v1 - S.inst[a].v0) > (S.inst[b2].v1 - S.inst[b2].v0);
    });
    S.fixed_nodes = all_fixed;
    S.user_fixed = user_fixed;
    EOLC_HIP(dalloc(c, &S.d_inst, (size_t)B));
    EOLC_HIP(hipMemcpy(S.d_inst, S.inst.data(), sizeof(eolc::BatchInst) * B, hipMemcpyHostToDevice));
    EOLC_HIP(upload(c, &S.d_order, S.order));
    EOLC_HIP(upload(c, &S.d_user_fixed, S.user_fixed));
    EOLC_HIP(dalloc(c, &S.d_scripted, 3 * S.user_fixed.size()));
    EOLC_HIP(hipMemset(S.d_scripted, 0, sizeof(double) * 3 * std::max<size_t>(S.user_fixed.size(), 1)));
    EOLC_HIP(dalloc(c, &S.slice_part, 4 * (size_t)c->topo.nslices));
    EOLC_HIP(dalloc(c, &S.stats, 4 * (size_t)B));
    EOLC_HIP(dalloc(c, &S.dstats, 4 * (size_t)B));
    EOLC_HIP(upload(c, &S.d_all_fixed, all_fixed));
    S.n_all_fixed = (int)all_fixed.size();
    EOLC_HIP(eolc::batch_set_fixmask(c, S.d_all_fixed, S.n_all_fixed));
    EOLC_HIP(hipStreamSynchronize(c->stream));
    batch_invalidate_graph(c);
    if (vert_off)
        for (int b = 0; b <= B; ++b) vert_off[b] = off[b];
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "set_batch_topology");
}

int eolc_set_batch_material(eolc_ctx *c, const double *E, const double *beta)
try {
    if (!c) return EOLC_ERR_INVALID;
    eolc::BatchState &S = c->batch;
    if (S.B <= 0) return fail(c, EOLC_ERR_INVALID, "set_batch_material: no batch topology");
    if (!E || !beta) return fail(c, EOLC_ERR_INVALID, "set_batch_material: null arrays");
    for (int b = 0; b < S.B; ++b)
        if (!(E[b] > 0.0) || !(beta[b] >= 0.0) || !std::isfinite(E[b]) || !std::isfinite(beta[b]))
            return fail(c, EOLC_ERR_INVALID, "set_batch_material: E must be > 0 and beta >= 0");
    EOLC_HIP(hipSetDevice(c->device));
    const Topology &t = c->topo;
    // element -> instance: faces and hinges are stored instance by instance, and
    // every element's first vertex lies in its instance's vertex range
    std::vector<double> ef(t.F), bh(t.H);
    auto inst_of = [&](int v) {
        int lo = 0, hi = S.B - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (S.inst[mid].v0 <= v) lo = mid; else hi = mid - 1;
        }
        return lo;
    };
    for (int f = 0; f < t.F; ++f) ef[f] = E[inst_of(c->topo_tri[3 * f])];
    c->e_batch_max = 0.0;
    for (int b = 0; b < S.B; ++b) c->e_batch_max = std::max(c->e_batch_max, E[b]);
    std::vector<int> hh(4 * (size_t)t.H);
    if (t.H) EOLC_HIP(hipMemcpy(hh.data(), c->hinge, sizeof(int) * 4 * t.H, hipMemcpyDeviceToHost));
    for (int h = 0; h < t.H; ++h) bh[h] = beta[inst_of(hh[4 * h])];
    EOLC_HIP(upload(c, &c->e_face, ef));
    EOLC_HIP(upload(c, &c->beta_hinge, bh));
    c->have_mat = true;
    c->have_forces = false;
    batch_invalidate_graph(c);
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "set_batch_material");
}

// One batched step: forces, rhs, fixed values, persistent solve, x += h v.
static int batch_step_enqueue(eolc_ctx *c, const double grav[3], double h, double tol, int maxit, int variant)
{
    EOLC_HIP(eolc::launch_face_eval(c, h, grav));
    EOLC_HIP(eolc::launch_hinge_eval(c, h));
    EOLC_HIP(eolc::launch_assemble(c, eolc::kLaySoA));
    EOLC_HIP(eolc::launch_rhs(c, h));
    EOLC_HIP(eolc::batch_solve_async(c, tol, maxit, variant));
    EOLC_HIP(eolc::launch_update_x(c, h));
    return EOLC_OK;
}

static int batch_ready(eolc_ctx *c, const double grav[3], double h, double tol, int maxit)
{
    if (c->batch.B <= 0) return fail(c, EOLC_ERR_INVALID, "batch: no batch topology");
    if (!c->have_mat || !c->have_state || !c->e_face) return fail(c, EOLC_ERR_INVALID, "batch: material and state must be set");
    if (!grav || !(h > 0.0) || !(tol > 0.0) || maxit <= 0) return fail(c, EOLC_ERR_INVALID, "batch: bad arguments");
    return EOLC_OK;
}

static int batch_collect(eolc_ctx *c, eolc_solve_stats *stats, double ms)
{
    eolc::BatchState &S = c->batch;
    std::vector<int> is(4 * (size_t)S.B);
    std::vector<double> ds(4 * (size_t)S.B);
    EOLC_HIP(hipMemcpy(is.data(), S.stats, sizeof(int) * 4 * S.B, hipMemcpyDeviceToHost));
    EOLC_HIP(hipMemcpy(ds.data(), S.dstats, sizeof(double) * 4 * S.B, hipMemcpyDeviceToHost));
    int worst = EOLC_OK;
    for (int b = 0; b < S.B; ++b) {
        if (is[4 * b + 2] != EOLC_OK && worst == EOLC_OK) worst = is[4 * b + 2];
        if (stats) {
            eolc_solve_stats &st = stats[b];
            std::memset(&st, 0, sizeof(st));
            st.iters = is[4 * b];
            st.refinements = is[4 * b + 1];
            st.status = is[4 * b + 2];
            st.relres = ds[4 * b + 2];
            st.true_relres = ds[4 * b] > 0 ? std::sqrt(ds[4 * b + 1] / ds[4 * b]) : 0.0;
            st.ms_solve = ms;
        }
    }
    if (worst != EOLC_OK) c->err = "batch_step: at least one instance did not converge (see its stats)";
    return worst;
}

int eolc_batch_step(eolc_ctx *c, const double grav[3], double h, double tol, int maxit, int flags,
                    eolc_solve_stats *stats)
try {
    if (!c) return EOLC_ERR_INVALID;
    int rc = batch_ready(c, grav, h, tol, maxit);
    if (rc) return rc;
    EOLC_HIP(hipSetDevice(c->device));
    eolc::BatchState &S = c->batch;
    hipEvent_t e0, e1;
    EOLC_HIP(hipEventCreate(&e0));
    EOLC_HIP(hipEventCreate(&e1));
    EOLC_HIP(hipEventRecord(e0, c->stream));
    if (flags & EOLC_BATCH_EAGER) {
        rc = batch_step_enqueue(c, grav, h, tol, maxit, S.variant);
    } else {
        const bool same = S.graph && S.graph_h == h && S.graph_tol == tol && S.graph_maxit == maxit &&
                          S.graph_variant == S.variant && S.graph_ml == S.ml && S.graph_grav[0] == grav[0] &&
                          S.graph_grav[1] == grav[1] &&
                          S.graph_grav[2] == grav[2];
        if (!same) {
            batch_invalidate_graph(c);
            hipGraph_t g = nullptr;
            EOLC_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            rc = batch_step_enqueue(c, grav, h, tol, maxit, S.variant);
            hipError_t ec = hipStreamEndCapture(c->stream, &g);
            if (rc) {
                if (g) (void)hipGraphDestroy(g);
                return rc;
            }
            EOLC_HIP(ec);
            hipError_t ei = hipGraphInstantiate(&S.graph, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            EOLC_HIP(ei);
            S.graph_h = h;
            S.graph_tol = tol;
            S.graph_maxit = maxit;
            S.graph_variant = S.variant;
            S.graph_ml = S.ml;
            for (int k = 0; k < 3; ++k) S.graph_grav[k] = grav[k];
        }
        EOLC_HIP(hipGraphLaunch(S.graph, c->stream));
    }
    if (rc) return rc;
    EOLC_HIP(hipEventRecord(e1, c->stream));
    EOLC_HIP(hipEventSynchronize(e1));
    const double ms = elapsed_ms(e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    c->have_forces = true;
    c->h_last = h;
    return batch_collect(c, stats, ms);
} catch (...) {
    return eolc::report_exception(c, "batch_step");
}

int eolc_batch_autotune(eolc_ctx *c, const double grav[3], double h, double tol, int maxit, int *best, double *ms)
try {
    if (!c) return EOLC_ERR_INVALID;
    int rc = batch_ready(c, grav, h, tol, maxit);
    if (rc) return rc;
    EOLC_HIP(hipSetDevice(c->device));
    eolc::BatchState &S = c->batch;
    const size_t n3 = 3 * (size_t)c->topo.N;
    double *x0 = nullptr, *v0 = nullptr;
    EOLC_HIP(hipMalloc((void **)&x0, n3 * sizeof(double)));
    EOLC_HIP(hipMalloc((void **)&v0, n3 * sizeof(double)));
    EOLC_HIP(hipMemcpyAsync(x0, c->x, n3 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    EOLC_HIP(hipMemcpyAsync(v0, c->v, n3 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    // forces once, then the solve of the same system with each workgroup size
    EOLC_HIP(eolc::launch_face_eval(c, h, grav));
    EOLC_HIP(eolc::launch_hinge_eval(c, h));
    EOLC_HIP(eolc::launch_assemble(c, eolc::kLaySoA));
    EOLC_HIP(eolc::launch_rhs(c, h));
    hipEvent_t e0, e1;
    EOLC_HIP(hipEventCreate(&e0));
    EOLC_HIP(hipEventCreate(&e1));
    int arg = S.variant;
    double tb = 0.0;
    for (int v = 0; v < eolc::kBatchVariants; ++v) {
        EOLC_HIP(hipMemcpyAsync(c->v, v0, n3 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        EOLC_HIP(hipEventRecord(e0, c->stream));
        EOLC_HIP(eolc::batch_solve_async(c, tol, maxit, v));
        EOLC_HIP(hipEventRecord(e1, c->stream));
        EOLC_HIP(hipEventSynchronize(e1));
        const double t = elapsed_ms(e0, e1);
        if (ms) ms[v] = t;
        if (v == 0 || t < tb) {
            tb = t;
            arg = v;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    EOLC_HIP(hipMemcpyAsync(c->x, x0, n3 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    EOLC_HIP(hipMemcpyAsync(c->v, v0, n3 * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    EOLC_HIP(hipStreamSynchronize(c->stream));
    (void)hipFree(x0);
    (void)hipFree(v0);
    S.variant = arg;
    batch_invalidate_graph(c);
    c->have_forces = false;
    if (best) *best = arg;
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "batch_autotune");
}

int eolc_set_batch_scripted(eolc_ctx *c, const double *vel)
try {
    if (!c) return EOLC_ERR_INVALID;
    eolc::BatchState &S = c->batch;
    if (S.B <= 0) return fail(c, EOLC_ERR_INVALID, "set_batch_scripted: no batch topology");
    const size_t n = 3 * S.user_fixed.size();
    std::vector<double> h(n, 0.0);
    if (vel)
        for (size_t i = 0; i < n; ++i) {
            if (!std::isfinite(vel[i])) return fail(c, EOLC_ERR_INVALID, "set_batch_scripted: non-finite velocity");
            h[i] = vel[i];
        }
    if (n == 0) return EOLC_OK;
    EOLC_HIP(hipSetDevice(c->device));
    // ordered after every step already enqueued (eager or replayed graph); the
    // captured step reads the buffer at each replay, so no re-capture is needed
    EOLC_HIP(hipMemcpyAsync(S.d_scripted, h.data(), sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    EOLC_HIP(hipStreamSynchronize(c->stream));
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "set_batch_scripted");
}

int eolc_set_batch_multilevel(eolc_ctx *c, int on)
try {
    if (!c) return EOLC_ERR_INVALID;
    eolc::BatchState &S = c->batch;
    if (S.B <= 0) return fail(c, EOLC_ERR_INVALID, "set_batch_multilevel: no batch topology");
    EOLC_HIP(hipSetDevice(c->device));
    if (on && !S.d_ml_lo) {
        int rc = ml_prepare(c);
        if (rc) return rc;
        const eolc::MLState &m = c->ml;
        constexpr int St = eolc::kMLMaxLevels + 1;
        std::vector<int> lo((size_t)(S.B + 1) * St, 0);
        for (int l = 1; l <= m.lv.L; ++l) {
            int expect = 0;
            for (int b = 0; b < S.B; ++b) {
                int mn = INT_MAX, mx = -1;
                for (int v = S.inst[b].v0; v < S.inst[b].v1; ++v) {
                    mn = std::min(mn, m.maps[l][v]);
                    mx = std::max(mx, m.maps[l][v]);
                }
                if (mn != expect) return fail(c, EOLC_ERR_INVALID, "set_batch_multilevel: aggregates span instances");
                lo[(size_t)b * St + l] = mn;
                expect = mx + 1;
            }
            if (expect != m.lv.n[l]) return fail(c, EOLC_ERR_INVALID, "set_batch_multilevel: aggregates outside instances");
            lo[(size_t)S.B * St + l] = m.lv.n[l];
        }
        EOLC_HIP(upload(c, &S.d_ml_lo, lo));
    }
    S.ml = on != 0 && c->ml.lv.L > 0;
    batch_invalidate_graph(c);
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "set_batch_multilevel");
}

int eolc_set_batch_variant(eolc_ctx *c, int variant)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (variant < 0 || variant >= eolc::kBatchVariants) return fail(c, EOLC_ERR_INVALID, "set_batch_variant: bad variant");
    c->batch.variant = variant;
    batch_invalidate_graph(c);
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "set_batch_variant");
}

int eolc_time_spmv(eolc_ctx *c, int precision, int reps, double *avg_ms)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (!c->have_forces || (precision != 32 && precision != 64) || reps <= 0 || !avg_ms)
        return fail(c, EOLC_ERR_INVALID, "time_spmv: bad arguments or no assembled matrix");
    EOLC_HIP(hipSetDevice(c->device));
    float ms = 0.f;
    EOLC_HIP(eolc::time_spmv(c, precision == 32, -1, reps, &ms));
    *avg_ms = ms / reps;
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "time_spmv");
}

int eolc_spmv_layout_info(eolc_ctx *c, int64_t *info, int n)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (!c->have_topo || !info || n <= 0) return fail(c, EOLC_ERR_INVALID, "spmv_layout_info: no topology or no output");
    const Topology &t = c->topo;
    int64_t cross = 0, up = 0;
    if (t.pt_ok)
        for (size_t g = 0; g < t.pt_row.size(); ++g) {
            if (t.pt_row[g] < 0) continue;
            const unsigned cnt = (unsigned)t.pt_cnt[g];
            up += cnt & 0xffffu;
        }
    // (pt_lo padding holds kPatchCross too: only the entries rows read count)
    if (t.pt_ok) {
        const int C = eolc::kSlice;
        for (size_t g = 0; g < t.pt_row.size(); ++g) {
            if (t.pt_row[g] < 0) continue;
            const int sl = (int)(g / C), l = (int)(g % C), nlo = (int)((unsigned)t.pt_cnt[g] >> 16);
            for (int k = 0; k < nlo; ++k) cross += t.pt_lo[(size_t)t.pt_loff[sl] + (size_t)k * C + l] == eolc::kPatchCross;
        }
    }
    const int64_t v[8] = {t.pt_ok ? 1 : 0, up + cross, t.pt_ok ? (t.nb - t.N) - up : 0,
                          t.pt_ok ? (int64_t)t.pt_row.size() : 0, cross, t.pt_slots,
                          t.pt_ok ? (int64_t)t.pt_woff.back() : 0, t.pt_wmax};
    for (int k = 0; k < std::min(n, 8); ++k) info[k] = v[k];
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "spmv_layout_info");
}

int eolc_set_ml_linear(eolc_ctx *c, int where)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (where < 0 || where > 2) return fail(c, EOLC_ERR_INVALID, "set_ml_linear: 0, 1 or 2");
    c->ml_linear = where;
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "set_ml_linear");
}

int eolc_set_mixed_stall(eolc_ctx *c, int first_cycle, int min_cycle, int sticky)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (first_cycle <= 0 || min_cycle <= 0 || sticky < 0)
        return fail(c, EOLC_ERR_INVALID, "set_mixed_stall: need positive cycle limits and sticky >= 0");
    c->stall_first = first_cycle;
    c->stall_min = min_cycle;
    c->stall_sticky = sticky;
    c->fp64_sticky = std::min(c->fp64_sticky, sticky);
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "set_mixed_stall");
}

int eolc_set_stiff_threshold(eolc_ctx *c, double rel)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (!(rel >= 0.0) || !std::isfinite(rel)) return fail(c, EOLC_ERR_INVALID, "set_stiff_threshold: need rel >= 0");
    c->stiff_rel = rel;
    c->s32_valid = 0;   // the FP32 patch records depend on it
    drop_batch_graph(c);
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "set_stiff_threshold");
}

int eolc_set_mixed_delta(eolc_ctx *c, double delta)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (!(delta > 0.0 && delta < 1.0)) return fail(c, EOLC_ERR_INVALID, "set_mixed_delta: need 0 < delta < 1");
    c->mixed_delta = delta;
    c->mixed_delta_ml = delta;
    c->batch_delta = delta;
    drop_batch_graph(c);
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "set_mixed_delta");
}

int eolc_stream_copy(eolc_ctx *c, int64_t bytes, int reps, double *avg_ms)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (bytes < 16 || reps <= 0 || !avg_ms) return fail(c, EOLC_ERR_INVALID, "stream_copy: bad arguments");
    EOLC_HIP(hipSetDevice(c->device));
    float ms = 0.f;
    EOLC_HIP(eolc::stream_copy(c, (size_t)bytes, reps, &ms, 0));
    *avg_ms = ms / reps;
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "stream_copy");
}

int eolc_stream_read(eolc_ctx *c, int64_t bytes, int reps, double *avg_ms)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (bytes < 16 || reps <= 0 || !avg_ms) return fail(c, EOLC_ERR_INVALID, "stream_read: bad arguments");
    EOLC_HIP(hipSetDevice(c->device));
    float ms = 0.f;
    EOLC_HIP(eolc::stream_copy(c, (size_t)bytes, reps, &ms, 1));
    *avg_ms = ms / reps;
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "stream_read");
}

int eolc_autotune(eolc_ctx *c, int precision, int reps, int *best, double *ms)
try {
    if (!c) return EOLC_ERR_INVALID;
    if (!c->have_forces || (precision != 32 && precision != 64) || reps <= 0)
        return fail(c, EOLC_ERR_INVALID, "autotune: bad arguments or no assembled matrix");
    EOLC_HIP(hipSetDevice(c->device));
    int arg = 0;
    double tbest = 0.0;
    for (int v = 0; v < eolc::kSpmvVariants; ++v) {
        if (!eolc::spmv_usable(c->topo, v)) {
            if (ms) ms[v] = INFINITY;   // not applicable to this mesh
            continue;
        }
        float t = 0.f;
        EOLC_HIP(eolc::time_spmv(c, precision == 32, v, reps, &t));
        const double avg = (double)t / reps;
        if (ms) ms[v] = avg;
        if (v == 0 || avg < tbest) {
            tbest = avg;
            arg = v;
        }
    }
    c->spmv_variant[precision == 32 ? 0 : 1] = arg;
    if (best) *best = arg;
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "autotune");
}

int eolc_set_spmv_variant(eolc_ctx *c, int precision, int variant)
try {
    if (!c) return EOLC_ERR_INVALID;
    if ((precision != 32 && precision != 64) || variant < 0 || variant >= eolc::kSpmvVariants)
        return fail(c, EOLC_ERR_INVALID, "set_spmv_variant: bad precision or variant");
    if (eolc::spmv_staged(variant) && (!c->have_topo || c->topo.wmax > eolc::kStageVerts))
        return fail(c, EOLC_ERR_INVALID, "set_spmv_variant: the LDS-staged variants need every workgroup's vertex "
                                         "window within " + std::to_string(eolc::kStageVerts) + " vertices");
    if (eolc::spmv_patch(variant) && (!c->have_topo || !c->topo.pt_ok))
        return fail(c, EOLC_ERR_INVALID, "set_spmv_variant: the patch layout does not apply to this mesh" +
                                             (c->have_topo ? ": " + c->topo.pt_why : std::string()));
    c->spmv_variant[precision == 32 ? 0 : 1] = variant;
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(c, "set_spmv_variant");
}

}  // extern "C"
