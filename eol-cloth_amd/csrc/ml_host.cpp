// Aggregation hierarchy of the additive multilevel preconditioner (host code, built
// once per topology and rebuilt with it, like the SELL layout and the gather maps).
//
// Level 1 aggregates the vertex graph of the block pattern: two vertices are
// adjacent when their 3x3 block is stored, so the bending stencil's second ring
// counts.  Level l + 1 aggregates the graph of level-l aggregates, two aggregates
// being adjacent when any of their members are.  On every level the seeds are taken
// greedily in index order such that no two lie within 2 hops of each other (3 above
// the first level);
// every node then joins the seed that a breadth-first search started from all seeds
// at once reaches first (a graph Voronoi partition).  The aggregates are connected
// and compact: about 12 vertices on the first level of a triangulated grid, then
// about 16 children per aggregate.  Levels stop at one aggregate per connected
// component, when a level merges nothing, or at kMLMaxLevels.
//
// Numbering is hierarchical (top-down, stable in the original order), so that every
// aggregate's members, and every aggregate's children, are contiguous ranges.
#include <algorithm>
#include <numeric>
#include <string>
#include <vector>

#include "eolc_internal.h"

namespace eolc {
namespace {

constexpr int kSeedDist = 2;        // level 1 (vertices)
constexpr int kSeedDistUpper = 3;   // levels >= 2: fewer, larger aggregates (fewer coarse levels)

struct Graph {
    int n = 0;
    std::vector<int> ip, ix;   // CSR adjacency without self loops
};

// Greedy seeds (pairwise distance > d) and breadth-first Voronoi regions.
std::vector<int> voronoi(const Graph &g, int d, int &na)
{
    const int n = g.n;
    std::vector<int> agg(n, -1), stamp(n, -1), depth(n, 0), q;
    std::vector<char> blocked(n, 0);
    std::vector<int> seeds;
    q.reserve(n);
    for (int s = 0; s < n; ++s) {
        if (blocked[s]) continue;
        seeds.push_back(s);
        q.clear();
        q.push_back(s);
        stamp[s] = s;
        depth[s] = 0;
        blocked[s] = 1;
        for (size_t h = 0; h < q.size(); ++h) {
            const int u = q[h];
            if (depth[u] == d) continue;
            for (int k = g.ip[u]; k < g.ip[u + 1]; ++k) {
                const int v = g.ix[k];
                if (stamp[v] == s) continue;
                stamp[v] = s;
                depth[v] = depth[u] + 1;
                blocked[v] = 1;
                q.push_back(v);
            }
        }
    }
    q.clear();
    for (int k = 0; k < (int)seeds.size(); ++k) {
        agg[seeds[k]] = k;
        q.push_back(seeds[k]);
    }
    for (size_t h = 0; h < q.size(); ++h) {
        const int u = q[h];
        for (int k = g.ip[u]; k < g.ip[u + 1]; ++k) {
            const int v = g.ix[k];
            if (agg[v] < 0) {
                agg[v] = agg[u];
                q.push_back(v);
            }
        }
    }
    na = (int)seeds.size();
    return agg;
}

// Graph of the aggregates (neighbour lists sorted).
Graph coarsen(const Graph &g, const std::vector<int> &agg, int na)
{
    std::vector<int> mp(na + 1, 0), mem(g.n);
    for (int u = 0; u < g.n; ++u) mp[agg[u] + 1]++;
    for (int a = 0; a < na; ++a) mp[a + 1] += mp[a];
    std::vector<int> fill(mp.begin(), mp.end() - 1);
    for (int u = 0; u < g.n; ++u) mem[fill[agg[u]]++] = u;
    Graph c;
    c.n = na;
    c.ip.assign(na + 1, 0);
    std::vector<int> mark(na, -1);
    for (int a = 0; a < na; ++a) {
        for (int m = mp[a]; m < mp[a + 1]; ++m) {
            const int u = mem[m];
            for (int k = g.ip[u]; k < g.ip[u + 1]; ++k) {
                const int b = agg[g.ix[k]];
                if (b != a && mark[b] != a) {
                    mark[b] = a;
                    c.ix.push_back(b);
                }
            }
        }
        c.ip[a + 1] = (int)c.ix.size();
        std::sort(c.ix.begin() + c.ip[a], c.ix.end());
    }
    return c;
}

// Stable counting sort of ids 0..n-1 by key (0..nk-1): position of each id.
std::vector<int> rank_by_key(const std::vector<int> &key, int nk)
{
    std::vector<int> cnt(nk + 1, 0);
    for (int k : key) cnt[k + 1]++;
    for (int b = 0; b < nk; ++b) cnt[b + 1] += cnt[b];
    std::vector<int> pos(key.size());
    for (size_t a = 0; a < key.size(); ++a) pos[a] = cnt[key[a]]++;
    return pos;
}

}  // namespace

std::string ml_build(const Topology &t, std::vector<std::vector<int>> &maps, std::vector<int> &hord,
                     std::vector<int> &ptr1, std::vector<int> &agg1, std::vector<int> &cptr, std::vector<int> &par,
                     std::vector<unsigned char> &lam, MLLevels &lv)
{
    const int N = t.N;
    Graph g;
    g.n = N;
    g.ip.assign(N + 1, 0);
    for (int i = 0; i < N; ++i) {
        for (int k = t.rowptr[i]; k < t.rowptr[i + 1]; ++k)
            if (t.col[k] != i) g.ix.push_back(t.col[k]);
        g.ip[i + 1] = (int)g.ix.size();
    }
    // P[l]: level-l id -> level-(l+1) id in the order the levels were built (P[0]: vertices)
    std::vector<std::vector<int>> P;
    std::vector<int> n{N};
    while ((int)P.size() < kMLMaxLevels) {
        int na = 0;
        std::vector<int> a = voronoi(g, P.empty() ? kSeedDist : kSeedDistUpper, na);
        if (na >= g.n) break;   // nothing merged: no edges left
        P.push_back(std::move(a));
        n.push_back(na);
        if (na == 1) break;
        g = coarsen(g, P.back(), na);
        if (g.ix.empty()) break;   // one aggregate per connected component
    }
    const int L = (int)P.size();
    lv = MLLevels{};
    lv.L = L;
    maps.clear();
    hord.clear(); ptr1.clear(); agg1.clear(); cptr.clear(); par.clear(); lam.clear();
    if (L == 0) return "";
    // hierarchical ids, top-down: nid[l][build id] = final id
    std::vector<std::vector<int>> nid(L + 1);
    nid[L].resize(n[L]);
    std::iota(nid[L].begin(), nid[L].end(), 0);
    for (int l = L - 1; l >= 1; --l) {
        std::vector<int> key(n[l]);
        for (int a = 0; a < n[l]; ++a) key[a] = nid[l + 1][P[l][a]];
        nid[l] = rank_by_key(key, n[l + 1]);
    }
    for (int l = 0; l <= L; ++l) lv.n[l] = n[l];
    lv.off[1] = 0;
    for (int l = 1; l <= L; ++l) lv.off[l + 1] = lv.off[l] + n[l];
    lv.coff[2] = 0;
    for (int l = 2; l < L; ++l) lv.coff[l + 1] = lv.coff[l] + n[l] + 1;
    // vertices
    agg1.resize(N);
    for (int i = 0; i < N; ++i) agg1[i] = nid[1][P[0][i]];
    {
        const std::vector<int> pos = rank_by_key(agg1, n[1]);
        hord.resize(N);
        for (int i = 0; i < N; ++i) hord[pos[i]] = i;
        ptr1.assign(n[1] + 1, 0);
        for (int i = 0; i < N; ++i) ptr1[agg1[i] + 1]++;
        for (int a = 0; a < n[1]; ++a) ptr1[a + 1] += ptr1[a];
    }
    // parents of levels 1..L-1 and child ranges of levels 2..L (final ids)
    par.assign(std::max(lv.off[L], 1), -1);
    cptr.clear();
    for (int l = 1; l < L; ++l)
        for (int a = 0; a < n[l]; ++a) par[lv.off[l] + nid[l][a]] = nid[l + 1][P[l][a]];
    for (int l = 2; l <= L; ++l) {
        std::vector<int> cnt(n[l] + 1, 0);
        for (int a = 0; a < n[l - 1]; ++a) cnt[par[lv.off[l - 1] + a] + 1]++;
        for (int b = 0; b < n[l]; ++b) cnt[b + 1] += cnt[b];
        for (int a = 1; a < n[l - 1]; ++a)   // children are contiguous by construction
            if (par[lv.off[l - 1] + a] < par[lv.off[l - 1] + a - 1]) return "multilevel: child ranges not contiguous";
        cptr.insert(cptr.end(), cnt.begin(), cnt.end());
    }
    if (cptr.empty()) cptr.push_back(0);
    // per-vertex aggregate at every level (final ids)
    maps.assign(L + 1, std::vector<int>());
    for (int l = 1; l <= L; ++l) maps[l].resize(N);
    for (int i = 0; i < N; ++i) {
        int o = P[0][i];
        for (int l = 1; l <= L; ++l) {
            maps[l][i] = nid[l][o];
            if (l < L) o = P[l][o];
        }
    }
    // first common level of every stored block (SELL entry order)
    lam.assign((size_t)t.nbs, 255);
    if (t.sellpos.empty()) return "";
    for (int i = 0; i < N; ++i)
        for (int k = t.rowptr[i]; k < t.rowptr[i + 1]; ++k) {
            const int j = t.col[k];
            int lm = 0;
            if (j != i) {
                lm = L + 1;
                for (int l = 1; l <= L; ++l)
                    if (maps[l][i] == maps[l][j]) {
                        lm = l;
                        break;
                    }
            }
            lam[(size_t)t.sellpos[k]] = (unsigned char)lm;
        }
    return "";
}

}  // namespace eolc

// Host-only form for callers and tests without a context: the hierarchy of a block
// pattern given as vertex CSR (rowptr: N + 1, col: columns of the stored blocks).
extern "C" int eolc_ml_partition(int N, const int *rowptr, const int *col, int *levels, int *counts, int *agg)
try {
    if (N <= 0 || !rowptr || !col || !levels) return EOLC_ERR_INVALID;
    eolc::Topology t;
    t.N = N;
    t.rowptr.assign(rowptr, rowptr + N + 1);
    if (t.rowptr[0] != 0) return EOLC_ERR_INVALID;
    for (int i = 0; i < N; ++i)
        if (t.rowptr[i + 1] < t.rowptr[i]) return EOLC_ERR_INVALID;
    t.col.assign(col, col + t.rowptr[N]);
    for (int j : t.col)
        if (j < 0 || j >= N) return EOLC_ERR_INVALID;
    std::vector<std::vector<int>> maps;
    std::vector<int> hord, ptr1, agg1, cptr, par;
    std::vector<unsigned char> lam;
    eolc::MLLevels lv{};
    if (!eolc::ml_build(t, maps, hord, ptr1, agg1, cptr, par, lam, lv).empty()) return EOLC_ERR_INVALID;
    *levels = lv.L;
    if (counts)
        for (int l = 0; l <= lv.L; ++l) counts[l] = lv.n[l];
    if (agg)
        for (int l = 1; l <= lv.L; ++l) std::copy(maps[l].begin(), maps[l].end(), agg + (size_t)(l - 1) * N);
    return EOLC_OK;
} catch (...) {
    return eolc::report_exception(nullptr, "ml_partition");
}
