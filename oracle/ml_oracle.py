"""Multilevel additive preconditioner: FP64 restatement used as the checker.

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import it.  The preconditioner has no counterpart in the
reference, whose solver "none" runs Eigen's conjugate gradient
(GeneralizedSolver.cpp:107-172); it is an option of this library's solve.  The
restatement follows the definition

    P = D0^-1 + sum_l Z_l [diag_blocks(Z_l^T A Z_l)]^-1 Z_l^T,

with Z_l the piecewise-constant prolongation of the level-l aggregates (zero rows on
fixed DOFs) and the Galerkin products formed as explicit sparse products.  For the
single-cloth solve the levels >= LINEAR_FROM carry linear fields instead: Z_l holds,
per aggregate and component, the basis (1, (X - c) / s) of the material coordinates
(linear_prolongation) and the 9 x 9 blocks are inverted on the basis functions they
keep (chol_pinv); the batched solve keeps the translations (no X).  The
aggregation from its rule: on every level, seeds taken greedily in index order at
pairwise hop distance > 2 (> 3 above the first level), every node joined to the
seed a breadth-first search from all seeds reaches first; the next level
aggregates the graph of the aggregates.

Pins (tests/test_ml_oracle.py): for an unconstrained cloth the single top-level block
equals the total mass times I (translations span the stiffness null space); P is
symmetric and positive definite, with and without the linear levels; every level
has the Voronoi property (each node's seed is a nearest seed) and connected
aggregates; the host library's partition is the same partition; the linear basis is
centred and of unit mean square on every aggregate and spans the linear fields;
chol_pinv is numpy's inverse on a positive definite block and on the block without a
dependent basis function.
"""
from collections import deque

import numpy as np
import scipy.sparse as sp

SEED_DIST = 2         # level 1
SEED_DIST_UPPER = 3   # levels >= 2
MAX_LEVELS = 8
LINEAR_FROM = 3       # first level whose aggregates carry linear fields (single-cloth solve)
PIVOT_TOL = 1e-10     # scaled pivot at or below which a basis function is left out


def vertex_graph(rowptr, col):
    """Adjacency lists (sorted, no self loops) of a vertex CSR block pattern."""
    N = len(rowptr) - 1
    return [sorted({int(j) for j in col[rowptr[i]:rowptr[i + 1]] if j != i}) for i in range(N)]


def voronoi(adj, d=SEED_DIST):
    """Greedy seeds at pairwise distance > d, then breadth-first Voronoi regions."""
    n = len(adj)
    blocked = np.zeros(n, bool)
    seeds = []
    for s in range(n):
        if blocked[s]:
            continue
        seeds.append(s)
        dist = {s: 0}
        q = deque([s])
        blocked[s] = True
        while q:
            u = q.popleft()
            if dist[u] == d:
                continue
            for v in adj[u]:
                if v not in dist:
                    dist[v] = dist[u] + 1
                    blocked[v] = True
                    q.append(v)
    agg = -np.ones(n, dtype=np.int64)
    q = deque()
    for k, s in enumerate(seeds):
        agg[s] = k
        q.append(s)
    while q:
        u = q.popleft()
        for v in adj[u]:
            if agg[v] < 0:
                agg[v] = agg[u]
                q.append(v)
    return agg, seeds


def coarse_graph(adj, agg, na):
    nb = [set() for _ in range(na)]
    for u, vs in enumerate(adj):
        for v in vs:
            if agg[u] != agg[v]:
                nb[agg[u]].add(int(agg[v]))
    return [sorted(s) for s in nb]


def hierarchy(rowptr, col, d=SEED_DIST, d_upper=SEED_DIST_UPPER, max_levels=MAX_LEVELS):
    """Per-vertex aggregate maps of levels 1..L (labels in build order), with the
    library's stopping rules: stop when a level merges nothing, at one aggregate,
    when the aggregates are the connected components, or at max_levels."""
    adj = vertex_graph(rowptr, col)
    maps, cur = [], np.arange(len(adj))
    while len(maps) < max_levels:
        agg, seeds = voronoi(adj, d if not maps else d_upper)
        na = len(seeds)
        if na >= len(adj):
            break
        cur = agg[cur]
        maps.append(cur.copy())
        if na == 1:
            break
        adj = coarse_graph(adj, agg, na)
        if not any(adj):
            break
    return maps


def same_partition(a, b):
    """True when two labelings define the same partition."""
    a, b = np.asarray(a), np.asarray(b)
    if a.shape != b.shape:
        return False
    pairs = set(zip(a.tolist(), b.tolist()))
    return len(pairs) == len(set(a.tolist())) == len(set(b.tolist()))


def prolongation(agg, N, fixed_dof=()):
    """Z (3N x 3 n_agg): identity blocks from each vertex to its aggregate, zero rows on
    fixed DOFs."""
    agg = np.asarray(agg)
    na = int(agg.max()) + 1
    rows = np.arange(3 * N)
    keep = np.ones(3 * N, bool)
    keep[np.asarray(fixed_dof, dtype=np.int64)] = False
    rows = rows[keep]
    return sp.csr_matrix((np.ones(rows.size), (rows, 3 * agg[rows // 3] + rows % 3)), shape=(3 * N, 3 * na))


def inv_blocks(D):
    """Inverse of (n, 3, 3) symmetric blocks; components with a non-positive diagonal
    (every DOF of the aggregate fixed) left out, non-PD blocks give 0."""
    out = np.zeros_like(D)
    for a in range(D.shape[0]):
        S = 0.5 * (D[a] + D[a].T)
        keep = np.diag(S) > 0
        Sk = S[np.ix_(keep, keep)]
        if keep.any() and np.all(np.linalg.eigvalsh(Sk) > 0):
            out[a][np.ix_(keep, keep)] = np.linalg.inv(Sk)
    return out


def diag_blocks(A, n):
    D = np.zeros((n, 3, 3))
    c = A.tocoo()
    k = (c.row // 3) == (c.col // 3)
    np.add.at(D, (c.row[k] // 3, c.row[k] % 3, c.col[k] % 3), c.data[k])
    return D


def block_jacobi(A, N, fixed_dof=()):
    """Inverse diagonal blocks with fixed DOFs' rows and columns of the inverse zero."""
    D = diag_blocks(A, N)
    fixed = np.zeros(3 * N, bool)
    fixed[np.asarray(fixed_dof, dtype=np.int64)] = True
    for i in range(N):
        for c in range(3):
            if fixed[3 * i + c]:
                D[i, c, :] = 0.0
                D[i, :, c] = 0.0
                D[i, c, c] = 1.0
    Di = np.linalg.inv(D)
    for i in range(N):
        for c in range(3):
            if fixed[3 * i + c]:
                Di[i, c, :] = 0.0
                Di[i, :, c] = 0.0
    return Di


def coarse_levels(A, N, maps, fixed_dof=(), X=None, linear_from=LINEAR_FROM):
    """[(Z_l, inverse Galerkin diagonal blocks)] for every level.  With the material
    coordinates X, the levels from linear_from on carry linear fields
    (linear_prolongation, 9 x 9 blocks inverted by chol_pinv); without X every level
    carries the translations alone (3 x 3 blocks), as the batched solve does."""
    out = []
    for l, agg in enumerate(maps, start=1):
        if X is not None and l >= linear_from:
            Z = linear_prolongation(agg, N, X, fixed_dof)
            G = (Z.T @ A @ Z).tocoo()
            na = Z.shape[1] // 9
            Gd = np.zeros((na, 9, 9))
            k = (G.row // 9) == (G.col // 9)
            np.add.at(Gd, (G.row[k] // 9, G.row[k] % 9, G.col[k] % 9), G.data[k])
            out.append((Z, chol_pinv(Gd)))
        else:
            Z = prolongation(agg, N, fixed_dof)
            G = (Z.T @ A @ Z).tocsr()
            out.append((Z, inv_blocks(diag_blocks(G, G.shape[0] // 3))))
    return out


def aggregate_geometry(agg, X):
    """Centroid and RMS radius of every aggregate in the material coordinates X (N x 2);
    a radius of 0 (one vertex) counts as 1."""
    agg = np.asarray(agg)
    X = np.asarray(X, dtype=np.float64).reshape(len(agg), -1)[:, :2]
    na = int(agg.max()) + 1
    cnt = np.bincount(agg, minlength=na).astype(np.float64)
    c = np.stack([np.bincount(agg, X[:, j], minlength=na) for j in range(2)], 1) / cnt[:, None]
    d = X - c[agg]
    s = np.sqrt(np.bincount(agg, (d * d).sum(1), minlength=na) / cnt)
    s[s == 0.0] = 1.0
    return c, s


def linear_prolongation(agg, N, X, fixed_dof=()):
    """Z (3N x 9 n_agg): column 9A + 3m + a holds phi_m(i) on DOF (i, a) of each member
    i of aggregate A, phi = (1, (X_i - c_A) / s_A) (c_A, s_A from aggregate_geometry);
    zero rows on fixed DOFs.  The columns span the fields that are linear in the
    material coordinates on each aggregate."""
    agg = np.asarray(agg)
    c, s = aggregate_geometry(agg, X)
    X = np.asarray(X, dtype=np.float64).reshape(N, -1)[:, :2]
    u = (X - c[agg]) / s[agg][:, None]
    phi = np.stack([np.ones(N), u[:, 0], u[:, 1]], 1)
    keep = np.ones(3 * N, bool)
    keep[np.asarray(fixed_dof, dtype=np.int64)] = False
    rows, cols, vals = [], [], []
    for m in range(3):
        for a in range(3):
            rr = 3 * np.arange(N) + a
            k = keep[rr]
            rows.append(rr[k])
            cols.append((9 * agg + 3 * m + a)[k])
            vals.append(phi[k, m])
    na = int(agg.max()) + 1
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(3 * N, 9 * na))


def chol_pinv(G, tol=PIVOT_TOL):
    """Inverse of each symmetric positive semi-definite block of G (n, K, K) on the
    basis functions it keeps: with s_p = G_pp^-1/2 the scaled block s G s has a unit
    diagonal; an LDL^T factorisation of it in index order leaves out every basis
    function whose diagonal is not positive or whose pivot is at or below tol (its
    row and column of the inverse are zero), and the result is s (S_kept)^-1 s."""
    G = 0.5 * (np.asarray(G, dtype=np.float64) + np.transpose(G, (0, 2, 1)))
    n, K, _ = G.shape
    out = np.zeros_like(G)
    for b in range(n):
        g = np.diag(G[b]).copy()
        sc = np.where((g > 0) & np.isfinite(g), 1.0 / np.sqrt(np.where(g > 0, g, 1.0)), 0.0)
        S = sc[:, None] * G[b] * sc[None, :]
        Lm = np.zeros((K, K))
        d = np.zeros(K)
        kept = np.zeros(K, bool)
        for k in range(K):
            dk = S[k, k] - np.sum(Lm[k, :k] ** 2 * d[:k])
            if not (sc[k] > 0 and dk > tol and np.isfinite(dk)):
                continue
            kept[k] = True
            d[k] = dk
            Lm[k, k] = 1.0
            for i in range(k + 1, K):
                Lm[i, k] = (S[i, k] - np.sum(Lm[i, :k] * Lm[k, :k] * d[:k])) / dk
        idx = np.flatnonzero(kept)
        if idx.size:
            out[b][np.ix_(idx, idx)] = sc[idx, None] * np.linalg.inv(S[np.ix_(idx, idx)]) * sc[None, idx]
    return out


def apply(A, N, maps, r, fixed_dof=(), D0=None, levels=None, X=None):
    """z = P r (FP64); r's fixed entries are ignored, z's are 0."""
    r = np.array(r, dtype=np.float64)
    fd = np.asarray(fixed_dof, dtype=np.int64)
    r[fd] = 0.0
    if D0 is None:
        D0 = block_jacobi(A, N, fixed_dof)
    if levels is None:
        levels = coarse_levels(A, N, maps, fixed_dof, X)
    z = np.einsum('nij,nj->ni', D0, r.reshape(N, 3)).reshape(-1)
    for Z, Di in levels:
        R = (Z.T @ r).reshape(-1, Di.shape[1])
        z += Z @ np.einsum('nij,nj->ni', Di, R).reshape(-1)
    z[fd] = 0.0
    return z


def pcg(A, b, x0, P, fixed_dof=(), tol=1e-8, maxit=100000):
    """Dirichlet-reduced PCG (fixed DOFs keep x0's values): iterations to ||r|| <= tol ||b_red||."""
    fd = np.asarray(fixed_dof, dtype=np.int64)
    free = np.ones(A.shape[0], bool)
    free[fd] = False
    x = np.array(x0, dtype=np.float64)
    xf = np.zeros_like(x)
    xf[fd] = x[fd]
    b_red = b - A @ xf
    b_red[fd] = 0.0
    x[fd] = 0.0
    r = b_red - A @ x
    r[fd] = 0.0
    nb = np.linalg.norm(b_red)
    z = P(r)
    p = z.copy()
    rz = r @ z
    for k in range(1, maxit + 1):
        q = A @ p
        q[fd] = 0.0
        a = rz / (p @ q)
        x += a * p
        r -= a * q
        if np.linalg.norm(r) <= tol * nb:
            x[fd] = xf[fd]
            return k, x
        z = P(r)
        rz2 = r @ z
        p = z + (rz2 / rz) * p
        rz = rz2
    x[fd] = xf[fd]
    return maxit, x
