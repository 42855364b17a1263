"""Multilevel additive preconditioner: FP64 restatement used as the checker.

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import it.  The preconditioner has no counterpart in the
reference, whose solver "none" runs Eigen's conjugate gradient
(GeneralizedSolver.cpp:107-172); it is an option of this library's solve.  The
restatement follows the definition

    P = D0^-1 + sum_l Z_l [diag_blocks(Z_l^T A Z_l)]^-1 Z_l^T,

with Z_l the piecewise-constant prolongation of the level-l aggregates (zero rows on
fixed DOFs) and the Galerkin products formed as explicit sparse products, and the
aggregation from its rule: on every level, seeds taken greedily in index order at
pairwise hop distance > 2 (> 3 above the first level), every node joined to the
seed a breadth-first search from all seeds reaches first; the next level
aggregates the graph of the aggregates.

Pins (tests/test_ml_oracle.py): for an unconstrained cloth the single top-level block
equals the total mass times I (translations span the stiffness null space); P is
symmetric and positive definite; every level has the Voronoi property (each node's
seed is a nearest seed) and connected aggregates; the host library's partition is
the same partition.
"""
from collections import deque

import numpy as np
import scipy.sparse as sp

SEED_DIST = 2         # level 1
SEED_DIST_UPPER = 3   # levels >= 2
MAX_LEVELS = 8


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


def coarse_levels(A, N, maps, fixed_dof=()):
    """[(Z_l, inverse Galerkin diagonal blocks)] for every level."""
    out = []
    for agg in maps:
        Z = prolongation(agg, N, fixed_dof)
        G = (Z.T @ A @ Z).tocsr()
        out.append((Z, inv_blocks(diag_blocks(G, G.shape[0] // 3))))
    return out


def apply(A, N, maps, r, fixed_dof=(), D0=None, levels=None):
    """z = P r (FP64); r's fixed entries are ignored, z's are 0."""
    r = np.array(r, dtype=np.float64)
    fd = np.asarray(fixed_dof, dtype=np.int64)
    r[fd] = 0.0
    if D0 is None:
        D0 = block_jacobi(A, N, fixed_dof)
    if levels is None:
        levels = coarse_levels(A, N, maps, fixed_dof)
    z = np.einsum('nij,nj->ni', D0, r.reshape(N, 3)).reshape(-1)
    for Z, Di in levels:
        R = (Z.T @ r).reshape(-1, 3)
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
