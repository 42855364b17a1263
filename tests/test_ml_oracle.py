"""The multilevel preconditioner's FP64 oracle and the host library's aggregation
hierarchy (no GPU): the top-level Galerkin block against the total mass, symmetry
and positivity of P, the Voronoi property and connectivity of every level, the
library's partition against the restated rule, and the iteration gain over
block-Jacobi PCG."""
from collections import deque

import numpy as np
import scipy.sparse as sp

import eolc
import ml_oracle as ML
import oracle as O
from conftest import GRAV, H
from eolc import scene


def vertex_pattern(A, N):
    """Vertex CSR block pattern of a scalar 3N x 3N matrix."""
    c = A.tocoo()
    B = sp.csr_matrix((np.ones(c.nnz), (c.row // 3, c.col // 3)), shape=(N, N))
    B.sum_duplicates()
    B.sort_indices()
    return B.indptr.astype(np.int32), B.indices.astype(np.int32)


def system(n, sigma_x=1e-3, seed_x=1234):
    sc = scene.Scene(n, sigma_x=sigma_x, sigma_v=0.05, seed_x=seed_x)
    rc, f, M, A = O.forces_fill(sc.x0, sc.X, sc.tri, sc.hinge, O.material(), GRAV, H)
    assert rc == 0
    return sc, f, M.tocsr(), A.tocsr()


def test_top_level_block_is_total_mass():
    """Unconstrained cloth: translations are in the stiffness null space, so the
    Galerkin block of the single top-level aggregate is sum_ij M_ij = (rho * area) I."""
    sc, f, M, A = system(12, sigma_x=2e-2)
    rp, cl = vertex_pattern(A, sc.N)
    maps = ML.hierarchy(rp, cl)
    assert len(np.unique(maps[-1])) == 1
    Z = ML.prolongation(maps[-1], sc.N)
    G = (Z.T @ A @ Z).toarray()
    total = 0.1 * 1.0   # density 0.1 kg/m^2 on a 1 m x 1 m sheet
    assert np.max(np.abs(G - total * np.eye(3))) <= 1e-9 * total
    # the same from the mass matrix alone, and the top block of P's restatement
    assert abs(M.sum() / 3 - total) <= 1e-12 * total
    levels = ML.coarse_levels(A, sc.N, maps)
    assert np.allclose(levels[-1][1][0], np.eye(3) / total, rtol=1e-9)


def test_preconditioner_is_symmetric_positive_definite():
    sc, f, M, A = system(10)
    rp, cl = vertex_pattern(A, sc.N)
    maps = ML.hierarchy(rp, cl)
    fd = np.concatenate([np.arange(3 * i, 3 * i + 3) for i in sc.fixed])
    rng = np.random.default_rng(4)
    u, v = rng.normal(size=(2, 3 * sc.N))
    u[fd] = 0.0
    v[fd] = 0.0
    Pu = ML.apply(A, sc.N, maps, u, fd)
    Pv = ML.apply(A, sc.N, maps, v, fd)
    assert abs(u @ Pv - v @ Pu) <= 1e-12 * abs(u @ Pv)
    assert u @ Pu > 0 and v @ Pv > 0
    assert np.all(Pu[fd] == 0.0)


def bfs_dist(adj, src):
    d = {s: 0 for s in src}
    q = deque(src)
    while q:
        x = q.popleft()
        for y in adj[x]:
            if y not in d:
                d[y] = d[x] + 1
                q.append(y)
    return d


def test_voronoi_property_and_connected_aggregates():
    sc, f, M, A = system(16)
    rp, cl = vertex_pattern(A, sc.N)
    adj = ML.vertex_graph(rp, cl)
    agg, seeds = ML.voronoi(adj)
    # seeds pairwise farther than the seed distance
    for s in seeds[:40]:
        d = bfs_dist(adj, [s])
        assert all(d.get(t, 99) > ML.SEED_DIST for t in seeds if t != s)
    # every node's seed is a nearest seed
    dmin = bfs_dist(adj, seeds)
    for k, s in enumerate(seeds[:40]):
        ds = bfs_dist(adj, [s])
        for v in np.nonzero(agg == k)[0]:
            assert ds[v] == dmin[v]
    # aggregates connected
    for k in range(len(seeds)):
        mem = set(np.nonzero(agg == k)[0].tolist())
        start = next(iter(mem))
        seen, q = {start}, deque([start])
        while q:
            x = q.popleft()
            for y in adj[x]:
                if y in mem and y not in seen:
                    seen.add(y)
                    q.append(y)
        assert seen == mem


def test_library_partition_matches_rule():
    """eolc_ml_partition (host code of the library) gives the restated partition on
    every level, for the grid scene and for two disconnected sheets."""
    sc, f, M, A = system(20)
    rp, cl = vertex_pattern(A, sc.N)
    ref = ML.hierarchy(rp, cl)
    got = eolc.ml_partition(rp, cl)
    assert len(got) == len(ref) >= 3
    for a, b in zip(got, ref):
        assert ML.same_partition(a, b)
    # two copies side by side: no aggregate may span both
    N = sc.N
    B = sp.block_diag([sp.csr_matrix((np.ones(cl.size), cl, rp), shape=(N, N))] * 2).tocsr()
    got2 = eolc.ml_partition(B.indptr.astype(np.int32), B.indices.astype(np.int32))
    ref2 = ML.hierarchy(B.indptr, B.indices)
    assert len(got2) == len(ref2)
    for a, b in zip(got2, ref2):
        assert ML.same_partition(a, b)
        assert not set(a[:N].tolist()) & set(a[N:].tolist())
    assert len(np.unique(got2[-1])) == 2


def test_multilevel_pcg_needs_fewer_iterations():
    """On one step's system the coarse levels cut the block-Jacobi PCG iterations
    (the oracle's own solve: same tolerance, same answer)."""
    sc, f, M, A = system(24)
    N = sc.N
    rp, cl = vertex_pattern(A, N)
    maps = ML.hierarchy(rp, cl)
    fd = np.concatenate([np.arange(3 * i, 3 * i + 3) for i in sc.fixed])
    b = M @ sc.v0 + H * f
    x0 = sc.v0.copy()
    x0[fd] *= 0.99
    D0 = ML.block_jacobi(A, N, fd)
    levels = ML.coarse_levels(A, N, maps, fd)
    bj = lambda r: ML.apply(A, N, [], r, fd, D0=D0, levels=[])
    mlp = lambda r: ML.apply(A, N, maps, r, fd, D0=D0, levels=levels)
    k_bj, x_bj = ML.pcg(A, b, x0, bj, fd, tol=1e-10)
    k_ml, x_ml = ML.pcg(A, b, x0, mlp, fd, tol=1e-10)
    assert k_ml < 0.7 * k_bj
    assert np.max(np.abs(x_ml - x_bj)) <= 1e-6 * np.max(np.abs(x_bj))


def test_linear_basis_geometry_and_span():
    """The linear coarse basis: on every aggregate the coordinate features are centred
    and of unit mean square (Z^T Z per aggregate is [[n, 0], [0, S]] per component
    with trace S = n), and a field a + B (X - c) that is linear on the aggregate lies
    in the span of its columns."""
    sc = scene.Scene(20)
    N = sc.N
    rng = np.random.default_rng(7)
    agg = rng.integers(0, 9, N)
    agg[:9] = np.arange(9)
    Z = ML.linear_prolongation(agg, N, sc.X)
    G = (Z.T @ Z).toarray()
    for a in range(9):
        mem = np.flatnonzero(agg == a)
        B = G[9 * a:9 * a + 9, 9 * a:9 * a + 9]
        n = len(mem)
        for c in range(3):
            sub = B[c::3, c::3]
            assert abs(sub[0, 0] - n) <= 1e-12 * n
            assert np.max(np.abs(sub[0, 1:])) <= 1e-10 * n
            assert abs(np.trace(sub[1:, 1:]) - n) <= 1e-10 * n
        assert np.max(np.abs(B - B[np.ix_(np.arange(9), np.arange(9))])) == 0.0
        # off-component entries vanish
        assert np.all(B[0::3, 1::3] == 0.0)
        # a linear field on the aggregate's members
        Xm = sc.X[mem]
        w = np.zeros((N, 3))
        w[mem] = rng.normal(size=3) + (Xm - Xm.mean(0)) @ rng.normal(size=(2, 3))
        cols = Z[:, 9 * a:9 * a + 9].toarray()
        y, *_ = np.linalg.lstsq(cols, w.reshape(-1), rcond=None)
        assert np.max(np.abs(cols @ y - w.reshape(-1))) <= 1e-10 * np.max(np.abs(w))


def test_chol_pinv_against_numpy():
    """chol_pinv: the inverse of a positive definite block; with a basis function that
    depends on the earlier ones, the inverse of the block without it (zero row and
    column); a zero block gives zero."""
    rng = np.random.default_rng(5)
    Q = rng.normal(size=(9, 9))
    S = Q @ Q.T + 0.1 * np.eye(9)
    assert np.allclose(ML.chol_pinv(S[None])[0], np.linalg.inv(S), rtol=1e-10, atol=1e-12)
    # basis function 4 = f1 - 2 f3 over 8 independent ones: G = C^T H C
    C = np.zeros((8, 9))
    for k in range(9):
        if k < 4:
            C[k, k] = 1.0
        elif k > 4:
            C[k - 1, k] = 1.0
    C[1, 4], C[3, 4] = 1.0, -2.0
    W = rng.normal(size=(8, 8))
    S2 = C.T @ (W @ W.T + 0.5 * np.eye(8)) @ C
    B = np.delete(np.arange(9), 4)
    Pi = ML.chol_pinv(S2[None])[0]
    ref = np.zeros((9, 9))
    ref[np.ix_(B, B)] = np.linalg.inv(S2[np.ix_(B, B)])
    assert np.allclose(Pi, ref, rtol=1e-8, atol=1e-10 * np.max(np.abs(ref)))
    assert np.all(ML.chol_pinv(np.zeros((1, 9, 9))) == 0.0)


def test_linear_levels_symmetric_positive_and_fewer_iterations():
    """With the linear fields on the levels >= 3, P stays symmetric positive definite,
    the top-level block's translation part is still the total mass, and the PCG needs
    fewer iterations than with the translations alone (same answer)."""
    sc, f, M, A = system(40)
    N = sc.N
    rp, cl = vertex_pattern(A, N)
    maps = ML.hierarchy(rp, cl)
    assert len(maps) >= 3
    fd = np.concatenate([np.arange(3 * i, 3 * i + 3) for i in sc.fixed])
    rng = np.random.default_rng(8)
    u, v = rng.normal(size=(2, 3 * N))
    u[fd] = v[fd] = 0.0
    Pu = ML.apply(A, N, maps, u, fd, X=sc.X)
    Pv = ML.apply(A, N, maps, v, fd, X=sc.X)
    assert abs(u @ Pv - v @ Pu) <= 1e-12 * abs(u @ Pv)
    assert u @ Pu > 0 and v @ Pv > 0 and np.all(Pu[fd] == 0.0)
    # unconstrained: translation part of the top block = total mass
    Zt = ML.linear_prolongation(maps[-1], N, sc.X)
    Gt = (Zt.T @ A @ Zt).toarray()
    if len(np.unique(maps[-1])) == 1:
        assert np.max(np.abs(Gt[:3, :3] - 0.1 * np.eye(3))) <= 1e-9 * 0.1
    b = M @ sc.v0 + H * f
    x0 = sc.v0.copy()
    x0[fd] *= 0.99
    D0 = ML.block_jacobi(A, N, fd)
    lt = ML.coarse_levels(A, N, maps, fd)
    ll = ML.coarse_levels(A, N, maps, fd, sc.X)
    assert [Di.shape[1] for _, Di in ll][:2] == [3, 3] and all(Di.shape[1] == 9 for _, Di in ll[2:])
    k_t, x_t = ML.pcg(A, b, x0, lambda r: ML.apply(A, N, maps, r, fd, D0=D0, levels=lt), fd, tol=1e-10)
    k_l, x_l = ML.pcg(A, b, x0, lambda r: ML.apply(A, N, maps, r, fd, D0=D0, levels=ll), fd, tol=1e-10)
    assert k_l < k_t
    assert np.max(np.abs(x_l - x_t)) <= 1e-6 * np.max(np.abs(x_t))
