"""CPU checks of the oracle's proximity-query restatement."""
import numpy as np

import oracle as O
from eolc import scene


def test_jitter_matches_libstdcxx_stream():
    """std::mt19937 (seed 1) + std::uniform_real_distribution<double>(-1, 1) as libstdc++
    draws them; the fixture comes from tests/golden/make_jitter.cpp."""
    import os
    gold = np.load(os.path.join(os.path.dirname(__file__), "golden", "jitter_seed1.npy"))
    assert np.array_equal(O.mt19937_uniform(1, gold.size), gold)


def test_create_edges_structure():
    """createEdges on a small grid: every interior edge is paired with the two faces
    that share it, boundary edges are singletons, order is by (max, min) vertex."""
    sc = scene.Scene(7, sigma_x=1e-3)
    x = sc.x0.reshape(-1, 3)
    edges = O.create_edges(sc.x0, sc.tri)
    nint = sum(e.internal for e in edges)
    assert nint == sc.H and len(edges) - nint == 4 * 7
    keys = [(max(e.verts[0], e.verts[1]), min(e.verts[0], e.verts[1])) for e in edges]
    assert keys == sorted(keys)
    for e in edges:
        f0 = set(sc.tri[e.faces[0]])
        assert {e.verts[0], e.verts[1], e.verts[2]} == f0
        if e.internal:
            f1 = set(sc.tri[e.faces[1]])
            assert {e.verts[0], e.verts[1], e.verts[3]} == f1
            n0, n1 = np.array(e.normals[:3]), np.array(e.normals[3:])
            assert abs(np.arccos(np.clip(n0 @ n1, -1, 1)) - e.angle) < 1e-10   # acos is ill-conditioned near 0


def test_create_edges_hash_wraps_like_int32():
    """For 3F (max index + 1) >= 2^31 the reference's int hash overflows; the order is
    then the stable order of the wrapped 32-bit values (two's complement)."""
    n = 300
    X, tri = scene.grid(n)
    F = len(tri)
    a = tri.reshape(-1); b = tri[:, [1, 2, 0]].reshape(-1)
    kmin = np.minimum(a, b).astype(np.int64) + 1
    kmax = np.maximum(a, b).astype(np.int64) + 1
    true_hash = kmin + (3 * F + 1) * kmax
    assert true_hash.max() > 2 ** 31                      # the wrap really happens here
    wrapped = ((true_hash + 2 ** 31) % 2 ** 32 - 2 ** 31).astype(np.int64)
    order = np.argsort(wrapped, kind="stable")
    x = np.zeros((X.shape[0], 3)); x[:, :2] = X
    edges = O.create_edges(x.reshape(-1), tri)
    # first record of every edge follows the wrapped order
    firsts = []
    k = 0
    ws = wrapped[order]
    while k < len(order):
        firsts.append(order[k])
        k += 2 if (k + 1 < len(order) and ws[k + 1] == ws[k]) else 1
    assert len(firsts) == len(edges)
    for e, r in zip(edges[:2000], firsts[:2000]):
        f, i = divmod(int(r), 3)
        assert e.faces[0] == f and e.verts[0] == tri[f][i] and e.verts[1] == tri[f][(i + 1) % 3]


def test_raytri_against_reference_build():
    """The oracle's intersect_triangle3_inc against the reference's raytri.cpp compiled
    from its source (oracle/_ref): same hit decision and bit-identical t on random rays,
    on rays aimed at points within a few 1e-6 of the edges and vertices (where the
    "-ZERO" = +1e-6 borders decide), on rays parallel to the plane, and with both
    orientations (det > 0 and det < 0)."""
    R = O.ref_lib()
    if R is None:
        import pytest
        pytest.skip("oracle/_ref not built")
    rng = np.random.default_rng(4)
    cases = []
    for _ in range(2000):
        cases.append(tuple(rng.normal(size=3) for _ in range(5)))
    for k in range(3000):
        v0, v1, v2 = rng.normal(size=3), rng.normal(size=3), rng.normal(size=3)
        # a target point with barycentric coordinates on or near the border
        b = rng.choice([0.0, 1.0, 0.5, 0.25, 0.3, 1e-6, -1e-6, 2e-6, 1 - 1e-6, 1 + 1e-6], size=2)
        b = b * rng.choice([1.0, 1.0 + 1e-7, 1.0 - 1e-7], size=2)
        p = v0 + b[0] * (v1 - v0) + b[1] * (v2 - v0)
        o = p + rng.normal(size=3)
        d = p - o
        if k % 5 == 0:
            d = (v1 - v0) + 1e-9 * rng.normal(size=3)     # (nearly) parallel to the plane
        if k % 2:
            v1, v2 = v2, v1                               # the other orientation
        cases.append((o, d, v0, v1, v2))
    hits = 0
    for o, d, v0, v1, v2 in cases:
        t, u, v = np.zeros(1), np.zeros(1), np.zeros(1)
        hit = bool(R.ref_tri3_inc(o, d, v0, v1, v2, t, u, v))
        ohit, ot = O.tri3_inc(o, d, v0, v1, v2)
        assert hit == ohit, (o, d, v0, v1, v2)
        if hit:
            hits += 1
            assert ot == t[0]
    assert 400 < hits < len(cases) - 400, hits
