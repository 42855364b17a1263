"""Velocity transfer after a remesh on the GPU (grid search) against the oracle's
brute-force scans of Cloth::velocityTransfer: source vertex, enclosing face,
barycentric coordinates and velocities bitwise equal, on a remesh-like point set
(surviving vertices, edge midpoints on shared edges, random interior points,
near-vertex points either side of 1e-6, points just outside the sheet within
is_inside's tolerance), and an uncovered point reported.  At 1024^2 the GPU result
is checked through a linear velocity field, which the transfer reproduces."""
import numpy as np
import pytest

import eolc
import oracle as O
from eolc import scene

pytestmark = pytest.mark.gpu


def remesh_points(sc, rng, n_rand):
    X, tri = sc.X, sc.tri
    e = np.concatenate([tri[:, [0, 1]], tri[:, [1, 2]], tri[:, [2, 0]]])
    mid = 0.5 * (X[e[:, 0]] + X[e[:, 1]])
    surv = X[rng.choice(len(X), len(X) // 2, replace=False)] + rng.normal(scale=1e-8, size=(len(X) // 2, 2))
    lo, hi = X.min(0), X.max(0)
    near = X[rng.choice(len(X), 200)] + rng.choice([-1, 1], size=(200, 2)) * rng.uniform(6e-7, 2e-6, size=(200, 1))
    near = np.clip(near, lo, hi)   # off the sheet by more than is_inside's tolerance would find no face
    inside = rng.uniform(lo, hi, size=(n_rand, 2))
    edge = np.stack([rng.uniform(lo[0], hi[0], 50), np.full(50, lo[1] - 3e-9)], 1)
    return np.concatenate([surv, mid, near, inside, edge])


@pytest.mark.parametrize("n", [16, 40])
def test_transfer_matches_bruteforce(ctx, n):
    rng = np.random.default_rng(n)
    sc = scene.Scene(n, sigma_x=1e-3, sigma_v=0.05)
    P = remesh_points(sc, rng, 2000)
    print(f"transfer: old {n}x{n} ({sc.N} vertices), {len(P)} new points", flush=True)
    vo, so, fo, bo, missing = O.velocity_transfer(sc.X, sc.v0, sc.tri, P)
    rc, vg, sg, fg, bg, ms = ctx.velocity_transfer(sc.X, sc.v0, sc.tri, P, allow=(eolc.ERR_INVALID,))
    print(f"  {np.sum(sg >= 0)} surviving, {np.sum(fg >= 0)} in faces, missing {missing} "
          f"(gpu {np.sum((sg < 0) & (fg < 0))}), {ms:.3f} ms", flush=True)
    assert missing == 0 and rc == eolc.OK
    assert np.array_equal(sg, so) and np.array_equal(fg, fo)
    assert np.array_equal(bg, bo) and np.array_equal(vg, vo)
    assert np.sum(sg >= 0) > 0 and np.sum(fg >= 0) > 0


def test_uncovered_point_is_reported(ctx):
    sc = scene.Scene(8)
    P = np.array([[0.51, 0.47], [3.0, 3.0]])   # inside a face; far off the sheet
    rc, vg, sg, fg, bg, ms = ctx.velocity_transfer(sc.X, sc.v0, sc.tri, P, allow=(eolc.ERR_INVALID,))
    assert rc == eolc.ERR_INVALID and fg[1] == -1 and sg[1] == -1 and np.all(vg[3:] == 0.0)
    assert fg[0] >= 0
    _, so, fo, _, missing = O.velocity_transfer(sc.X, sc.v0, sc.tri, P)
    assert missing == 1 and fo[0] == fg[0] and so[0] == sg[0] == -1


def test_transfer_at_scale():
    """1024^2 old mesh (1,050,625 vertices), 1.5 M new points: every point found, a
    linear field reproduced to 1e-12, survivors take their vertex's velocity."""
    sc = scene.Scene(1024)
    rng = np.random.default_rng(5)
    A = rng.normal(size=(3, 2))
    V = sc.X @ A.T + 1.0
    P = np.concatenate([sc.X[::2] + 1e-9, rng.uniform(0.0, 1.0, size=(1_000_000, 2))])
    with eolc.Context(0) as c:
        rc, vg, sg, fg, bg, ms = c.velocity_transfer(sc.X, V, sc.tri, P)
    print(f"transfer 1024^2 -> {len(P)} points: {ms:.2f} ms on the device", flush=True)
    assert rc == eolc.OK
    n_s = len(sc.X[::2])
    assert np.array_equal(sg[:n_s], np.arange(0, sc.N, 2))
    r_s, r_f = sg[n_s:], fg[n_s:]
    assert np.all((r_s >= 0) | (r_f >= 0))   # a few random points fall within 1e-6 of a vertex
    inf = n_s + np.nonzero(r_f >= 0)[0]
    assert np.max(np.abs(vg.reshape(-1, 3)[inf] - (P[inf] @ A.T + 1.0))) < 1e-12
    onv = n_s + np.nonzero(r_s >= 0)[0]
    assert np.array_equal(vg.reshape(-1, 3)[onv], V[sg[onv]])


def test_transfer_with_eol_nodes_matches_bruteforce(ctx):
    """EOL nodes in the old mesh (the boundary and a few interior nodes): the GPU's
    EOL correction of new LAG vertices equals the oracle's, bit for bit, and changes
    the velocities of the points whose old face has an EOL vertex."""
    n = 16
    rng = np.random.default_rng(5)
    sc = scene.Scene(n, sigma_x=1e-3, sigma_v=0.05)
    ij = np.array([(k % (n + 1), k // (n + 1)) for k in range(sc.N)])
    eol = np.where((ij[:, 0] == 0) | (ij[:, 0] == n) | (ij[:, 1] == 0) | (ij[:, 1] == n))[0]
    eol = np.concatenate([eol, [5 * (n + 1) + 5, 9 * (n + 1) + 3]])
    eidx = -np.ones(sc.N, dtype=np.int32)
    eidx[eol] = np.arange(len(eol))
    V = rng.normal(scale=0.05, size=(len(eol), 2))
    P = remesh_points(sc, rng, 1000)
    print(f"transfer with {len(eol)} EOL nodes: {len(P)} new points", flush=True)
    vo, so, fo, bo, missing = O.velocity_transfer_eol(sc.X, sc.x0, sc.v0, eidx, V, sc.tri, P)
    rc, vg, sg, fg, bg, ms = ctx.velocity_transfer_eol(sc.X, sc.x0, sc.v0, eidx, V, sc.tri, P)
    assert missing == 0 and rc == eolc.OK
    assert np.array_equal(sg, so) and np.array_equal(fg, fo) and np.array_equal(bg, bo)
    assert np.array_equal(vg, vo)
    vp, *_ = O.velocity_transfer(sc.X, sc.v0, sc.tri, P)
    touched = np.array([f >= 0 and np.any(eidx[sc.tri[f]] >= 0) for f in fo])
    print(f"  {touched.sum()} points in faces with an EOL vertex", flush=True)
    assert touched.sum() > 0
    assert not np.array_equal(vg.reshape(-1, 3)[touched], vp.reshape(-1, 3)[touched])
    assert np.array_equal(vg.reshape(-1, 3)[~touched], vp.reshape(-1, 3)[~touched])
