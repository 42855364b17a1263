"""The oracle's EOL extension of Forces::fill (Forces.cpp:177-329, 399-497, 580-683,
746-884) against an independent dense formulation.

For an element with vertices v and per-vertex deformation gradient F (the face's
Dx DX^-1; for a hinge the mean over its two faces), the EOL blocks the reference
pushes (fillXMI ... fillxXB with their transposes) are exactly the blocks of
T^T B T with T = blockdiag over the element's vertices of [I3 | -F] for an EOL
vertex (and I3 otherwise), B the element's M + dampingB h^2 K (faces) or
dampingB h^2 K (hinges); the X rows of f are T^T f.  An EOL hinge also adds its
bending force to the x rows (the non-EOL branch does not).  Here that sum is
formed densely in numpy, element by element, from the oracle's own element
routines, and compared with the oracle's triplet assembly (1e-12 of the largest
entry); with no EOL node the EOL fill equals the plain fill bit for bit."""
import numpy as np
import pytest

import oracle as O
from eolc import scene

GRAV = np.array([0.0, 0.0, -9.8])
H = 5e-3


def grad(x, X, a, b, c):
    Dx = np.stack([x[b] - x[a], x[c] - x[a]], axis=1)
    DX = np.stack([X[b] - X[a], X[c] - X[a]], axis=1)
    return Dx @ np.linalg.inv(DX)


def dense_eol(sc, mat, eol, flags=0):
    N = sc.N
    E = len(eol)
    n = 3 * N + 2 * E
    eidx = -np.ones(N, dtype=int)
    eidx[eol] = np.arange(E)
    x = sc.x0.reshape(-1, 3)
    X = sc.X
    Md, Kd = np.zeros((n, n)), np.zeros((n, n))
    f = np.zeros(n)
    s = mat.dampingB * H * H

    def tmap(verts, F):
        cols = []
        blocks = []
        for v in verts:
            cols.append(list(range(3 * v, 3 * v + 3)))
            if eidx[v] >= 0:
                cols[-1] += [3 * N + 2 * eidx[v], 3 * N + 2 * eidx[v] + 1]
        T = np.zeros((3 * len(verts), sum(len(c) for c in cols)))
        o = 0
        for k, v in enumerate(verts):
            T[3 * k:3 * k + 3, o:o + 3] = np.eye(3)
            if eidx[v] >= 0:
                T[3 * k:3 * k + 3, o + 3:o + 5] = -F
            o += len(cols[k])
        return T, sum(cols, [])

    for t in sc.tri:
        a, b, c = t
        P, Q = O.face_PQ(x[a], x[b], x[c], X[a], X[b], X[c])
        _, fm, Km = O.membrane(x[a], x[b], x[c], X[a], X[b], X[c], mat.e, mat.nu, P, Q)
        _, fi, Mi = O.inertial(x[a], x[b], x[c], X[a], X[b], X[c], GRAV, mat.density)
        T, idx = tmap(t, grad(x, X, a, b, c))
        ix = np.ix_(idx, idx)
        Md[ix] += T.T @ Mi @ T
        Kd[ix] += T.T @ (Mi + s * Km) @ T
        f[idx] += T.T @ (fm + fi)
    for q in sc.hinge:
        _, fb, Kb = O.bending(*[x[v] for v in q], *[X[v] for v in q], mat.beta)
        is_eol = np.any(eidx[q] >= 0)
        F = 0.5 * (grad(x, X, q[0], q[1], q[2]) + grad(x, X, q[1], q[0], q[3]))
        T, idx = tmap(q, F)
        Kd[np.ix_(idx, idx)] += T.T @ (s * Kb) @ T
        if is_eol:
            f[idx] += T.T @ fb
        elif flags & O.INCLUDE_BENDING_FORCE:
            f[(3 * np.asarray(q)[:, None] + np.arange(3)).reshape(-1)] += fb
    return f, Md, Kd


@pytest.mark.parametrize("which", ["boundary", "mixed"])
def test_eol_fill_is_the_mapped_element_sum(which):
    sc = scene.Scene(6, sigma_x=1e-3, sigma_v=0.0, seed_x=5)
    mat = O.material(0.1, 500.0, 0.3, 1e-5, 0.0, 1.0)
    N = sc.N
    ij = np.array([(k % 7, k // 7) for k in range(N)])
    bnd = np.where((ij[:, 0] == 0) | (ij[:, 0] == 6) | (ij[:, 1] == 0) | (ij[:, 1] == 6))[0]
    eol = bnd if which == "boundary" else np.concatenate([bnd[::3], [17, 24, 25]])
    eol = np.asarray(eol, dtype=np.int32)
    rc, f, M, MDK = O.forces_fill_eol(sc.x0, sc.X, sc.tri, sc.hinge, mat, GRAV, H, eol)
    assert rc == 0
    fd, Md, Kd = dense_eol(sc, mat, eol)
    n = 3 * N + 2 * len(eol)
    assert f.shape == (n,) and M.shape == (n, n) and MDK.shape == (n, n)
    for A, B in ((M.toarray(), Md), (MDK.toarray(), Kd)):
        assert np.max(np.abs(A - B)) <= 1e-12 * np.max(np.abs(B))
        assert np.max(np.abs(A - A.T)) <= 1e-13 * np.max(np.abs(B))
    assert np.max(np.abs(f - fd)) <= 1e-12 * np.max(np.abs(fd))
    # the X rows are not empty and the x part of an EOL hinge's force is there
    assert np.max(np.abs(f[3 * N:])) > 0 and np.max(np.abs(MDK.toarray()[3 * N:, :3 * N])) > 0
    rc0, f0, M0, MDK0 = O.forces_fill(sc.x0, sc.X, sc.tri, sc.hinge, mat, GRAV, H)
    assert np.max(np.abs(f[:3 * N] - f0)) > 0   # EOL hinges add their bending force
    # the x-x blocks do not change
    assert np.array_equal(MDK.toarray()[:3 * N, :3 * N], MDK0.toarray())
    assert np.array_equal(M.toarray()[:3 * N, :3 * N], M0.toarray())


def test_eol_fill_without_eol_nodes_is_the_plain_fill():
    sc = scene.Scene(5, sigma_x=1e-3, sigma_v=0.0, seed_x=9)
    mat = O.material(0.1, 500.0, 0.3, 1e-5, 0.0, 1.0)
    for flags in (0, O.INCLUDE_BENDING_FORCE):
        rc, f, M, MDK = O.forces_fill_eol(sc.x0, sc.X, sc.tri, sc.hinge, mat, GRAV, H, [], flags=flags)
        rc0, f0, M0, MDK0 = O.forces_fill(sc.x0, sc.X, sc.tri, sc.hinge, mat, GRAV, H, flags=flags)
        assert np.array_equal(f, f0)
        for A, B in ((M, M0), (MDK, MDK0)):
            assert np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)
            assert np.array_equal(A.data, B.data)
    assert O.forces_fill_eol(sc.x0, sc.X, sc.tri, sc.hinge, mat, GRAV, H, [3, 3])[0] == 1
