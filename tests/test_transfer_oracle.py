"""Velocity transfer after a remesh (Cloth::velocityTransfer, non-EOL): the oracle's
brute-force restatement against closed forms (barycentric coordinates of a known
triangle, a linear velocity field reproduced exactly) and the reference's tie and
tolerance rules (first old vertex within 1e-6, first enclosing face, is_inside's
float thresholds)."""
import numpy as np

import oracle as O


def one_tri():
    u = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0]])
    v = np.array([[1.0, 2.0, 3.0], [4.0, 5.0, 6.0], [7.0, 8.0, 9.0]])
    return u, v, np.array([[0, 1, 2]])


def test_known_barycentric_coordinates():
    u, v, tri = one_tri()
    vn, src, face, bary, missing = O.velocity_transfer(u, v, tri, np.array([[0.25, 0.25], [0.0, 0.0]]))
    assert missing == 0
    assert src[0] == -1 and face[0] == 0
    assert np.array_equal(bary[:3], [0.5, 0.25, 0.25])
    assert np.array_equal(vn[:3], 0.5 * v[0] + 0.25 * v[1] + 0.25 * v[2])
    assert src[1] == 0 and face[1] == -1 and np.array_equal(vn[3:], v[0])


def test_linear_field_is_reproduced():
    rng = np.random.default_rng(3)
    n = 12
    g = np.linspace(0, 1, n + 1)
    U = np.stack(np.meshgrid(g, g), -1).reshape(-1, 2)
    tri = []
    for j in range(n):
        for i in range(n):
            a, b, c, d = j * (n + 1) + i, j * (n + 1) + i + 1, (j + 1) * (n + 1) + i, (j + 1) * (n + 1) + i + 1
            tri += [[a, b, d], [a, d, c]]
    tri = np.array(tri)
    A = rng.normal(size=(3, 2))
    c0 = rng.normal(size=3)
    V = U @ A.T + c0
    P = rng.uniform(0.001, 0.999, size=(500, 2))
    vn, src, face, bary, missing = O.velocity_transfer(U, V, tri, P)
    assert missing == 0 and np.all(face >= 0) and np.all(src == -1)
    assert np.max(np.abs(vn.reshape(-1, 3) - (P @ A.T + c0))) < 1e-12
    assert np.allclose(bary.reshape(-1, 3).sum(1), 1.0, atol=1e-15)


def test_tie_and_tolerance_rules():
    # two faces sharing the edge (1,0)-(0,1): a point on it goes to the first face
    u = np.array([[0.0, 0.0], [1.0, 0.0], [0.0, 1.0], [1.0, 1.0]])
    v = np.arange(12, dtype=float).reshape(4, 3)
    tri = np.array([[1, 3, 2], [0, 1, 2]])
    _, src, face, _, _ = O.velocity_transfer(u, v, tri, np.array([[0.5, 0.5]]))
    assert face[0] == 0
    # slightly outside a boundary edge: accepted within is_inside's float tolerance
    _, _, face, _, missing = O.velocity_transfer(u, v, tri, np.array([[0.5, -5e-7], [0.5, -1e-4]]))
    assert face[0] == 1 and face[1] == -1 and missing == 1
    # vertex matches: first old vertex in index order within 1e-6, strictly below
    u2 = np.array([[0.0, 0.0], [5e-7, 0.0], [1.0, 0.0], [0.0, 1.0]])
    v2 = np.arange(12, dtype=float).reshape(4, 3)
    tri2 = np.array([[0, 2, 3]])
    vn, src, _, _, _ = O.velocity_transfer(u2, v2, tri2, np.array([[4e-7, 0.0], [0.0, 1.5e-6]]))
    assert src[0] == 0 and np.array_equal(vn[:3], v2[0])
    assert src[1] == -1


def test_eol_correction_closed_form():
    """An old face whose vertices were EOL nodes: the new vertex's velocity is
    sum_i b_i (v_i - F V_i) with F = Dx DX^-1 of the old face (numpy's inverse here);
    V = 0 gives the plain transfer bit for bit; a vertex matched to an old vertex
    keeps its v."""
    rng = np.random.default_rng(11)
    u, v, tri = one_tri()
    u = u * np.array([0.7, 0.4]) + np.array([0.1, 0.2])
    x = np.concatenate([u, np.zeros((3, 1))], 1) + rng.normal(scale=0.05, size=(3, 3))
    P = np.array([[0.3, 0.3], [0.1, 0.2]])
    for eidx in ([0, 1, 2], [-1, 0, -1]):
        nE = sum(e >= 0 for e in eidx)
        V = rng.normal(size=(max(nE, 1), 2))
        vn, src, face, bary, missing = O.velocity_transfer_eol(u, x, v, eidx, V, tri, P)
        assert missing == 0 and face[0] == 0 and src[1] == 0
        F = np.stack([x[1] - x[0], x[2] - x[0]], 1) @ np.linalg.inv(np.stack([u[1] - u[0], u[2] - u[0]], 1))
        vw = v.copy()
        for i, e in enumerate(eidx):
            if e >= 0:
                vw[i] = v[i] - F @ V[e]
        ref = bary[:3] @ vw
        assert np.max(np.abs(vn[:3] - ref)) <= 1e-14 * np.max(np.abs(ref))
        assert np.array_equal(vn[3:], v[0])
        vz, *_ = O.velocity_transfer_eol(u, x, v, eidx, np.zeros_like(V), tri, P)
        vp, *_ = O.velocity_transfer(u, v, tri, P)
        assert np.array_equal(vz, vp)
