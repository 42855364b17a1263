"""The oracle's element functions against the reference's own compiled kernels
(golden vectors in tests/golden/elements.npz), finite differences and invariants."""
import os

import numpy as np
import pytest

import oracle as O
from conftest import rel_err

GOLD = os.path.join(os.path.dirname(__file__), "golden", "elements.npz")
G = np.array([0.0, 0.0, -9.8])


@pytest.fixture(scope="module")
def gold():
    return np.load(GOLD)


def scaled_err(a, ref, scale):
    """max |a - ref| / max(max |ref|, scale): relative, with the element's natural
    magnitude as the floor so that rest/flat states (ref ~ rounding noise) compare
    on the same footing."""
    return float(np.max(np.abs(np.asarray(a) - ref)) / max(np.max(np.abs(ref)), scale))


def tri_scales(Xs, e=500.0):
    Xa, Xb, Xc = Xs
    A = 0.5 * abs((Xb[0] - Xa[0]) * (Xc[1] - Xa[1]) - (Xb[1] - Xa[1]) * (Xc[0] - Xa[0]))
    L = max(np.linalg.norm(Xb - Xa), np.linalg.norm(Xc - Xa), np.linalg.norm(Xc - Xb))
    return e * A / L, e * A / L ** 2        # force per unit strain, stiffness


def hinge_scales(Xs, beta=1e-5):
    X0, X1, X2, X3 = Xs
    A0 = 0.5 * ((X1[0] - X0[0]) * (X2[1] - X0[1]) - (X1[1] - X0[1]) * (X2[0] - X0[0]))
    A1 = 0.5 * ((X0[0] - X1[0]) * (X3[1] - X1[1]) - (X0[1] - X1[1]) * (X3[0] - X1[0]))
    L = np.linalg.norm(X1 - X0)
    w = 1.5 * beta * L ** 2 / (A0 + A1)
    return w, w / L, w / L ** 2


def test_membrane_golden(gold):
    """Tolerance: 1e-12 of max(|ref|, natural scale) for f and K."""
    for k in range(len(gold["mem_W"])):
        xs, Xs = gold["tri_x"][k], gold["tri_X"][k]
        W, f, K = O.membrane(*xs, *Xs, 500.0, 0.3, gold["tri_P"][k], gold["tri_Q"][k])
        sf, sk = tri_scales(Xs)
        assert scaled_err(f, gold["mem_f"][k], sf) < 1e-12, k
        assert scaled_err(K, gold["mem_K"][k], sk) < 1e-12, k
        assert abs(W - gold["mem_W"][k]) <= 1e-10 * max(abs(gold["mem_W"][k]), sf * 1e-6)


def test_inertial_golden(gold):
    for k in range(len(gold["ine_W"])):
        xs, Xs = gold["tri_x"][k], gold["tri_X"][k]
        W, f, M = O.inertial(*xs, *Xs, G, 0.1)
        assert rel_err(f, gold["ine_f"][k]) < 1e-14
        assert rel_err(M, gold["ine_M"][k]) < 1e-14
        assert abs(W - gold["ine_W"][k]) <= 1e-10 * abs(gold["ine_W"][k]) + 1e-300


def test_bending_golden(gold):
    """Tolerance: 1e-10 of max(|ref|, natural scale) for f and K; W = w (1 - phi)
    loses digits to cancellation near flat hinges, so W is checked against w."""
    for k in range(len(gold["hin_W"])):
        xs, Xs = gold["hin_x"][k], gold["hin_X"][k]
        W, f, K = O.bending(*xs, *Xs, 1e-5)
        w, sf, sk = hinge_scales(Xs)
        assert scaled_err(f, gold["hin_f"][k], sf) < 1e-10, k
        assert scaled_err(K, gold["hin_K"][k], sk) < 1e-10, k
        assert abs(W - gold["hin_W"][k]) <= 1e-12 * w, k


def test_reference_build_agrees_with_oracle():
    """When oracle/_ref is built, compare on fresh random elements too."""
    R = O.ref_lib()
    if R is None:
        pytest.skip("oracle/_ref not built (reference sources absent)")
    rng = np.random.default_rng(7)
    for _ in range(200):
        Xs = [np.zeros(2), np.array([1, 0.]) + rng.normal(0, .2, 2), np.array([0, 1.]) + rng.normal(0, .2, 2)]
        xs = [np.r_[X, 0] + rng.normal(0, .3, 3) for X in Xs]
        P, Q = O.face_PQ(*xs, *Xs)
        a = O.membrane(*xs, *Xs, 500, .3, P, Q)
        b = O.membrane(*xs, *Xs, 500, .3, P, Q, impl=R)
        assert rel_err(a[2], b[2]) < 1e-12 and rel_err(a[1], b[1]) < 1e-12
        Xh = [np.zeros(2), np.array([1, 0.]), np.array([.5, .8]) + rng.normal(0, .1, 2),
              np.array([.5, -.8]) + rng.normal(0, .1, 2)]
        xh = [np.r_[X, 0] + rng.normal(0, .3, 3) for X in Xh]
        a = O.bending(*xh, *Xh, 1e-5)
        b = O.bending(*xh, *Xh, 1e-5, impl=R)
        assert rel_err(a[2], b[2]) < 1e-10 and rel_err(a[1], b[1]) < 1e-10


def _fd_grad(fun, x, eps):
    g = np.zeros_like(x)
    for i in range(x.size):
        d = np.zeros_like(x); d[i] = eps
        g[i] = (fun(x + d) - fun(x - d)) / (2 * eps)
    return g


def test_membrane_finite_differences():
    rng = np.random.default_rng(3)
    Xs = [np.zeros(2), np.array([1.0, 0.1]), np.array([0.2, 0.9])]
    x = np.concatenate([np.r_[X, 0] + rng.normal(0, .2, 3) for X in Xs])
    P, Q = O.face_PQ(x[0:3], x[3:6], x[6:9], *Xs)   # P, Q frozen (as in the reference)

    def W(y): return O.membrane(y[0:3], y[3:6], y[6:9], *Xs, 500, .3, P, Q)[0]

    def f(y): return O.membrane(y[0:3], y[3:6], y[6:9], *Xs, 500, .3, P, Q)[1]
    _, f0, K0 = O.membrane(x[0:3], x[3:6], x[6:9], *Xs, 500, .3, P, Q)
    assert rel_err(-_fd_grad(W, x, 1e-6), f0) < 1e-7
    Kfd = np.array([-_fd_grad(lambda y: f(y)[i], x, 1e-6) for i in range(9)])
    assert rel_err(Kfd, K0) < 1e-7
    assert np.allclose(K0, K0.T, rtol=0, atol=1e-14 * np.abs(K0).max())
    assert np.linalg.eigvalsh(K0).min() > -1e-10 * np.abs(K0).max()   # PSD for A > 0
    t = np.tile([0.3, -0.2, 0.7], 3)
    assert np.abs(K0 @ t).max() < 1e-12 * np.abs(K0).max()             # translation null space


def test_membrane_rest_state_and_rotation():
    Xs = [np.zeros(2), np.array([1.0, 0.1]), np.array([0.2, 0.9])]
    xs = [np.r_[X, 0.0] for X in Xs]
    P, Q = O.face_PQ(*xs, *Xs)
    W, f, K = O.membrane(*xs, *Xs, 500, .3, P, Q)
    assert abs(W) < 1e-12 and np.abs(f).max() < 1e-10
    # rigid rotation of a deformed state leaves W unchanged (P, Q recomputed)
    rng = np.random.default_rng(5)
    xd = [x + rng.normal(0, 0.1, 3) for x in xs]
    P, Q = O.face_PQ(*xd, *Xs)
    W1 = O.membrane(*xd, *Xs, 500, .3, P, Q)[0]
    A, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    A *= np.sign(np.linalg.det(A))
    xr = [A @ x for x in xd]
    P, Q = O.face_PQ(*xr, *Xs)
    W2 = O.membrane(*xr, *Xs, 500, .3, P, Q)[0]
    assert abs(W1 - W2) < 1e-10 * abs(W1)


def test_poldec_cases():
    rng = np.random.default_rng(11)
    for _ in range(50):
        M = rng.normal(size=(2, 2))
        Q = O.poldec(M.T.reshape(-1)).reshape(2, 2).T
        if np.linalg.det(M) > 0:
            assert np.allclose(Q.T @ Q, np.eye(2), atol=1e-13)
            assert np.linalg.det(Q) > 0
            S = Q.T @ M
            assert abs(S[0, 1] - S[1, 0]) < 1e-12 * np.abs(M).max()     # Q^T M symmetric
        else:
            assert np.allclose(Q.T @ Q, np.eye(2), atol=1e-13)
    Q = O.poldec(np.array([1.0, 2.0, 2.0, 4.0]))                           # det = 0: sign 0
    assert np.allclose(Q, np.array([1.0, 2.0, 2.0, 4.0]) / np.sqrt(5.0))


def test_inertial_closed_form():
    Xs = [np.zeros(2), np.array([2.0, 0.0]), np.array([0.0, 1.0])]
    xs = [np.r_[X, 0.0] for X in Xs]
    W, f, M = O.inertial(*xs, *Xs, G, 0.1)
    m = 0.1 * 1.0
    assert np.allclose(f.reshape(3, 3).sum(0), m * G, rtol=1e-15)
    assert abs(M.sum() - 3 * m) < 1e-15                                     # total mass per axis x3
    assert np.allclose(M[0:3, 0:3], m / 6 * np.eye(3)) and np.allclose(M[0:3, 3:6], m / 12 * np.eye(3))


def test_bending_finite_differences_and_invariants():
    rng = np.random.default_rng(9)
    Xs = [np.zeros(2), np.array([1.0, 0.0]), np.array([0.4, 0.8]), np.array([0.6, -0.7])]
    x = np.concatenate([np.r_[X, 0] + rng.normal(0, .15, 3) for X in Xs])

    def parts(y): return O.bending(y[0:3], y[3:6], y[6:9], y[9:12], *Xs, 1e-5)
    W0, f0, K0 = parts(x)
    assert rel_err(-_fd_grad(lambda y: parts(y)[0], x, 1e-6), f0) < 1e-6
    Kfd = np.array([-_fd_grad(lambda y: parts(y)[1][i], x, 1e-6) for i in range(12)])
    assert rel_err(Kfd, K0) < 1e-6
    assert np.abs(K0 - K0.T).max() < 1e-12 * np.abs(K0).max()
    t = np.tile([0.3, -0.2, 0.7], 4)
    assert np.abs(K0 @ t).max() < 1e-10 * np.abs(K0).max()
    assert np.abs(f0.reshape(4, 3).sum(0)).max() < 1e-12 * np.abs(f0).max()  # no net force
    flat = np.concatenate([np.r_[X, 0] for X in Xs])
    Wf, ff, _ = parts(flat)
    assert abs(Wf) < 1e-18 and np.abs(ff).max() < 1e-15
