"""The FP64 oracle of the learned preconditioner: gradient against central finite
differences, SPD of the preconditioner, block-Jacobi at zero output."""
import numpy as np
import scipy.sparse.linalg as spla

import mpn_oracle as MO
import oracle as O
from conftest import GRAV, H
from eolc import scene


def small_system(n=6, seed=3):
    sc = scene.Scene(n, sigma_x=2e-2, sigma_v=0.05)
    rc, f, M, A = O.forces_fill(sc.x0, sc.X, sc.tri, sc.hinge, O.material(), GRAV, H)
    N = sc.N
    mass = np.array([M[3 * i, 3 * i] for i in range(N)])
    fixmask = np.zeros(N, dtype=np.int64)
    fixmask[sc.fixed] = 7
    rng = np.random.default_rng(seed)
    r = rng.normal(size=3 * N)
    for i in sc.fixed:
        r[3 * i:3 * i + 3] = 0
    free = np.ones(3 * N, bool)
    for i in sc.fixed:
        free[3 * i:3 * i + 3] = False
    Af = A.tocsr()[free][:, free]
    e = np.zeros(3 * N)
    e[free] = spla.spsolve(Af.tocsc(), r[free])
    # the loss sees A restricted to free DOFs: zero fixed rows/cols, identity diagonal
    import scipy.sparse as sp
    Dm = sp.diags(free.astype(float))
    Ar = (Dm @ A @ Dm + sp.diags((~free).astype(float))).tocsr()
    return Ar, mass, N, fixmask, r, e


def test_zero_output_is_block_jacobi():
    A, mass, N, fixmask, r, e = small_system()
    X, W, B = MO.features(A, mass, N, fixmask)
    P, _ = MO.precond_from_outputs(np.zeros((N, MO.NOUT)), B, N, fixmask)
    for i in range(N):
        D = B[(i, i)]
        if fixmask[i] == 0:
            assert np.allclose(P[i], np.linalg.inv(D), rtol=1e-12, atol=1e-12 * np.abs(np.linalg.inv(D)).max())


def test_preconditioner_spd_for_any_output():
    A, mass, N, fixmask, r, e = small_system()
    X, W, B = MO.features(A, mass, N, fixmask)
    rng = np.random.default_rng(0)
    P, _ = MO.precond_from_outputs(rng.normal(scale=5.0, size=(N, MO.NOUT)), B, N, fixmask)
    for i in range(N):
        if fixmask[i] == 0:
            w = np.linalg.eigvalsh(0.5 * (P[i] + P[i].T))
            assert w.min() > 0 and np.allclose(P[i], P[i].T)


def test_gradient_matches_finite_differences():
    A, mass, N, fixmask, r, e = small_system()
    theta = MO.init_params(77)
    rng = np.random.default_rng(1)
    theta = theta + rng.normal(scale=0.05, size=theta.size)     # non-zero outputs
    loss, grad, _ = MO.loss_and_grad(theta, A, mass, N, fixmask, r, e)
    assert 0.0 < loss < 1.0
    for trial in range(6):
        d = rng.normal(size=theta.size)
        eps = 1e-6
        lp, _, _ = MO.loss_and_grad(theta + eps * d, A, mass, N, fixmask, r, e)
        lm, _, _ = MO.loss_and_grad(theta - eps * d, A, mass, N, fixmask, r, e)
        fd = (lp - lm) / (2 * eps)
        assert abs(fd - grad @ d) <= 1e-6 * max(abs(fd), 1e-3), (fd, grad @ d)
