"""Scene-step pieces on the CPU: the oracle pinned by the mathematics, and the
library's host-only entry points (eolc_rigid_integrate, eolc_box_step) against it.

Bars: the rigid-body update equals E0 expm(h [w^ v; 0 0]) to 1e-13 (scipy's matrix
exponential of the twist is the definition; the Rodrigues closed form must agree);
library vs oracle to 1e-14; schedule and row order as worked by hand."""
import numpy as np
import pytest
import scipy.linalg

import eolc
import scene_oracle as SO


def hat(w):
    return np.array([[0.0, -w[2], w[1]], [w[2], 0.0, -w[0]], [-w[1], w[0], 0.0]])


def twist_exp(phi, h):
    T = np.zeros((4, 4))
    T[:3, :3] = hat(phi[:3])
    T[:3, 3] = phi[3:6]
    return scipy.linalg.expm(h * T)


def random_frame(rng):
    Q, R = np.linalg.qr(rng.normal(size=(3, 3)))
    Q = Q @ np.diag(np.sign(np.diag(R)))
    if np.linalg.det(Q) < 0:
        Q[:, 0] = -Q[:, 0]
    E = np.eye(4)
    E[:3, :3] = Q
    E[:3, 3] = rng.normal(size=3)
    return E


def test_rigid_integrate_is_the_twist_exponential():
    rng = np.random.default_rng(5)
    for _ in range(20):
        E0 = random_frame(rng)
        phi = rng.normal(size=6) * np.array([3, 3, 3, 1, 1, 1])
        h = float(rng.uniform(1e-3, 0.2))
        ref = E0 @ twist_exp(phi, h)
        o = SO.rigid_integrate(E0, phi, h)
        assert np.max(np.abs(o - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))
        g = eolc.rigid_integrate(E0, phi, h)
        assert np.max(np.abs(g - o)) <= 1e-14 * max(1.0, np.max(np.abs(o)))
        R = g[:3, :3]
        assert np.allclose(R.T @ R, np.eye(3), atol=1e-14) and abs(np.linalg.det(R) - 1) < 1e-14
        assert np.array_equal(g[3], [0, 0, 0, 1])


def test_small_rotation_branch_is_a_translation():
    rng = np.random.default_rng(6)
    E0 = random_frame(rng)
    phi = np.array([1e-11, -3e-12, 2e-12, 0.4, -0.2, 0.1])
    h = 5e-3
    g = eolc.rigid_integrate(E0, phi, h)
    T = np.eye(4)
    T[:3, 3] = h * phi[3:]
    assert np.array_equal(g, E0 @ T) or np.max(np.abs(g - E0 @ T)) <= 1e-15
    assert np.max(np.abs(g - SO.rigid_integrate(E0, phi, h))) <= 1e-15
    # one step away from the branch the closed form takes over continuously
    phi2 = phi.copy()
    phi2[:3] = [2e-10, 0, 0]
    assert np.max(np.abs(eolc.rigid_integrate(E0, phi2, h) - E0 @ twist_exp(phi2, h))) <= 1e-13


def test_box_step_moves_in_the_world_frame():
    """Box::v is a world-frame velocity (w, u) of the box centre p: each step is the
    spatial screw exp(h [w^, u - w x p; 0 0]) applied on the left (rotation about
    the axis through the centre parallel to w, centre velocity u at the step's
    start), so after k steps the rotation is expm(k h w^) R0."""
    rng = np.random.default_rng(7)
    E = random_frame(rng)
    w = np.array([0.3, -1.1, 2.0])
    u = np.array([0.05, 0.4, -0.2])
    vel = np.concatenate([w, u])
    h, k = 5e-3, 40
    Eg, Eo = E.copy(), E.copy()
    for _ in range(k):
        Eg = eolc.box_step(Eg, vel, h)
        Eo = SO.box_step(Eo, vel, h)
    assert np.max(np.abs(Eg - Eo)) <= 1e-13
    Rk = scipy.linalg.expm(k * h * hat(w)) @ E[:3, :3]
    assert np.max(np.abs(Eg[:3, :3] - Rk)) <= 1e-12
    Es = E.copy()
    for _ in range(k):
        S = np.zeros((4, 4))
        S[:3, :3] = hat(w)
        S[:3, 3] = u - np.cross(w, Es[:3, 3])
        Es = scipy.linalg.expm(h * S) @ Es
    assert np.max(np.abs(Eg - Es)) <= 1e-12
    # no rotation: the centre moves by k h u exactly as a sum of steps
    Et = E.copy()
    for _ in range(k):
        Et = eolc.box_step(Et, np.concatenate([[0, 0, 0], u]), h)
    assert np.max(np.abs(Et[:3, 3] - (E[:3, 3] + k * h * u))) <= 1e-13
    assert np.array_equal(Et[:3, :3], E[:3, :3])
    # the library agrees with rigid_integrate on the body twist
    phi = np.concatenate([E[:3, :3].T @ w, E[:3, :3].T @ u])
    assert np.max(np.abs(eolc.box_step(E, vel, h) - eolc.rigid_integrate(E, phi, h))) <= 1e-15


def test_motion_rejects_bad_arguments():
    with pytest.raises(eolc.EolcError):
        eolc.rigid_integrate(np.eye(4), [np.nan, 0, 0, 0, 0, 0], 0.1)
    with pytest.raises(eolc.EolcError):
        eolc.box_step(np.eye(4), [0, 0, 0, 0, 0, 0], np.inf)


def test_update_fix_and_fixed_rows():
    whens = [0.0, 0.0475, 0.1]
    t, idx, switches = 0.0, 0, []
    for k in range(30):
        new = SO.update_fix(t, whens, idx)
        if new != idx:
            switches.append(k)
        idx = new
        t += 5e-3
    # t > 0.0475 first at k = 10 (t = 0.05); t > 0.1 first at k = 21 (the sum of
    # twenty 5e-3 steps is not above 0.1 in floating point, twenty-one are)
    s20 = 0.0
    for _ in range(20):
        s20 += 5e-3
    assert switches == ([10, 20] if s20 > 0.1 else [10, 21])
    spec = [[-1] * 6, [1, 1, 1, 0.1, 0, 0], [0, 0, 1, 0, 0, 0.05], [1, 0, 1, -0.2, 0, 0.3]]
    dofs, vel = SO.fixed_rows([0, 7, 9, 4], spec)
    assert dofs.tolist() == [21, 22, 23, 29, 12, 14]
    assert vel.tolist() == [0.1, 0.0, 0.0, 0.05, -0.2, 0.3]
