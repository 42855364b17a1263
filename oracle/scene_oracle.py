"""Scene-step pieces around the cloth step: CPU restatement (numpy).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
scene-step entry points.  Nothing in the product imports it.

  rigid_integrate  Rigid::integrate, Rigid.cpp:79-112
  box_step         Box::step, Box.cpp:103-114
  update_fix       Cloth::updateFix, Cloth.cpp:415-420
  fixed_rows       the addFixed rows of Constraints::fill, Constraints.cpp:114-119, 472-496
  scene_step       one Scene::step (Scene.cpp:72-98; remesh and EOL off, solver "none")
                   on the FP64 oracle's forces and PCG

Pinned in tests/test_scene_oracle.py: rigid_integrate against scipy's matrix
exponential of the twist (the definition of the rigid-body update), the
small-rotation branch against the pure translation, box_step against the
world-frame motion it must produce (a world-fixed rotation axis and velocity),
and update_fix / fixed_rows against hand-worked schedules.
"""
import numpy as np


def rigid_integrate(E0, phi, h):
    """E0 (4x4) times exp(h phi^) for the body twist phi = (w, v), written as
    Rigid.cpp:79-112 computes it: Rodrigues for the unit axis w/|w| over the angle
    |w| h, translation (w.v)|w|h w + (I - R)(w x v) with w, v divided by |w|; a
    plain translation h v when |w| <= 1e-10."""
    E0 = np.asarray(E0, dtype=np.float64).reshape(4, 4)
    w = np.array(phi[:3], dtype=np.float64)
    v = np.array(phi[3:6], dtype=np.float64)
    T = np.eye(4)
    T[:3, 3] = h * v
    wlen = np.linalg.norm(w)
    if wlen > 1e-10:
        w = w / wlen
        v = v / wlen
        c, s = np.cos(wlen * h), np.sin(wlen * h)
        c1 = 1.0 - c
        wX, wY, wZ = w
        R = np.array([[c + wX * wX * c1, -wZ * s + wX * wY * c1, wY * s + wX * wZ * c1],
                      [wZ * s + wX * wY * c1, c + wY * wY * c1, -wX * s + wY * wZ * c1],
                      [-wY * s + wX * wZ * c1, wX * s + wY * wZ * c1, c + wZ * wZ * c1]])
        d = (np.eye(3) - R) @ np.cross(w, v)
        T[:3, :3] = R
        T[:3, 3] = (w @ v) * wlen * h * w + d
    return E0 @ T


def box_step(E, vel, h):
    """Box::step: the twist (angular, linear) = Box::v mapped into the body frame by
    blockdiag(R^T, R^T), then rigid_integrate."""
    E = np.asarray(E, dtype=np.float64).reshape(4, 4)
    Rt = E[:3, :3].T
    phi = np.concatenate([Rt @ np.asarray(vel[:3], dtype=np.float64), Rt @ np.asarray(vel[3:6], dtype=np.float64)])
    return rigid_integrate(E, phi, h)


def update_fix(t, whens, fsindex):
    """Cloth::updateFix: the next list becomes current once t exceeds its `when`
    (one switch per call)."""
    if fsindex + 1 >= len(whens):
        return fsindex
    return fsindex + 1 if t > whens[fsindex + 1] else fsindex


def fixed_rows(corners, spec):
    """addFixed rows of one FixedList: corners (4 node indices), spec (4 x 6; a row
    starting with -1 is unused).  Returns (dofs, scripted velocities) in
    Constraints::fill's order (corner 1..4, axis x, y, z)."""
    dofs, vel = [], []
    for k in range(4):
        c = spec[k]
        if c[0] == -1:
            continue
        for a in range(3):
            if c[a] == 1.0:
                dofs.append(3 * corners[k] + a)
                vel.append(c[a + 3])
    return np.array(dofs, dtype=np.int32), np.array(vel, dtype=np.float64)


def scene_step(O, x, v, X, tri, hinge, mat, grav, h, dofs, scripted, tol):
    """One cloth step with single-DOF fixed rows on the FP64 oracle (module O):
    rhs = M v + h f, v[dof] = 0.99 v_old[dof] + scripted, PCG warm-started from
    v_old, x += h v.  x and v are updated in place."""
    rc, f, M, MDK = O.forces_fill(x, X, tri, hinge, mat, grav, h)
    assert rc == 0
    rhs = M @ v + h * f
    vals = (1 - 0.01) * v[dofs] + scripted
    vn, st = O.velocity_solve(MDK, rhs, dofs, vals, v0=v, tol=tol)
    assert st.status == 0
    v[:] = vn
    x += h * v
    return st
