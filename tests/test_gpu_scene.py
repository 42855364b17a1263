"""The scene step around the cloth step (Scene::step, Scene.cpp:72-98, remesh and EOL
off, solver "none"): FixedList schedules with scripted and per-axis corners
(Cloth::updateFix, Constraints.cpp:472-496), the obstacle motion (Box::step) and the
CD2 records handed to the host QP; and the scripted velocities of the batched
graph step.

Bars: positions after 20 steps within 1e-10 of the FP64 oracle run with the same
schedule (tolerance 1e-12 on both sides); box frames within 1e-13 of the oracle's
restatement; the final CD2 records bit-exact against the oracle's brute force on the
final positions and frame; batched instances with scripted velocities within the
batched path's bars of the single-cloth step, and graph replays after a schedule
switch bitwise equal to eager steps."""
import os
import subprocess

import numpy as np
import pytest

import eolc
import oracle as O
import scene_oracle as SO
from conftest import GRAV, H
from eolc import scene
from test_gpu_collide import oracle_cd, rec_tuple

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "eol-cloth_amd", "host", "cloth_step")


def schedule(n):
    """The two FixedLists of cloth_step's scene: the y = 1 corners fixed, then from
    t > 0.0475 corner A scripted at x = 0.1 m/s and corner B free but for z at
    0.05 m/s."""
    cA, cB = n * (n + 1), (n + 1) * (n + 1) - 1
    unused = [-1.0] * 6
    f0 = [[1, 1, 1, 0, 0, 0], [1, 1, 1, 0, 0, 0], unused, unused]
    f1 = [[1, 1, 1, 0.1, 0, 0], [0, 0, 1, 0, 0, 0.05], unused, unused]
    return [0.0, 0.0475], [f0, f1], [cA, cB, 0, 0]


def oracle_scene(n, steps, tol=1e-12):
    sc = scene.Scene(n)
    whens, specs, corners = schedule(n)
    mat = O.material(0.1, 500.0, 0.3, 1e-5, 0.0, 1.0)
    x, v = sc.x0.copy(), sc.v0.copy()
    E = np.eye(4)
    E[:3, 3] = [0.5, 0.45, -0.08]
    vel = np.array([0.0, 0.0, 2.0, 0.0, 0.0, 0.4])
    t, idx, switched = 0.0, 0, -1
    for k in range(steps):
        new = SO.update_fix(t, whens, idx)
        if new != idx:
            switched = k
        idx = new
        dofs, scr = SO.fixed_rows(corners, specs[idx])
        SO.scene_step(O, x, v, sc.X, sc.tri, sc.hinge, mat, GRAV, H, dofs, scr, tol)
        E = SO.box_step(E, vel, H)
        t += H
    return sc, x, v, E, switched


def test_step_dofs_schedule_matches_oracle(ctx):
    n, steps = 20, 20
    print(f"scene schedule through eolc_step_dofs, {n}x{n}, {steps} steps", flush=True)
    sc, xo, vo, _, switched = oracle_scene(n, steps)
    assert switched == 10
    whens, specs, corners = schedule(n)
    ctx.set_topology(sc.X, sc.tri, sc.hinge)
    ctx.set_material(0.1, 500.0, 0.3, 1e-5, 0.0, 1.0)
    ctx.set_options(0)   # the session's context may carry another test's options
    ctx.upload_state(sc.x0, sc.v0)
    t, idx = 0.0, 0
    for k in range(steps):
        idx = SO.update_fix(t, whens, idx)
        dofs, scr = SO.fixed_rows(corners, specs[idx])
        rc, st = ctx.step_dofs(GRAV, H, dofs, scr, tol=1e-12)
        assert rc == eolc.OK and st.true_relres <= 1e-12
        t += H
    x, v = ctx.download_state()
    err = np.max(np.abs(x - xo))
    print(f"  max |x - oracle| {err:.2e}", flush=True)
    assert err <= 1e-10
    # the scripted corner moved with its velocity, the z-only corner kept x and y free
    cA, cB = corners[0], corners[1]
    assert abs(v[3 * cA] - (0.99 ** 10 * 0 + 0.1 * sum(0.99 ** j for j in range(10)))) <= 1e-12
    assert abs(v[3 * cB + 2] - 0.05 * sum(0.99 ** j for j in range(10))) <= 1e-12
    assert abs(v[3 * cB]) > 1e-6 or abs(v[3 * cB + 1]) > 1e-6


def test_cpp_scene_driver(tmp_path):
    n, steps = 20, 20
    out = str(tmp_path / "scene.bin")
    print(f"C++ Scene driver: {n}x{n}, {steps} steps", flush=True)
    r = subprocess.run([EXE, str(n), str(steps), out, "scene"], capture_output=True, text=True, timeout=120)
    print("  " + r.stdout.strip(), flush=True)
    assert r.returncode == 0, r.stderr
    assert "switch_step=10" in r.stdout
    sc, xo, vo, Eo, _ = oracle_scene(n, steps)
    data = np.fromfile(out, dtype=np.float64)
    x, Ec = data[:3 * sc.N], data[3 * sc.N:].reshape(4, 4).T   # frame stored column-major
    assert np.max(np.abs(x - xo)) <= 1e-10
    assert np.max(np.abs(Ec - Eo)) <= 1e-13
    raw = open(out + ".cd", "rb").read()
    nrec = len(raw) // C_SIZE()
    recs = (eolc.CollisionRec * nrec).from_buffer_copy(raw) if nrec else []
    pts = np.array([[0.3, 0.3, -0.004], [0.7, 0.6, 0.003]])
    nrm = np.array([[0.0, 0.0, 1.0], [0.0, 0.0, -1.0]])
    o = oracle_cd(5e-3, pts, nrm, [np.array([0.4, 0.3, 0.1])], [Ec], x, sc.tri)
    print(f"  final CD2 records: driver {nrec}, oracle {len(o)}", flush=True)
    assert nrec == len(o) and nrec > 0
    for a, b in zip(recs, o):
        assert rec_tuple(a) == rec_tuple(b)


def C_SIZE():
    import ctypes
    return ctypes.sizeof(eolc.CollisionRec)


def test_batch_scripted_velocities():
    bs = scene.BatchScene(B=4, side_lo=6, side_hi=16, seed=91)
    rng = np.random.default_rng(3)
    nfix = [len(sc.fixed) for sc in bs.inst]
    scripted = [rng.normal(scale=0.05, size=3 * m) for m in nfix]
    print("batched step with scripted fixed nodes, sides", bs.sides, flush=True)
    with eolc.Context(0) as c:
        bs.setup(c)
        x0, v0 = c.download_state()
        c.set_batch_scripted(np.concatenate(scripted))
        rc, st = c.batch_step(GRAV, H, tol=1e-10, maxit=20000, eager=True)
        assert rc == eolc.OK
        eager = c.download_state()
        xs, vs = bs.unpack(eager[0]), bs.unpack(eager[1])
        for b, sc in enumerate(bs.inst):
            with eolc.Context(0) as s:
                s.set_topology(sc.X, sc.tri, sc.hinge)
                s.set_material(0.1, float(bs.E[b]), 0.3, float(bs.beta[b]), 0.0, 1.0)
                s.upload_state(sc.x0, sc.v0)
                s.step(GRAV, H, sc.fixed, scripted=scripted[b], tol=1e-10, maxit=20000, mode=eolc.SOLVE_MIXED_BJ)
                x1, v1 = s.download_state()
            fd = (3 * np.asarray(sc.fixed)[:, None] + np.arange(3)).reshape(-1)
            assert np.array_equal(vs[b][fd], v1[fd])   # the fixed values, bit for bit
            err = np.max(np.abs(vs[b] - v1)) / np.max(np.abs(v1))
            print(f"  instance {b}: v rel err vs single cloth {err:.2e}", flush=True)
            assert err < 1e-6
            assert np.max(np.abs(xs[b] - x1)) <= 1e-9 * np.max(np.abs(x1))
        # graph: captured with the first velocities, replayed after a switch to new
        # ones (no re-capture) equals the eager step with the new ones
        c.set_batch_scripted(np.concatenate(scripted))
        c.upload_state(x0, v0)
        c.batch_step(GRAV, H, tol=1e-10, maxit=20000)
        g1 = c.download_state()
        assert np.array_equal(g1[0], eager[0]) and np.array_equal(g1[1], eager[1])
        new = 2.0 * np.concatenate(scripted)
        c.set_batch_scripted(new)
        for replay in range(2):
            c.upload_state(x0, v0)
            c.download_state()
            c.batch_step(GRAV, H, tol=1e-10, maxit=20000)
            g = c.download_state()
            c.upload_state(x0, v0)
            c.batch_step(GRAV, H, tol=1e-10, maxit=20000, eager=True)
            e = c.download_state()
            print(f"  replay {replay} after the switch: bitwise equal to eager", flush=True)
            assert np.array_equal(g[0], e[0]) and np.array_equal(g[1], e[1])
        assert not np.array_equal(g[1], eager[1])
        c.set_batch_scripted(None)
