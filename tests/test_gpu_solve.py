"""Velocity solve and step parity against the CPU oracle.

Solved velocities: both sides solved to 1e-13 relative residual (or dense
elimination), compared to 1e-10 relative.  Trajectory: 200 steps of the 20x20
configuration, positions within 1e-9 of the oracle's (relative to max |x|)."""
import numpy as np
import pytest

import eolc
import oracle as O
from conftest import GRAV, H, rel_err
from eolc import scene

pytestmark = pytest.mark.gpu


def setup(ctx, sc, flags=0):
    ctx.set_topology(sc.X, sc.tri, sc.hinge)
    ctx.set_material(0.1, 500.0, 0.3, 1e-5, 0.0, 1.0)
    ctx.set_options(flags)
    ctx.upload_state(sc.x0, sc.v0)


def fixed_dofs(sc, v):
    d = np.concatenate([3 * sc.fixed[:, None] + np.arange(3)[None, :]]).reshape(-1).astype(np.int32)
    return d, 0.99 * v[d]


@pytest.mark.parametrize("mode", [eolc.SOLVE_FP64_BJ, eolc.SOLVE_FP64_JACOBI])
def test_velocity_solve_vs_dense(ctx, mode):
    sc = scene.Scene(20, sigma_x=1e-3, sigma_v=0.05)
    print(f"solve vs dense mode={mode}", flush=True)
    setup(ctx, sc)
    ctx.forces_fill(GRAV, H)
    d, val = fixed_dofs(sc, sc.v0)
    rc, st = ctx.velocity_solve(d, val, tol=1e-13, maxit=50000, mode=mode)
    _, vg = ctx.download_state()
    _, fo, Mo, Ko = O.forces_fill(sc.x0, sc.X, sc.tri, sc.hinge, O.material(), GRAV, H)
    rhs = Mo @ sc.v0 + H * fo
    rcd, vd = O.dense_velocity_solve(Ko, rhs, d, val)
    assert rcd == 0
    assert st.true_relres < 1e-12
    assert rel_err(vg, vd) < 1e-10
    assert np.array_equal(vg[d], val)


def test_mixed_solve_reaches_tolerance(ctx):
    sc = scene.Scene(48, sigma_x=1e-3, sigma_v=0.05)
    setup(ctx, sc)
    ctx.forces_fill(GRAV, H)
    d, val = fixed_dofs(sc, sc.v0)
    rc, st = ctx.velocity_solve(d, val, tol=1e-8, maxit=50000, mode=eolc.SOLVE_MIXED_BJ)
    _, vm = ctx.download_state()
    print("mixed", st.as_dict(), flush=True)
    assert st.true_relres <= 1e-8 and st.refinements >= 1
    # every working iteration was enqueued; the host waits once per chunk and once
    # per reliable update
    assert st.launched >= st.iters and st.host_syncs >= st.refinements + 1
    ctx.upload_state(sc.x0, sc.v0)
    ctx.forces_fill(GRAV, H)
    rc, st64 = ctx.velocity_solve(d, val, tol=1e-13, maxit=50000, mode=eolc.SOLVE_FP64_BJ)
    _, v64 = ctx.download_state()
    assert rel_err(vm, v64) < 1e-5


def test_trajectory_20_steps(ctx):
    """Before the hanging sheet starts to buckle the trajectories agree tightly."""
    sc = scene.Scene(20)
    setup(ctx, sc)
    x, v = sc.x0.copy(), sc.v0.copy()
    mat = O.material()
    for k in range(20):
        rc, st = ctx.step(GRAV, H, sc.fixed, tol=1e-12, maxit=20000, mode=eolc.SOLVE_FP64_BJ)
        rco, sto, _ = O.step(x, v, sc.X, sc.tri, sc.hinge, mat, GRAV, H, sc.fixed, tol=1e-12, maxit=20000)
        assert rc == 0 and rco == 0
    xg, vg = ctx.download_state()
    print("20-step err x", rel_err(xg, x), "v", rel_err(vg, v), flush=True)
    assert rel_err(xg, x) < 1e-10
    assert rel_err(vg, v) < 1e-8


def test_trajectory_200_steps(ctx):
    """200 steps of the 20x20 configuration.  The sheet buckles after ~40 steps and
    amplifies rounding-level differences: the oracle run at solver tolerances 1e-12
    and 1e-13 itself ends ~3e-2 apart.  Stated tolerance: the GPU trajectory stays
    within 20x that self-sensitivity of the oracle's, the fixed corners stay put and
    every value is finite."""
    sc = scene.Scene(20)
    setup(ctx, sc)
    mat = O.material()
    runs = []
    for tol in (1e-12, 1e-13):
        x, v = sc.x0.copy(), sc.v0.copy()
        for k in range(200):
            rco, sto, _ = O.step(x, v, sc.X, sc.tri, sc.hinge, mat, GRAV, H, sc.fixed, tol=tol, maxit=20000)
            assert rco == 0
        runs.append(x)
    for k in range(200):
        rc, st = ctx.step(GRAV, H, sc.fixed, tol=1e-12, maxit=20000, mode=eolc.SOLVE_FP64_BJ)
        assert rc == 0
    xg, vg = ctx.download_state()
    envelope = rel_err(runs[1], runs[0])
    err = rel_err(xg, runs[0])
    print(f"200-step err x {err:.3e}, oracle self-sensitivity {envelope:.3e}", flush=True)
    assert np.all(np.isfinite(xg))
    d = (3 * sc.fixed[:, None] + np.arange(3)).reshape(-1)
    assert np.array_equal(xg[d], sc.x0[d])
    assert err <= max(20 * envelope, 1e-9)


def test_solve_deterministic(ctx):
    sc = scene.Scene(24, sigma_x=2e-3, sigma_v=0.05)
    out = []
    for _ in range(2):
        setup(ctx, sc)
        rc, st = ctx.step(GRAV, H, sc.fixed, tol=1e-9, maxit=20000, mode=eolc.SOLVE_FP64_BJ)
        out.append(ctx.download_state())
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_iteration_cap_reports_maxit(ctx):
    sc = scene.Scene(20, sigma_x=1e-3, sigma_v=0.05)
    setup(ctx, sc)
    rc, st = ctx.step(GRAV, H, sc.fixed, tol=1e-14, maxit=5, mode=eolc.SOLVE_FP64_BJ,
                      allow=(eolc.ERR_MAXIT,))
    assert rc == eolc.ERR_MAXIT and st.iters == 5
    assert st.launched == 5 and st.host_syncs >= 1


def test_spmv_variants_bitwise_identical_and_autotune(ctx):
    """Every SpMV launch variant gives the same bits (same arithmetic, same order);
    the autotuner times them all and keeps one."""
    sc = scene.Scene(40, sigma_x=1e-3, sigma_v=0.05)
    for prec, mode in ((32, eolc.SOLVE_MIXED_BJ), (64, eolc.SOLVE_FP64_BJ)):
        out = []
        for var in range(eolc.SPMV_VARIANTS):
            print(f"spmv variant {var} precision {prec}", flush=True)
            setup(ctx, sc)
            ctx.set_spmv_variant(prec, var)
            rc, st = ctx.step(GRAV, H, sc.fixed, tol=1e-9, maxit=20000, mode=mode)
            assert rc == 0
            out.append((ctx.download_state(), st.iters))
        for (x, v), it in out[1:]:
            assert it == out[0][1]
            assert np.array_equal(x, out[0][0][0]) and np.array_equal(v, out[0][0][1])
        setup(ctx, sc)
        ctx.forces_fill(GRAV, H)
        best, ms = ctx.autotune(prec, reps=5)
        print(f"autotune fp{prec}: best {best}, ms {np.round(ms * 1e3, 1)} us", flush=True)
        assert 0 <= best < eolc.SPMV_VARIANTS and np.all(ms > 0) and ms[best] == ms.min()


def test_headline_path_50_steps_vs_oracle(ctx):
    """The benchmark's path (mixed precision, learned fine blocks at their initial
    weights, multilevel levels) for 50 steps of a 32x32 sheet with the benchmark's
    noise, against 50 steps of the oracle's FP64 block-Jacobi, both solved to 1e-10.
    Reported: position and velocity error after 50 steps, and the force error of the
    GPU's FP64 evaluation at the final state.  Bars: positions and velocities within
    20x the oracle's own sensitivity to its solver tolerance (1e-10 vs 1e-11), with
    floors 1e-9 and 1e-7; forces within 1e-10."""
    from eolc import learned
    sc = scene.Scene(32, sigma_x=1e-3, sigma_v=0.05)
    print("headline path: 50 steps 32x32 vs oracle", flush=True)
    mat = O.material()
    runs = []
    for tol in (1e-10, 1e-11):
        x, v = sc.x0.copy(), sc.v0.copy()
        for k in range(50):
            rco, sto, _ = O.step(x, v, sc.X, sc.tri, sc.hinge, mat, GRAV, H, sc.fixed, tol=tol, maxit=20000)
            assert rco == 0
        runs.append((x, v))
    setup(ctx, sc)
    ctx.learned_set_weights(learned.init_weights(7))
    head = eolc.SOLVE_MIXED_LEARNED | eolc.SOLVE_MULTILEVEL
    for k in range(50):
        rc, st = ctx.step(GRAV, H, sc.fixed, tol=1e-10, maxit=20000, mode=head)
        assert rc == 0 and st.true_relres <= 1e-10
    xg, vg = ctx.download_state()
    (xo, vo), (x11, v11) = runs
    ex, ev = rel_err(xg, xo), rel_err(vg, vo)
    sx, sv = rel_err(x11, xo), rel_err(v11, vo)
    ctx.forces_fill(GRAV, H)
    fg = ctx.download_forces()
    rcf, fo, _, _ = O.forces_fill(xg, sc.X, sc.tri, sc.hinge, mat, GRAV, H)
    ef = rel_err(fg, fo)
    print(f"  50 steps: x err {ex:.2e} (oracle sensitivity {sx:.2e}), v err {ev:.2e} ({sv:.2e}), "
          f"forces at the final state {ef:.2e}", flush=True)
    assert ex <= max(20 * sx, 1e-9) and ev <= max(20 * sv, 1e-7)
    assert ef <= 1e-10
