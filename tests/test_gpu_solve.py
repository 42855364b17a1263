"""Velocity solve and step parity against the CPU oracle.

Solved velocities: both sides solved to 1e-13 relative residual (or dense
elimination), compared to 1e-10 relative.  Trajectories: 20 steps of the 20x20
configuration from the same start within 1e-10; 200 steps pinned segment by segment
(the GPU restarts from the oracle's state every 10 steps, 1e-10 per segment); the
benchmark's mixed path over 50 steps at 32x32, segment by segment at the bar its
solver tolerance sets."""
import numpy as np
import pytest

import eolc
import oracle as O
from conftest import GRAV, H, rel_err
from eolc import learned, scene

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


def segment_parity(ctx, sc, steps, seg, mode, tol_gpu, tol_oracle, mat):
    """Segment-wise trajectory parity: the oracle steps the whole horizon; at the
    start of every segment of `seg` steps its (x, v) is uploaded to the GPU, both
    sides step `seg` steps, and the GPU's end-of-segment state is compared with the
    oracle's.  The oracle also re-runs each segment at a 10x tighter tolerance, so
    every segment reports its own sensitivity to the solver tolerance.  Returns a
    list of (x err, v err, x sensitivity, v sensitivity) per segment."""
    x, v = sc.x0.copy(), sc.v0.copy()
    out = []
    for s0 in range(0, steps, seg):
        ctx.upload_state(x, v)
        xt, vt = x.copy(), v.copy()
        for k in range(seg):
            rc, st = ctx.step(GRAV, H, sc.fixed, tol=tol_gpu, maxit=20000, mode=mode)
            assert rc == 0 and st.true_relres <= tol_gpu
            rco, sto, _ = O.step(x, v, sc.X, sc.tri, sc.hinge, mat, GRAV, H, sc.fixed, tol=tol_oracle, maxit=20000)
            assert rco == 0
            rco, sto, _ = O.step(xt, vt, sc.X, sc.tri, sc.hinge, mat, GRAV, H, sc.fixed, tol=tol_oracle / 10,
                                 maxit=20000)
            assert rco == 0
        xg, vg = ctx.download_state()
        out.append((rel_err(xg, x), rel_err(vg, v), rel_err(xt, x), rel_err(vt, v)))
    return out


def test_trajectory_200_steps(ctx):
    """200 steps of the 20x20 configuration (FP64 block-Jacobi, both sides at
    1e-12), pinned segment by segment: every 10 steps the GPU restarts from the
    oracle's state, and after each segment positions agree to 1e-10 and velocities to
    1e-8 (relative to the largest entry).  The sheet buckles after about 40 steps and
    amplifies differences over long horizons (the oracle at 1e-12 and 1e-13 ends 3e-2
    apart after 200 steps), so whole-horizon agreement is not a meaningful bar; per
    segment it is."""
    sc = scene.Scene(20)
    setup(ctx, sc)
    res = segment_parity(ctx, sc, 200, 10, eolc.SOLVE_FP64_BJ, 1e-12, 1e-12, O.material())
    for k, (ex, ev, sx, sv) in enumerate(res):
        print(f"  segment {k}: x err {ex:.2e} v err {ev:.2e} (oracle sensitivity {sx:.2e} / {sv:.2e})", flush=True)
    assert max(r[0] for r in res) <= 1e-10
    assert max(r[1] for r in res) <= 1e-8


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


FP64_DIFF_VARIANTS = (18, 22)   # FP64 difference form on symmetric records
PATCH_VARIANTS = (25, 26, 27, 28)   # patch layout (its dot products group rows by patch)


def test_spmv_variants_bitwise_identical_and_autotune(ctx):
    """Every SpMV launch variant of one arithmetic family gives the same bits (same
    arithmetic, same order; the LDS-staged variants 19-24 form the same p_new values
    in LDS that the others form per gather): all FP32 variants but the patch ones,
    the FP64 variants of the assembled form, and the FP64 variants of the difference
    form on symmetric records (18, 22: other rounding), whose solve meets the
    tolerance and agrees with the assembled form to the solver tolerance.  The patch
    variants 25-28 form every row's product from the same terms in the same order,
    but each workgroup's partial sum of p.q covers its patch instead of 1,024
    consecutive rows, so the dot products round differently: they agree with each
    other bit for bit and with the others to the solver tolerance.  The autotuner
    times them all and keeps one."""
    sc = scene.Scene(40, sigma_x=1e-3, sigma_v=0.05)
    for prec, mode in ((32, eolc.SOLVE_MIXED_BJ), (64, eolc.SOLVE_FP64_BJ)):
        out = []
        for var in range(eolc.SPMV_VARIANTS):
            print(f"spmv variant {var} precision {prec}", flush=True)
            setup(ctx, sc)
            ctx.set_spmv_variant(prec, var)
            rc, st = ctx.step(GRAV, H, sc.fixed, tol=1e-9, maxit=20000, mode=mode)
            assert rc == 0 and st.true_relres <= 1e-9
            out.append((ctx.download_state(), st.iters))
        special = (FP64_DIFF_VARIANTS + PATCH_VARIANTS) if prec == 64 else PATCH_VARIANTS
        families = [[o for k, o in enumerate(out) if k not in special]]
        if prec == 64:
            families += [[out[k] for k in FP64_DIFF_VARIANTS], [out[k] for k in PATCH_VARIANTS]]
        else:
            families.append([out[k] for k in PATCH_VARIANTS])
        for family in families:
            for (x, v), it in family[1:]:
                assert it == family[0][1]
                assert np.array_equal(x, family[0][0][0]) and np.array_equal(v, family[0][0][1])
        for k in (FP64_DIFF_VARIANTS + PATCH_VARIANTS) if prec == 64 else PATCH_VARIANTS:
            (xk, vk), itk = out[k]
            print(f"  variant {k}: {itk} iterations (variant 0: {out[0][1]}), v rel diff "
                  f"{rel_err(vk, out[0][0][1]):.2e}", flush=True)
            assert rel_err(vk, out[0][0][1]) <= 1e-5
        ctx.set_spmv_variant(prec, 0)
        setup(ctx, sc)
        ctx.forces_fill(GRAV, H)
        best, ms = ctx.autotune(prec, reps=5)
        print(f"autotune fp{prec}: best {best}, ms {np.round(ms * 1e3, 1)} us", flush=True)
        assert 0 <= best < eolc.SPMV_VARIANTS and np.all(ms > 0) and ms[best] == ms.min()
        ctx.set_spmv_variant(prec, 0)


def test_headline_path_50_steps_vs_oracle(ctx):
    """The benchmark's path (mixed precision, learned fine blocks, multilevel levels)
    for 50 steps of a 32x32 sheet with the benchmark's noise, pinned segment by
    segment against the oracle's FP64 block-Jacobi (both solved to 1e-10): every 10
    steps the GPU restarts from the oracle's state.  Bar per segment: positions and
    velocities within 20x the oracle's own change when its tolerance is 10x tighter
    over that segment, with floors 1e-9 (x) and 1e-7 (v) -- the bar the solver
    tolerance sets, since mixed and FP64 solves stop at different iterates of residual
    <= 1e-10.  Also: the GPU's FP64 forces at the final state within 1e-10."""
    from eolc import learned
    sc = scene.Scene(32, sigma_x=1e-3, sigma_v=0.05)
    print("headline path: 50 steps 32x32 vs oracle, segments of 10", flush=True)
    mat = O.material()
    setup(ctx, sc)
    ctx.learned_set_weights(learned.init_weights(7))
    head = eolc.SOLVE_MIXED_LEARNED | eolc.SOLVE_MULTILEVEL
    res = segment_parity(ctx, sc, 50, 10, head, 1e-10, 1e-10, mat)
    for k, (ex, ev, sx, sv) in enumerate(res):
        print(f"  segment {k}: x err {ex:.2e} v err {ev:.2e} (oracle sensitivity {sx:.2e} / {sv:.2e})", flush=True)
        assert ex <= max(20 * sx, 1e-9) and ev <= max(20 * sv, 1e-7)
    xg, _ = ctx.download_state()
    ctx.forces_fill(GRAV, H)
    fg = ctx.download_forces()
    rcf, fo, _, _ = O.forces_fill(xg, sc.X, sc.tri, sc.hinge, mat, GRAV, H)
    ef = rel_err(fg, fo)
    print(f"  forces at the final state {ef:.2e}", flush=True)
    assert ef <= 1e-10

def test_mixed_stall_continues_in_fp64():
    """A mixed solve whose FP32 iteration stalls finishes in FP64 (from its iterate, or
    from its initial guess after linear coarse fields) and meets the tolerance
    (stats.used_fallback bit 1); the context's next mixed solves
    start in FP64 for the sticky count.  The stall is forced from the host: a
    reliable-update trigger of 1e-4 (eolc_set_mixed_delta) keeps the first FP32 cycle
    going past its first 16-iteration chunk, and a first-cycle limit of 1 iteration
    (eolc_set_mixed_stall) then declares it stalled; the FP64 answer is compared with
    a plain FP64 solve of the same system.  (Own contexts: both are context settings.)"""
    sc = scene.Scene(48, sigma_x=1e-3, sigma_v=0.05)
    for mode, lin in ((eolc.SOLVE_MIXED_BJ, 0), (eolc.SOLVE_MIXED_BJ | eolc.SOLVE_MULTILEVEL, 0),
                      (eolc.SOLVE_MIXED_LEARNED | eolc.SOLVE_MULTILEVEL, 0),
                      (eolc.SOLVE_MIXED_BJ | eolc.SOLVE_MULTILEVEL, 1)):
        with eolc.Context(0) as c:
            setup(c, sc, eolc.OPT_SPD_PROJECT_BENDING)
            c.set_ml_linear(lin)   # 1: linear coarse fields in the FP32 iterations
            c.learned_set_weights(learned.init_weights(7))
            c.set_mixed_delta(1e-4)
            c.set_mixed_stall(first_cycle=1, min_cycle=1, sticky=2)
            rc, st = c.step(GRAV, H, sc.fixed, tol=1e-10, maxit=100000, mode=mode)
            print(f"  mode {mode} (linear fields {lin}): forced stall -> {st.as_dict()}", flush=True)
            assert rc == eolc.OK and st.true_relres <= 1e-10 and (st.used_fallback & 2)
            x1, v1 = c.download_state()
            # the next two solves start in FP64 directly, the third in FP32 again
            c.set_mixed_delta(0.5)
            for k, fp64 in ((0, True), (1, True), (2, False)):
                c.set_mixed_stall(first_cycle=10000, min_cycle=1000, sticky=2 if fp64 else 0)
                rc, stk = c.step(GRAV, H, sc.fixed, tol=1e-10, maxit=100000, mode=mode)
                assert rc == eolc.OK and bool(stk.used_fallback & 2) == fp64, (k, stk.as_dict())
            # against a plain FP64 solve of the first step's system
            setup(c, sc, eolc.OPT_SPD_PROJECT_BENDING)
            rc, st64 = c.step(GRAV, H, sc.fixed, tol=1e-13, maxit=100000, mode=eolc.SOLVE_FP64_BJ)
            x64, v64 = c.download_state()
            assert rel_err(v1, v64) <= 1e-7
