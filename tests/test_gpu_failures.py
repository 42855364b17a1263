"""Failure paths that the reference leaves unguarded, forced from the host on small
meshes.  Every kernel still runs to its normal end (all loops have fixed trip
counts; no index depends on a value), and each case must end promptly with its
status code and leave the resident state as it was:

- a non-finite entry in an uploaded state (the reference has no NaN checks,
  ComputeMembrane.cpp:48, Forces.cpp:362-364): EOLC_ERR_NONFINITE at the upload;
- a face collapsed to a point in world space, whose P frame divides 0 by 0
  (Forces.cpp:359-364): the step reports EOLC_ERR_NONFINITE from the solve within a
  few iterations, in every solve mode;
- the exact (unprojected) bending Hessian on a strongly perturbed sheet, where
  M + h^2 K is indefinite (GeneralizedSolver.cpp:140-142 aborts instead of solving):
  EOLC_ERR_NOT_SPD within a few iterations, not a stagnating loop to maxit."""
import numpy as np
import pytest

import eolc
from conftest import GRAV, H
from eolc import learned, scene

pytestmark = pytest.mark.gpu

MODES = [("fp64-bj", eolc.SOLVE_FP64_BJ), ("mixed-bj", eolc.SOLVE_MIXED_BJ),
         ("fp64-jacobi", eolc.SOLVE_FP64_JACOBI),
         ("mixed-bj-ml", eolc.SOLVE_MIXED_BJ | eolc.SOLVE_MULTILEVEL),
         ("mixed-learned-ml", eolc.SOLVE_MIXED_LEARNED | eolc.SOLVE_MULTILEVEL)]


def setup(ctx, sc, beta=1e-5, options=0):
    ctx.set_topology(sc.X, sc.tri, sc.hinge)
    ctx.set_material(0.1, 500.0, 0.3, beta, 0.0, 1.0)
    ctx.set_options(options)
    ctx.learned_set_weights(learned.init_weights(5))


def test_nonfinite_upload_rejected_and_state_kept(ctx):
    sc = scene.Scene(10, sigma_x=1e-3, sigma_v=0.05)
    setup(ctx, sc)
    ctx.upload_state(sc.x0, sc.v0)
    bad = sc.x0.copy()
    bad[7] = np.nan
    rc = eolc.lib().eolc_upload_state(ctx._h, eolc._d(bad), eolc._d(sc.v0))
    assert rc == eolc.ERR_NONFINITE
    badv = sc.v0.copy()
    badv[3] = np.inf
    rc = eolc.lib().eolc_upload_state(ctx._h, eolc._d(sc.x0), eolc._d(badv))
    assert rc == eolc.ERR_NONFINITE
    x, v = ctx.download_state()
    assert np.array_equal(x, sc.x0) and np.array_equal(v, sc.v0)


@pytest.mark.parametrize("name,mode", MODES)
def test_collapsed_face_reports_nonfinite(ctx, name, mode):
    sc = scene.Scene(10, sigma_x=1e-3, sigma_v=0.05)
    setup(ctx, sc)
    x0 = sc.x0.copy().reshape(-1, 3)
    a, b, c = sc.tri[37]
    x0[b] = x0[a]
    x0[c] = x0[a]
    x0 = x0.reshape(-1)
    print(f"collapsed face 37, mode {name}", flush=True)
    ctx.upload_state(x0, sc.v0)
    rc, st = ctx.step(GRAV, H, sc.fixed, tol=1e-8, maxit=5000, mode=mode,
                      allow=(eolc.ERR_NONFINITE,))
    assert rc == eolc.ERR_NONFINITE and st.status == eolc.ERR_NONFINITE
    assert st.iters <= 64, st.as_dict()
    x, v = ctx.download_state()
    assert np.array_equal(x, x0) and np.array_equal(v, sc.v0)


@pytest.mark.parametrize("name,mode", MODES)
def test_indefinite_exact_bending_reports_not_spd(ctx, name, mode):
    sc = scene.Scene(32, sigma_x=1e-2, sigma_v=0.05)
    setup(ctx, sc, beta=1e-3, options=0)
    print(f"indefinite 32x32 (sigma 1e-2, beta 1e-3), mode {name}", flush=True)
    ctx.upload_state(sc.x0, sc.v0)
    rc, st = ctx.step(GRAV, H, sc.fixed, tol=1e-8, maxit=20000, mode=mode,
                      allow=(eolc.ERR_NOT_SPD,))
    assert rc == eolc.ERR_NOT_SPD, st.as_dict()
    assert st.iters < 2000, st.as_dict()
    x, v = ctx.download_state()
    assert np.array_equal(x, sc.x0) and np.array_equal(v, sc.v0)
    # the same sheet with the projected hinge Hessians solves
    ctx.set_options(eolc.OPT_SPD_PROJECT_BENDING)
    rc, st = ctx.step(GRAV, H, sc.fixed, tol=1e-8, maxit=20000, mode=mode)
    assert rc == eolc.OK and st.true_relres <= 1e-8
    ctx.set_options(0)
