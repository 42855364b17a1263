"""Multilevel preconditioner on the GPU against its FP64 oracle (ml_oracle).

Bars: the hierarchy is the host rule's partition on every level; P r in FP64 within
1e-12 of the oracle's (relative to max |z|), with the translations alone and with
the linear fields on the levels >= 3, in the mixed modes within 1e-5 (FP32 fine
blocks and vectors, FP64 coarse levels); solves with the coarse levels reach the
FP64 true-residual tolerance with fewer iterations than block-Jacobi (and with the
linear fields fewer than with the translations), agree with the block-Jacobi
solution to 1e-6 relative, and the FP64 iteration counts are within 3% of the
oracle's own PCG with the same preconditioner."""
import numpy as np
import pytest

import eolc
import ml_oracle as ML
from conftest import GRAV, H, rel_err
from eolc import learned, scene

pytestmark = pytest.mark.gpu
MLF = eolc.SOLVE_MULTILEVEL


def setup(ctx, sc):
    ctx.set_topology(sc.X, sc.tri, sc.hinge)
    ctx.set_material(0.1, 500.0, 0.3, 1e-5, 0.0, 1.0)
    ctx.set_options(eolc.OPT_SPD_PROJECT_BENDING)
    ctx.upload_state(sc.x0, sc.v0)
    ctx.forces_fill(GRAV, H)


def pattern(A, N):
    import scipy.sparse as sp
    c = A.tocoo()
    B = sp.csr_matrix((np.ones(c.nnz), (c.row // 3, c.col // 3)), shape=(N, N))
    B.sum_duplicates()
    B.sort_indices()
    return B.indptr.astype(np.int32), B.indices.astype(np.int32)


@pytest.mark.parametrize("n", [20, 33])
def test_hierarchy_and_apply_match_oracle(ctx, n):
    sc = scene.Scene(n, sigma_x=1e-3, sigma_v=0.05)
    print(f"multilevel: hierarchy and P r, {n}x{n}", flush=True)
    setup(ctx, sc)
    A = ctx.download_csr(eolc.MAT_MDK).tocsr()
    rp, cl = pattern(A, sc.N)
    ref = ML.hierarchy(rp, cl)
    counts = ctx.ml_info()
    print(f"  aggregates per level {counts}", flush=True)
    assert counts[0] == sc.N and len(counts) == len(ref) + 1
    maps = [ctx.ml_aggregates(l) for l in range(1, len(counts))]
    for a, b in zip(maps, ref):
        assert ML.same_partition(a, b)
    fd = learned.fixed_dofs(sc.fixed)
    rng = np.random.default_rng(n)
    r = rng.normal(size=3 * sc.N)
    assert len(counts) >= 4   # levels >= 3 exist: the linear fields are exercised
    zlin = ML.apply(A, sc.N, maps, r, fd, X=sc.X)       # linear fields on the levels >= 3
    ztr = ML.apply(A, sc.N, maps, r, fd)                 # translations only
    # eolc_set_ml_linear 0: translations alone; 1 (the default): linear fields in
    # FP32 iterations only; 2: in every mode
    z32 = ctx.precond_apply(r, fd, mode=eolc.SOLVE_MIXED_BJ | MLF)
    z64m = ctx.precond_apply(r, fd, mode=eolc.SOLVE_FP64_BJ | MLF)
    ctx.set_ml_linear(0)
    z64 = ctx.precond_apply(r, fd, mode=eolc.SOLVE_FP64_BJ | MLF)
    z32t = ctx.precond_apply(r, fd, mode=eolc.SOLVE_MIXED_BJ | MLF)
    ctx.set_ml_linear(2)
    z64l = ctx.precond_apply(r, fd, mode=eolc.SOLVE_FP64_BJ | MLF)
    ctx.set_ml_linear(1)
    zbj = ctx.precond_apply(r, fd, mode=eolc.SOLVE_FP64_BJ)
    ebj = rel_err(zbj, ML.apply(A, sc.N, [], r, fd))
    e64, e64l, e32, e32t = rel_err(z64, ztr), rel_err(z64l, zlin), rel_err(z32, zlin), rel_err(z32t, ztr)
    print(f"  rel err FP64 translations {e64:.2e}, FP64 linear {e64l:.2e}, mixed linear {e32:.2e}, "
          f"mixed translations {e32t:.2e}, block-Jacobi only {ebj:.2e}", flush=True)
    assert e64 <= 1e-12 and e64l <= 1e-12 and ebj <= 1e-12 and e32 <= 1e-5 and e32t <= 1e-5
    assert np.array_equal(z64m, z64)   # the default keeps the translations in FP64
    for z in (z64, z64l, z32, z32t):
        assert np.all(z[fd] == 0.0)
    # the coarse levels change z (they are not a no-op), and so do the linear fields
    assert rel_err(zbj, ztr) > 1e-3 and rel_err(zlin, ztr) > 1e-3


def test_multilevel_solves(ctx):
    n = 48
    sc = scene.Scene(n, sigma_x=1e-3, sigma_v=0.05)
    print(f"multilevel: solves {n}x{n}", flush=True)
    setup(ctx, sc)
    fd = learned.fixed_dofs(sc.fixed)
    vals = 0.99 * sc.v0[fd]
    out = {}
    for name, mode, tol in (("fp64-bj", eolc.SOLVE_FP64_BJ, 1e-10), ("fp64-ml", eolc.SOLVE_FP64_BJ | MLF, 1e-10),
                            ("mixed-ml", eolc.SOLVE_MIXED_BJ | MLF, 1e-10), ("fp64-ml-lin", eolc.SOLVE_FP64_BJ | MLF, 1e-10),
                            ("mixed-ml-lin", eolc.SOLVE_MIXED_BJ | MLF, 1e-10)):
        ctx.set_ml_linear({"fp64-ml-lin": 2, "mixed-ml-lin": 1}.get(name, 0))
        ctx.upload_state(sc.x0, sc.v0)
        ctx.forces_fill(GRAV, H)
        rc, st = ctx.velocity_solve(fd, vals, tol=tol, mode=mode)
        _, v = ctx.download_state()
        out[name] = (st, v)
        print(f"  {name}: {st.iters} iterations, true relres {st.true_relres:.2e}", flush=True)
        assert rc == eolc.OK and st.true_relres <= tol
    ctx.set_ml_linear(1)
    with pytest.raises(eolc.EolcError):
        ctx.set_ml_linear(3)
    assert out["fp64-ml"][0].iters < 0.7 * out["fp64-bj"][0].iters
    assert out["fp64-ml-lin"][0].iters < out["fp64-ml"][0].iters
    assert out["mixed-ml-lin"][0].iters < out["mixed-ml"][0].iters
    vbj = out["fp64-bj"][1]
    for k in ("fp64-ml", "mixed-ml", "fp64-ml-lin", "mixed-ml-lin"):
        assert rel_err(out[k][1], vbj) <= 1e-6
    # the oracle's PCG with the same preconditioner
    A = ctx.download_csr(eolc.MAT_MDK).tocsr()
    M = ctx.download_csr(eolc.MAT_M).tocsr()
    f = ctx.download_forces()
    b = M @ sc.v0 + H * f
    maps = [ctx.ml_aggregates(l) for l in range(1, len(ctx.ml_info()))]
    D0 = ML.block_jacobi(A, sc.N, fd)
    x0 = sc.v0.copy()
    x0[fd] = vals
    for name, X in (("fp64-ml", None), ("fp64-ml-lin", sc.X)):
        levels = ML.coarse_levels(A, sc.N, maps, fd, X)
        k_ref, _ = ML.pcg(A, b, x0, lambda r: ML.apply(A, sc.N, maps, r, fd, D0=D0, levels=levels), fd, tol=1e-10)
        print(f"  oracle PCG with the preconditioner of {name}: {k_ref} iterations", flush=True)
        assert abs(out[name][0].iters - k_ref) <= max(2, 0.03 * k_ref)


def test_learned_with_multilevel(ctx):
    """Zero-output learned blocks equal block-Jacobi to FP32 rounding, so the learned
    mode with the coarse levels converges like mixed block-Jacobi with them."""
    sc = scene.Scene(32, sigma_x=1e-3, sigma_v=0.05)
    print("multilevel: learned blocks + coarse levels 32x32", flush=True)
    setup(ctx, sc)
    fd = learned.fixed_dofs(sc.fixed)
    vals = 0.99 * sc.v0[fd]
    ctx.learned_set_weights(learned.init_weights(3))
    res = []
    for mode in (eolc.SOLVE_MIXED_BJ | MLF, eolc.SOLVE_MIXED_LEARNED | MLF):
        ctx.upload_state(sc.x0, sc.v0)
        ctx.forces_fill(GRAV, H)
        rc, st = ctx.velocity_solve(fd, vals, tol=1e-9, mode=mode)
        assert rc == eolc.OK and st.true_relres <= 1e-9 and st.used_fallback == 0
        res.append(st.iters)
    print(f"  iterations mixed-bj+ml {res[0]}, learned+ml {res[1]}", flush=True)
    assert abs(res[0] - res[1]) <= max(2, 0.05 * res[0])


def test_multilevel_step_through_driver(ctx):
    """eolc_step with the flag: positions after 3 steps equal the FP64 block-Jacobi
    steps' to the solve tolerance."""
    sc = scene.Scene(24, sigma_x=1e-3, sigma_v=0.05)
    print("multilevel: 3 steps through eolc_step 24x24", flush=True)
    setup(ctx, sc)
    xs = []
    for mode in (eolc.SOLVE_FP64_BJ, eolc.SOLVE_FP64_BJ | MLF, eolc.SOLVE_MIXED_LEARNED | MLF):
        ctx.learned_set_weights(learned.init_weights(3))
        ctx.upload_state(sc.x0, sc.v0)
        for _ in range(3):
            rc, st = ctx.step(GRAV, H, sc.fixed, tol=1e-10, mode=mode)
            assert rc == eolc.OK
        xs.append(ctx.download_state()[0])
    for x in xs[1:]:
        assert np.max(np.abs(x - xs[0])) <= 1e-9


def two_sheets(n):
    """Two disconnected n x n sheets in one topology (the second shifted by 2 m)."""
    a = scene.Scene(n, sigma_x=1e-3, sigma_v=0.05)
    b = scene.Scene(n, sigma_x=1e-3, sigma_v=0.05, seed_x=77, seed_v=78)
    N = a.N
    X = np.vstack([a.X, b.X + [2.0, 0.0]])
    tri = np.vstack([a.tri, b.tri + N])
    hinge = np.vstack([a.hinge, b.hinge + N])
    x0 = np.concatenate([a.x0, b.x0 + np.tile([2.0, 0.0, 0.0], N)])
    v0 = np.concatenate([a.v0, b.v0])
    fixed = np.concatenate([a.fixed, b.fixed + N])
    return X, tri, hinge, x0, v0, fixed


@pytest.mark.parametrize("case", ["partial-fixing", "two-sheets", "tiny"])
def test_multilevel_edge_cases(ctx, case):
    """Component-wise fixed DOFs, a topology of two disconnected sheets, and a mesh
    small enough for a single level: P r against the oracle (1e-12) and a solve to
    1e-10 that agrees with block-Jacobi's to 1e-6."""
    print(f"multilevel edge case: {case}", flush=True)
    if case == "two-sheets":
        X, tri, hinge, x0, v0, fixed = two_sheets(14)
    else:
        sc = scene.Scene(18 if case == "partial-fixing" else 2, sigma_x=1e-3, sigma_v=0.05)
        X, tri, hinge, x0, v0, fixed = sc.X, sc.tri, sc.hinge, sc.x0, sc.v0, sc.fixed
    N = len(X)
    ctx.set_topology(X, tri, hinge)
    ctx.set_material(0.1, 500.0, 0.3, 1e-5, 0.0, 1.0)
    ctx.set_options(eolc.OPT_SPD_PROJECT_BENDING)
    ctx.upload_state(x0, v0)
    ctx.forces_fill(GRAV, H)
    fd = learned.fixed_dofs(fixed)
    if case == "partial-fixing":
        # z only of a column of vertices, x and y of another
        fd = np.concatenate([fd, 3 * np.arange(0, N, 19) + 2, 3 * np.arange(5, N, 23), 3 * np.arange(5, N, 23) + 1])
        fd = np.unique(fd)
    counts = ctx.ml_info()
    print(f"  aggregates per level {counts}", flush=True)
    A = ctx.download_csr(eolc.MAT_MDK).tocsr()
    maps = [ctx.ml_aggregates(l) for l in range(1, len(counts))]
    if case == "two-sheets":
        assert counts[-1] == 2 and not set(maps[-1][:N // 2]) & set(maps[-1][N // 2:])
    if case == "tiny":
        assert len(counts) == 2
    rng = np.random.default_rng(3)
    r = rng.normal(size=3 * N)
    ctx.set_ml_linear(2)
    z = ctx.precond_apply(r, fd, mode=eolc.SOLVE_FP64_BJ | MLF)
    ctx.set_ml_linear(1)
    zref = ML.apply(A, N, maps, r, fd, X=X)
    print(f"  P r rel err {rel_err(z, zref):.2e}", flush=True)
    assert rel_err(z, zref) <= 1e-12 and np.all(z[fd] == 0.0)
    vals = 0.99 * v0[fd]
    out = []
    for mode in (eolc.SOLVE_FP64_BJ, eolc.SOLVE_FP64_BJ | MLF, eolc.SOLVE_MIXED_BJ | MLF):
        ctx.upload_state(x0, v0)
        ctx.forces_fill(GRAV, H)
        rc, st = ctx.velocity_solve(fd, vals, tol=1e-10, mode=mode)
        assert rc == eolc.OK and st.true_relres <= 1e-10
        out.append(ctx.download_state()[1])
        print(f"  mode {mode}: {st.iters} iterations", flush=True)
    for v in out[1:]:
        assert rel_err(v, out[0]) <= 1e-6
        assert np.array_equal(v[fd], out[0][fd])
