"""The patch-layout SpMV (variants 25-28) on meshes other than the benchmark grid.

The patch layout cuts the vertices into compact patches by coordinate bisection
of X, whatever their numbering, and passes in-patch transposed products through
LDS.  These tests solve on an irregular mesh (diagonals flipped at random, faces
removed, vertex numbering shuffled) with the patch variants and with the SELL
variants, in both precisions, and require the same answer to the solver
tolerance; and they drive the stiff-block path (blocks applied from their FP64
values in both rows) from the host by lowering the threshold until most blocks
qualify."""
import numpy as np
import pytest

import eolc
from conftest import GRAV, H, rel_err
from eolc import scene

pytestmark = pytest.mark.gpu


def irregular_sheet(n=48, seed=3, drop=0.04):
    """A grid with random diagonals and a few faces removed, the vertices renumbered
    at random (unused vertices dropped), noise 1e-3 on x and 0.05 on v."""
    rng = np.random.default_rng(seed)
    X, _ = scene.grid(n)
    np1 = n + 1
    tris = []
    for j in range(n):
        for i in range(n):
            v00 = j * np1 + i
            v10, v01, v11 = v00 + 1, v00 + np1, v00 + np1 + 1
            if rng.random() < 0.5:
                q = [(v00, v10, v11), (v00, v11, v01)]
            else:
                q = [(v00, v10, v01), (v10, v11, v01)]
            for t in q:
                if rng.random() >= drop or (j == n - 1):   # keep the top row (fixed corners)
                    tris.append(t)
    tri = np.array(tris, dtype=np.int64)
    used = np.unique(tri)
    perm = rng.permutation(len(used))
    new_of = np.full(X.shape[0], -1, dtype=np.int64)
    new_of[used] = perm
    Xn = np.zeros((len(used), 2))
    Xn[perm] = X[used]
    tri = new_of[tri].astype(np.int32)
    hinge = scene.hinges(len(used), tri)
    N = len(used)
    x0 = np.zeros((N, 3))
    x0[:, :2] = Xn
    x0 = x0.reshape(-1) + scene.gaussian(11, 3 * N, 1e-3)
    v0 = scene.gaussian(12, 3 * N, 0.05)
    corners = np.array([new_of[n * np1], new_of[np1 * np1 - 1]], dtype=np.int32)
    return Xn, tri, hinge, x0, v0, corners


def solve(ctx, mesh, prec, variant, rel=None):
    Xn, tri, hinge, x0, v0, fixed = mesh
    ctx.set_topology(Xn, tri, hinge)
    ctx.set_material(0.1, 500.0, 0.3, 1e-5, 0.0, 1.0)
    ctx.set_options(eolc.OPT_SPD_PROJECT_BENDING)
    if rel is not None:
        ctx.set_stiff_threshold(rel)
    ctx.set_spmv_variant(prec, variant)
    ctx.upload_state(x0, v0)
    mode = eolc.SOLVE_MIXED_BJ if prec == 32 else eolc.SOLVE_FP64_BJ
    rc, st = ctx.step(GRAV, H, fixed, tol=1e-9, maxit=50000, mode=mode | eolc.SOLVE_MULTILEVEL)
    assert rc == 0 and st.true_relres <= 1e-9, st.as_dict()
    return ctx.download_state()[1], st


@pytest.mark.parametrize("prec", [32, 64])
def test_patch_variants_on_an_irregular_renumbered_mesh(ctx, prec):
    mesh = irregular_sheet()
    ctx.set_topology(mesh[0], mesh[1], mesh[2])
    info = ctx.spmv_layout_info()
    print(f"irregular mesh: {len(mesh[0])} vertices, {len(mesh[1])} faces, layout {info}", flush=True)
    assert info["applies"] == 1 and 0 < info["cross_blocks"] < info["records"]
    ref, st0 = solve(ctx, mesh, prec, 0)
    out = {}
    for var in (25, 26, 27, 28):
        out[var], st = solve(ctx, mesh, prec, var)
        print(f"  fp{prec} variant {var}: {st.iters} iterations (variant 0: {st0.iters}), "
              f"v rel diff {rel_err(out[var], ref):.2e}", flush=True)
        assert rel_err(out[var], ref) <= 1e-6
    for var in (26, 27, 28):   # launch shapes of one layout: the same bits
        assert np.array_equal(out[var], out[25])
    ctx.set_spmv_variant(prec, 0)


def test_stiff_blocks_applied_in_fp64(ctx):
    """Thresholds that send none, some and nearly all blocks to the FP64 path give
    the same mixed solve to the tolerance (the path runs in both rows of a block).
    The answer alone would not expose a wrong FP64 copy of a block: the mixed solve's
    FP64 true residual corrects an inconsistent FP32 operator, at the cost of more
    refinements or a fallback.  So the iteration and refinement counts must also
    match the run without stiff blocks (the operators differ by FP32 rounding only),
    and no run may fall back to FP64."""
    mesh = irregular_sheet(n=40, seed=5)
    ref, _ = solve(ctx, mesh, 64, 0)
    base = None
    for rel in (0.0, 1e-1, 1e-4, 1e-8):
        v, st = solve(ctx, mesh, 32, 25, rel=rel)
        print(f"  stiff threshold {rel:g}: {st.iters} iterations, {st.refinements} refinements, "
              f"fallback {st.used_fallback}, v rel diff vs FP64 {rel_err(v, ref):.2e}", flush=True)
        assert rel_err(v, ref) <= 1e-6
        assert st.used_fallback == 0
        if base is None:   # threshold 0: no stiff blocks
            base = (st.iters, st.refinements)
        else:
            assert abs(st.iters - base[0]) <= 0.1 * base[0] + 5, (st.iters, base)
            assert st.refinements <= base[1] + 3, (st.refinements, base)
    ctx.set_stiff_threshold(10.0)
    ctx.set_spmv_variant(32, 0)


def test_patch_layout_refused_falls_back(ctx):
    """A fan of 72 triangles around one vertex gives that row more than 64
    off-diagonal blocks: the patch layout does not apply, the patch variants are
    refused by name and reported infinite by the autotuner, and a context that had
    chosen one solves with its symmetric-record twin (variant 18, the same
    products)."""
    K = 72
    ang = 2 * np.pi * np.arange(K) / K
    X = np.concatenate([[[0.5, 0.5]], 0.5 + 0.4 * np.stack([np.cos(ang), np.sin(ang)], 1)])
    tri = np.array([[0, 1 + k, 1 + (k + 1) % K] for k in range(K)], dtype=np.int32)
    hinge = scene.hinges(len(X), tri)
    x0 = np.zeros((len(X), 3))
    x0[:, :2] = X
    x0 = x0.reshape(-1) + scene.gaussian(21, 3 * len(X), 1e-3)
    v0 = scene.gaussian(22, 3 * len(X), 0.05)
    fixed = np.array([1, 1 + K // 2], dtype=np.int32)
    # a patch variant chosen on a grid first
    sc = scene.Scene(16)
    ctx.set_topology(sc.X, sc.tri, sc.hinge)
    ctx.set_spmv_variant(32, 25)
    ctx.set_spmv_variant(64, 25)
    ctx.set_topology(X, tri, hinge)
    info = ctx.spmv_layout_info()
    print(f"fan of {K} triangles: layout {info}", flush=True)
    assert info["applies"] == 0
    with pytest.raises(eolc.EolcError):
        ctx.set_spmv_variant(32, 25)
    ctx.set_material(0.1, 500.0, 0.3, 1e-5, 0.0, 1.0)
    out = {}
    for prec, mode in ((32, eolc.SOLVE_MIXED_BJ), (64, eolc.SOLVE_FP64_BJ)):
        ctx.upload_state(x0, v0)
        rc, st = ctx.step(GRAV, H, fixed, tol=1e-10, maxit=20000, mode=mode)
        assert rc == 0 and st.true_relres <= 1e-10
        out[prec] = ctx.download_state()[1]
        ctx.upload_state(x0, v0)
        ctx.forces_fill(GRAV, H)
        best, ms = ctx.autotune(prec, reps=2)
        assert all(np.isinf(ms[v]) for v in (25, 26, 27, 28)) and best not in (25, 26, 27, 28)
    assert rel_err(out[32], out[64]) <= 1e-6


def test_patch_kernels_with_a_second_logical_workgroup(ctx):
    """At 730x730 the row kernels have 528 logical workgroups: 512 run, and the first
    16 also take logical workgroups 512-527 (a second patch of the SpMV, a second
    row range of the update and the prolongation).  The mixed multilevel step with
    the patch SpMV must meet the FP64 true residual (formed by the SELL kernel, so a
    wrong row in the second pass would show) and agree with the SELL variant."""
    n = 730
    sc = scene.Scene(n, sigma_x=1e-3, sigma_v=0.05)
    ctx.set_topology(sc.X, sc.tri, sc.hinge)
    ctx.set_material(0.1, 500.0, 0.3, 1e-5, 0.0, 1.0)
    ctx.set_options(eolc.OPT_SPD_PROJECT_BENDING)
    out = {}
    for var in (0, 25):
        ctx.set_spmv_variant(32, var)
        ctx.set_spmv_variant(64, var)
        for name, mode in (("mixed", eolc.SOLVE_MIXED_BJ), ("fp64", eolc.SOLVE_FP64_BJ)):
            ctx.upload_state(sc.x0, sc.v0)
            rc, st = ctx.step(GRAV, H, sc.fixed, tol=1e-8, maxit=60000, mode=mode | eolc.SOLVE_MULTILEVEL)
            print(f"  {n}^2 {name} variant {var}: rc {rc} {st.iters} iterations, true residual {st.true_relres:.2e}",
                  flush=True)
            assert rc == 0 and st.true_relres <= 1e-8
            out[(name, var)] = ctx.download_state()[1]
    for name in ("mixed", "fp64"):
        assert rel_err(out[(name, 25)], out[(name, 0)]) <= 1e-4
    ctx.set_spmv_variant(32, 0)
    ctx.set_spmv_variant(64, 0)
