"""Batched cloths: one persistent launch, per-instance convergence, HIP graph.

Each instance of a batch must match the same cloth stepped alone in its own
context (mixed-precision solve, tol 1e-10 on both sides: velocities within 1e-6
relative of each other -- the two solves stop at different iterates of
residual <= 1e-10 -- and positions within 1e-9).  Graph replays
are bitwise identical to the eager step, with eager work and host
synchronisation between replays; every workgroup-size variant gives the same
bits."""
import numpy as np
import pytest

import eolc
from conftest import GRAV, H
from eolc import scene

pytestmark = pytest.mark.gpu


def small_batch():
    return scene.BatchScene(B=5, side_lo=6, side_hi=20, seed=77)


def test_batch_instances_match_single_cloths():
    bs = small_batch()
    print("batch vs single cloths, sides", bs.sides, flush=True)
    with eolc.Context(0) as c:
        bs.setup(c)
        rc, st = c.batch_step(GRAV, H, tol=1e-10, maxit=20000, eager=True)
        assert rc == eolc.OK
        x, v = c.download_state()
    xs, vs = bs.unpack(x), bs.unpack(v)
    for b, sc in enumerate(bs.inst):
        assert st[b].status == eolc.OK and st[b].true_relres <= 1e-10, st[b].as_dict()
        with eolc.Context(0) as s:
            s.set_topology(sc.X, sc.tri, sc.hinge)
            s.set_material(0.1, float(bs.E[b]), 0.3, float(bs.beta[b]), 0.0, 1.0)
            s.upload_state(sc.x0, sc.v0)
            rc1, st1 = s.step(GRAV, H, sc.fixed, tol=1e-10, maxit=20000, mode=eolc.SOLVE_MIXED_BJ)
            x1, v1 = s.download_state()
        err = np.max(np.abs(vs[b] - v1)) / np.max(np.abs(v1))
        print(f"  instance {b} side {sc.n}: batch iters {st[b].iters} single {st1.iters} v rel err {err:.2e}",
              flush=True)
        assert err < 1e-6
        assert np.max(np.abs(xs[b] - x1)) <= 1e-9 * np.max(np.abs(x1))


def test_batch_graph_replays_match_eager_and_variants_agree():
    bs = small_batch()
    with eolc.Context(0) as c:
        bs.setup(c)
        x0, v0 = c.download_state()
        ref = None
        for variant in range(eolc.BATCH_VARIANTS):
            c.set_batch_variant(variant)
            c.upload_state(x0, v0)
            rc, st = c.batch_step(GRAV, H, tol=1e-9, eager=True)
            out = c.download_state()
            if ref is None:
                ref = out
            print(f"eager variant {variant}: iters {[s.iters for s in st]}", flush=True)
            assert np.array_equal(out[0], ref[0]) and np.array_equal(out[1], ref[1])
        for replay in range(3):
            c.upload_state(x0, v0)
            c.download_state()                   # host synchronisation between replays
            rc, st = c.batch_step(GRAV, H, tol=1e-9)   # graph (captured on the first replay)
            out = c.download_state()
            print(f"graph replay {replay}: iters {[s.iters for s in st]}", flush=True)
            assert rc == eolc.OK
            assert np.array_equal(out[0], ref[0]) and np.array_equal(out[1], ref[1])
        # two graph steps in a row equal two eager steps
        c.upload_state(x0, v0)
        c.batch_step(GRAV, H, tol=1e-9)
        c.batch_step(GRAV, H, tol=1e-9)
        g2 = c.download_state()
        c.upload_state(x0, v0)
        c.batch_step(GRAV, H, tol=1e-9, eager=True)
        c.batch_step(GRAV, H, tol=1e-9, eager=True)
        e2 = c.download_state()
        assert np.array_equal(g2[0], e2[0]) and np.array_equal(g2[1], e2[1])
        before = c.download_state()
        best, ms = c.batch_autotune(GRAV, H, tol=1e-9)
        print(f"batch autotune: best {best} ms {ms}", flush=True)
        assert 0 <= best < eolc.BATCH_VARIANTS and np.all(ms > 0)
        after = c.download_state()
        assert np.array_equal(after[0], before[0]) and np.array_equal(after[1], before[1])


def test_batch_reports_per_instance_maxit():
    bs = small_batch()
    with eolc.Context(0) as c:
        bs.setup(c)
        rc, st = c.batch_step(GRAV, H, tol=1e-14, maxit=3, eager=True, allow=(eolc.ERR_MAXIT,))
        assert rc == eolc.ERR_MAXIT
        assert all(s.status == eolc.ERR_MAXIT and s.iters <= 3 for s in st)


def test_batch_multilevel():
    """The multilevel coarse levels inside the persistent solve: every instance
    matches the same cloth stepped alone with EOLC_SOLVE_MULTILEVEL (tol 1e-10 both
    sides), needs fewer iterations than the block-Jacobi batch, graph replays equal
    the eager step bitwise, and every workgroup size gives the same bits."""
    bs = scene.BatchScene(B=4, side_lo=10, side_hi=40, seed=78)
    print("batch multilevel, sides", bs.sides, flush=True)
    with eolc.Context(0) as c:
        bs.setup(c)
        x0, v0 = c.download_state()
        rc, st_bj = c.batch_step(GRAV, H, tol=1e-10, maxit=20000, eager=True)
        c.set_batch_multilevel(True)
        ref = None
        for variant in range(eolc.BATCH_VARIANTS):
            c.set_batch_variant(variant)
            c.upload_state(x0, v0)
            rc, st = c.batch_step(GRAV, H, tol=1e-10, maxit=20000, eager=True)
            assert rc == eolc.OK
            out = c.download_state()
            if ref is None:
                ref, st_ml = out, st
            assert np.array_equal(out[0], ref[0]) and np.array_equal(out[1], ref[1])
        print(f"  iterations block-Jacobi {[s.iters for s in st_bj]}, multilevel {[s.iters for s in st_ml]}",
              flush=True)
        for replay in range(2):
            c.upload_state(x0, v0)
            c.download_state()
            rc, st = c.batch_step(GRAV, H, tol=1e-10, maxit=20000)
            out = c.download_state()
            assert rc == eolc.OK
            assert np.array_equal(out[0], ref[0]) and np.array_equal(out[1], ref[1])
    xs, vs = bs.unpack(ref[0]), bs.unpack(ref[1])
    assert sum(s.iters for s in st_ml) < 0.8 * sum(s.iters for s in st_bj)
    for b, sc in enumerate(bs.inst):
        assert st_ml[b].status == eolc.OK and st_ml[b].true_relres <= 1e-10
        with eolc.Context(0) as s:
            s.set_topology(sc.X, sc.tri, sc.hinge)
            s.set_material(0.1, float(bs.E[b]), 0.3, float(bs.beta[b]), 0.0, 1.0)
            s.upload_state(sc.x0, sc.v0)
            rc1, st1 = s.step(GRAV, H, sc.fixed, tol=1e-10, maxit=20000,
                              mode=eolc.SOLVE_MIXED_BJ | eolc.SOLVE_MULTILEVEL)
            x1, v1 = s.download_state()
        err = np.max(np.abs(vs[b] - v1)) / np.max(np.abs(v1))
        print(f"  instance {b} side {sc.n}: batch iters {st_ml[b].iters} single {st1.iters} v rel err {err:.2e}",
              flush=True)
        assert err < 1e-6
        assert np.max(np.abs(xs[b] - x1)) <= 1e-9 * np.max(np.abs(x1))


def test_batch_instances_match_the_oracle():
    """Each instance of a batch, with its own Young's modulus and bending stiffness
    (per-face E and per-hinge beta in the device kernels), against the oracle's
    Forces::fill with that instance's material (Forces.cpp:912-930; Material per cloth,
    Cloth.h:28-35): f, M and MDK to 1e-10 of the largest entry with the same
    sparsity; and the batched solve (tol 1e-10) against the oracle's dense solve of
    the oracle's own system, with the corners at 0.99 v_old, to 1e-6 relative (the
    bound that a 1e-10 residual gives at these systems' condition numbers)."""
    import oracle as O
    bs = small_batch()
    with eolc.Context(0) as c:
        off = bs.setup(c)
        rc, st = c.batch_step(GRAV, H, tol=1e-10, maxit=20000, eager=True)
        assert rc == eolc.OK
        x, v = c.download_state()
        f = c.download_forces()
        M = c.download_csr(eolc.MAT_M).tocsr()
        A = c.download_csr(eolc.MAT_MDK).tocsr()
    vs = bs.unpack(v)
    for b, sc in enumerate(bs.inst):
        lo, n3 = 3 * int(off[b]), 3 * sc.N
        mat = O.material(0.1, float(bs.E[b]), 0.3, float(bs.beta[b]), 0.0, 1.0)
        rco, fo, Mo, Ao = O.forces_fill(sc.x0, sc.X, sc.tri, sc.hinge, mat, GRAV, H)
        assert rco == 0
        errs = []
        for G, R in ((M, Mo), (A, Ao)):
            Gb = G[lo:lo + n3, lo:lo + n3].tocsr()
            R = R.tocsr()
            Gb.sort_indices(); R.sort_indices()
            assert np.array_equal(Gb.indptr, R.indptr) and np.array_equal(Gb.indices, R.indices)
            errs.append(np.max(np.abs(Gb.data - R.data)) / np.max(np.abs(R.data)))
        errs.append(np.max(np.abs(f[lo:lo + n3] - fo)) / np.max(np.abs(fo)))
        # outside its block the instance's rows are empty (no coupling between cloths)
        assert A[lo:lo + n3].nnz == A[lo:lo + n3, lo:lo + n3].nnz
        fd = (3 * np.asarray(sc.fixed)[:, None] + np.arange(3)).reshape(-1).astype(np.int32)
        rhs = Mo @ sc.v0 + H * fo
        rcd, vd = O.dense_velocity_solve(Ao, rhs, fd, 0.99 * sc.v0[fd])
        assert rcd == 0
        ev = np.max(np.abs(vs[b] - vd)) / np.max(np.abs(vd))
        print(f"  instance {b} side {sc.n} E {bs.E[b]:.3g} beta {bs.beta[b]:.3g}: M {errs[0]:.1e} "
              f"MDK {errs[1]:.1e} f {errs[2]:.1e}; v vs dense {ev:.1e} ({st[b].iters} iterations)", flush=True)
        assert max(errs) <= 1e-10
        assert ev <= 1e-6
