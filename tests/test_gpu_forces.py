"""Forces::fill parity: GPU (FP64) f, M and MDK against the CPU oracle.

Tolerance: 1e-10 relative to the largest entry of the oracle's array."""
import numpy as np
import pytest

import eolc
import oracle as O
from conftest import GRAV, H, rel_err
from eolc import scene

pytestmark = pytest.mark.gpu


def run_both(ctx, sc, x, v, flags=0, mat=(0.1, 500.0, 0.3, 1e-5, 0.0, 1.0)):
    ctx.set_topology(sc.X, sc.tri, sc.hinge)
    ctx.set_material(*mat)
    ctx.set_options(flags)
    ctx.upload_state(x, v)
    ctx.forces_fill(GRAV, H)
    fg = ctx.download_forces()
    Mg = ctx.download_csr(eolc.MAT_M)
    Kg = ctx.download_csr(eolc.MAT_MDK)
    oflags = O.INCLUDE_BENDING_FORCE if flags & eolc.OPT_INCLUDE_BENDING_FORCE else 0
    rc, fo, Mo, Ko = O.forces_fill(x, sc.X, sc.tri, sc.hinge, O.material(*mat), GRAV, H, oflags)
    assert rc == 0
    return (fg, Mg, Kg), (fo, Mo, Ko)


def same_pattern(A, B):
    A, B = A.tocsr(), B.tocsr()
    A.sort_indices(); B.sort_indices()
    return np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)


@pytest.mark.parametrize("n,sigma", [(20, 0.0), (20, 1e-3), (7, 1e-2), (33, 5e-3)])
def test_forces_parity(ctx, n, sigma):
    sc = scene.Scene(n, sigma_x=sigma, sigma_v=0.05)
    print(f"forces parity n={n} sigma={sigma}", flush=True)
    (fg, Mg, Kg), (fo, Mo, Ko) = run_both(ctx, sc, sc.x0, sc.v0)
    assert same_pattern(Mg, Mo) and same_pattern(Kg, Ko)
    assert rel_err(fg, fo) < 1e-10
    assert rel_err(Mg.data, Mo.data) < 1e-14
    assert rel_err(Kg.data, Ko.data) < 1e-10
    assert abs(Kg - Kg.T).max() == 0.0          # exactly symmetric by construction


def test_forces_include_bending(ctx):
    sc = scene.Scene(12, sigma_x=1e-2)
    (fg, _, _), (fo, _, _) = run_both(ctx, sc, sc.x0, sc.v0, flags=eolc.OPT_INCLUDE_BENDING_FORCE)
    assert rel_err(fg, fo) < 1e-10
    (fg0, _, _), _ = run_both(ctx, sc, sc.x0, sc.v0, flags=0)
    assert np.abs(fg - fg0).max() > 0      # the bending force is really there


def test_forces_mass_totals(ctx):
    """Sum of M = total mass rho * area in each coordinate; sum f = m g at rest."""
    sc = scene.Scene(16)
    (fg, Mg, _), _ = run_both(ctx, sc, sc.x0, sc.v0)
    assert abs(Mg.sum() - 3 * 0.1 * 1.0) < 1e-14
    assert np.allclose(fg.reshape(-1, 3).sum(0), 0.1 * GRAV, rtol=1e-12, atol=1e-15)


def test_forces_deterministic(ctx):
    sc = scene.Scene(24, sigma_x=3e-3)
    a, _ = run_both(ctx, sc, sc.x0, sc.v0)
    ctx.forces_fill(GRAV, H)
    b = (ctx.download_forces(), ctx.download_csr(eolc.MAT_M), ctx.download_csr(eolc.MAT_MDK))
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[2].data, b[2].data)


def test_degenerate_face_is_reported(ctx):
    sc = scene.Scene(4)
    tri = sc.tri.copy()
    tri[0] = tri[0][[0, 2, 1]]                  # clockwise face
    with pytest.raises(eolc.EolcError) as e:
        ctx.set_topology(sc.X, tri, sc.hinge)
    assert e.value.code == eolc.ERR_DEGENERATE


def test_spd_projection_matches_eigh(ctx):
    """EOLC_OPT_SPD_PROJECT_BENDING against a textbook restatement: per hinge,
    K9 = Q^T K Q (Q: orthonormal contrast basis of the complement of the
    translations), K+ = Q V max(L, 0) V^T Q^T with numpy's eigh, K from the oracle's
    bending Hessian.  The assembled difference MDK(projected) - MDK(exact) must equal
    h^2 sum_h (K+_h - K_h) to 1e-10 of the largest hinge Hessian entry."""
    import eolc
    from eolc import scene
    sc = scene.Scene(5, sigma_x=0.06, sigma_v=0.0, seed_x=99)
    x = sc.x0.reshape(-1, 3)
    mats = []
    for flags in (0, eolc.OPT_SPD_PROJECT_BENDING):
        ctx.set_topology(sc.X, sc.tri, sc.hinge)
        ctx.set_material(0.1, 500.0, 0.3, 1e-5, 0.0, 1.0)
        ctx.set_options(flags)
        ctx.upload_state(sc.x0, sc.v0)
        ctx.forces_fill(GRAV, H)
        mats.append(ctx.download_csr(eolc.MAT_MDK).toarray())
    ctx.set_options(0)
    diff = mats[1] - mats[0]
    r2, r6, r12 = 1 / np.sqrt(2), 1 / np.sqrt(6), 1 / np.sqrt(12)
    cw = np.array([[r2, -r2, 0, 0], [r6, r6, -2 * r6, 0], [r12, r12, r12, -3 * r12]])
    Q = np.kron(cw.T, np.eye(3))              # 12 x 9, orthonormal columns
    ref = np.zeros_like(diff)
    kmax, nneg = 0.0, 0
    for q in sc.hinge:
        _, _, K = O.bending(x[q[0]], x[q[1]], x[q[2]], x[q[3]], sc.X[q[0]], sc.X[q[1]], sc.X[q[2]], sc.X[q[3]], 1e-5)
        w, V = np.linalg.eigh(Q.T @ K @ Q)
        nneg += int(np.sum(w < 0))
        Kp = Q @ (V * np.maximum(w, 0.0)) @ V.T @ Q.T
        dofs = (3 * np.asarray(q)[:, None] + np.arange(3)).reshape(-1)
        ref[np.ix_(dofs, dofs)] += H * H * (Kp - K)
        kmax = max(kmax, H * H * np.abs(K).max())
    print(f"projection: {len(sc.hinge)} hinges, {nneg} negative eigenvalues clamped, "
          f"max|diff - ref| / max|h^2 K| = {np.abs(diff - ref).max() / kmax:.2e}", flush=True)
    assert nneg > 0
    assert np.abs(diff - ref).max() <= 1e-10 * kmax
