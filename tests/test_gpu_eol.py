"""EOL rows of Forces::fill on the GPU (eolc_set_eol) against the oracle's EOL fill
(orc_forces_fill_eol, itself pinned by the dense mapped element sum on the CPU).

Bars: f (3N + 2E), M and MDK ((3N + 2E) square) within 1e-10 of the largest entry
of the oracle's array, with the oracle's sparsity pattern; the x-x part equal bit
for bit to the GPU's fill without EOL nodes; solves refused while EOL nodes are
set; clearing them restores the plain fill bit for bit."""
import numpy as np
import pytest

import eolc
import oracle as O
from conftest import GRAV, H, rel_err
from eolc import scene

pytestmark = pytest.mark.gpu


def boundary(n):
    ij = np.array([(k % (n + 1), k // (n + 1)) for k in range((n + 1) ** 2)])
    return np.where((ij[:, 0] == 0) | (ij[:, 0] == n) | (ij[:, 1] == 0) | (ij[:, 1] == n))[0].astype(np.int32)


def fill(ctx, sc, flags, eol):
    ctx.set_topology(sc.X, sc.tri, sc.hinge)
    ctx.set_material(0.1, 500.0, 0.3, 1e-5, 0.0, 1.0)
    ctx.set_options(flags)
    ctx.set_eol(eol)
    ctx.upload_state(sc.x0, sc.v0)
    ctx.forces_fill(GRAV, H)
    return ctx.download_forces(), ctx.download_csr(eolc.MAT_M), ctx.download_csr(eolc.MAT_MDK)


@pytest.mark.parametrize("case", ["boundary", "mixed", "mixed-bending-force"])
def test_eol_fill_matches_oracle(ctx, case):
    n = 20
    sc = scene.Scene(n, sigma_x=1e-3, sigma_v=0.05)
    N = sc.N
    eol = boundary(n)
    if case != "boundary":
        eol = np.concatenate([eol[::3], [n + 3, 5 * (n + 1) + 7, 5 * (n + 1) + 8, N // 2]]).astype(np.int32)
    flags = eolc.OPT_INCLUDE_BENDING_FORCE if case == "mixed-bending-force" else 0
    print(f"EOL fill {case}: {n}x{n}, {len(eol)} EOL nodes", flush=True)
    fg, Mg, Kg = fill(ctx, sc, flags, eol)
    oflags = O.INCLUDE_BENDING_FORCE if flags else 0
    rc, fo, Mo, Ko = O.forces_fill_eol(sc.x0, sc.X, sc.tri, sc.hinge, O.material(0.1, 500.0, 0.3, 1e-5, 0.0, 1.0),
                                       GRAV, H, eol, flags=oflags)
    assert rc == 0
    n3 = 3 * N + 2 * len(eol)
    assert fg.shape == (n3,) and Mg.shape == (n3, n3) and Kg.shape == (n3, n3)
    ef, em, ek = rel_err(fg, fo), rel_err(Mg.toarray(), Mo.toarray()), rel_err(Kg.toarray(), Ko.toarray())
    print(f"  rel err f {ef:.2e}, M {em:.2e}, MDK {ek:.2e}; X-row entries {Kg[3 * N:].nnz}", flush=True)
    assert ef <= 1e-10 and em <= 1e-10 and ek <= 1e-10
    for A, B in ((Mg, Mo), (Kg, Ko)):
        A, B = A.tocsr(), B.tocsr()
        A.sort_indices(); B.sort_indices()
        assert np.array_equal(A.indptr, B.indptr) and np.array_equal(A.indices, B.indices)
    # the x-x part is the plain fill's, bit for bit; the x forces differ only by
    # the EOL hinges' bending force
    f0, M0, K0 = fill(ctx, sc, flags, [])
    assert np.array_equal(Kg[:3 * N, :3 * N].toarray(), K0.toarray())
    assert np.array_equal(Mg[:3 * N, :3 * N].toarray(), M0.toarray())
    if flags:
        assert np.array_equal(fg[:3 * N], f0)
    else:
        assert not np.array_equal(fg[:3 * N], f0)


def test_eol_refuses_solves_and_clears(ctx):
    sc = scene.Scene(12, sigma_x=1e-3, sigma_v=0.05)
    eol = boundary(12)
    fg, Mg, Kg = fill(ctx, sc, 0, eol)
    fd = (3 * np.asarray(sc.fixed)[:, None] + np.arange(3)).reshape(-1)
    rc, _ = ctx.velocity_solve(fd, 0.99 * sc.v0[fd], allow=(eolc.ERR_INVALID,))
    assert rc == eolc.ERR_INVALID
    rc, _ = ctx.step(GRAV, H, sc.fixed, allow=(eolc.ERR_INVALID,))
    assert rc == eolc.ERR_INVALID
    with pytest.raises(eolc.EolcError):
        ctx.set_eol([3, 3])
    with pytest.raises(eolc.EolcError):
        ctx.set_eol([sc.N])
    ctx.set_eol([])
    ctx.forces_fill(GRAV, H)
    f1, K1 = ctx.download_forces(), ctx.download_csr(eolc.MAT_MDK)
    with eolc.Context(0) as c2:
        f2, M2, K2 = fill(c2, sc, 0, [])
    assert np.array_equal(f1, f2) and np.array_equal(K1.toarray(), K2.toarray())
    rc, st = ctx.velocity_solve(fd, 0.99 * sc.v0[fd], tol=1e-10)
    assert rc == eolc.OK
