"""Proximity-query parity: GPU CD2 / CD records against the oracle's brute-force
restatement of btc::pointTriCollision and btc::boxTriCollision.

Bar: bit-exact -- every field of every record (indices, types, and the FP64
distances, normals, positions and weights compared by bit pattern), in the
reference's order."""
import numpy as np
import pytest

import eolc
import oracle as O
from eolc import scene

pytestmark = pytest.mark.gpu

FIELDS = ("dist", "nor1", "nor2", "pos1", "pos2", "pos1_", "count1", "count2", "verts1", "verts2",
          "weights1", "weights2", "tri1", "tri2", "nedge1", "edge2", "edgeDir")


def rec_tuple(r):
    out = []
    for f in FIELDS:
        v = getattr(r, f)
        if isinstance(v, float):
            out.append(np.float64(v).tobytes())
        elif isinstance(v, int):
            out.append(v)
        else:
            a = np.array(list(v))
            out.append(a.tobytes() if a.dtype.kind == "f" else tuple(a.tolist()))
    out.append(tuple(list(r.edge1)[: r.nedge1]))
    return tuple(out)


def oracle_cd(thr, pts, nrm, whd, E, x, tri, eol=False, box_npts=8, box_nedges=12):
    n, recs = O.point_tri(thr, pts, nrm, x, tri, eol=eol)
    assert n >= 0
    P = len(pts)
    out = list(recs)
    for b in range(len(whd)):
        nb, br = O.box_tri(thr, whd[b], E[b], x, tri, eol=False)
        assert nb >= 0
        if eol:
            base = P + b * box_npts + b * box_nedges
            for r in br:
                if r.count1 == 1 and r.count2 == 3:
                    r.verts1[0] = base + r.verts1[0]
                for e in range(r.nedge1):
                    r.edge1[e] = base + (box_npts + r.edge1[e])
        out.extend(br)
    return out


def compare(ctx, sc, x, thr, pts, nrm, whd, E, eol=False):
    ctx.set_topology(sc.X, sc.tri, sc.hinge)
    ctx.upload_state(x, np.zeros_like(x))
    g = ctx.collide(thr, pts, nrm, whd, E, eol_points=eol)
    o = oracle_cd(thr, pts, nrm, whd, E, x, sc.tri, eol=eol)
    kinds = {}
    for r in o:
        kinds[(r.count1, r.count2)] = kinds.get((r.count1, r.count2), 0) + 1
    print(f"records gpu {len(g)} oracle {len(o)} kinds {kinds}", flush=True)
    assert len(g) == len(o)
    for k, (a, b) in enumerate(zip(g, o)):
        assert rec_tuple(a) == rec_tuple(b), f"record {k} differs"
    return kinds


def thin_box_scene(n, sigma):
    sc = scene.Scene(n, sigma_x=sigma)
    pts, nrm, whd, E = scene.obstacles(seed=1236, P=64, B=4)
    # one thin box whose corners reach the cloth (V1-T2), one tilted box
    whd = np.vstack([whd, [0.2, 0.15, 0.006]])
    E = list(E) + [np.eye(4)]
    E[-1][0, 3], E[-1][1, 3], E[-1][2, 3] = 0.37, 0.61, 0.0011
    R = np.eye(4)
    c, s = np.cos(0.4), np.sin(0.4)
    R[:3, :3] = [[c, -s, 0], [s * 0.9, c * 0.9, 0.43589], [-s * 0.43589, -c * 0.43589, 0.9]]
    R[:3, :3] = np.linalg.qr(R[:3, :3])[0]
    R[0, 3], R[1, 3], R[2, 3] = 0.7, 0.3, 0.01
    whd = np.vstack([whd, [0.12, 0.1, 0.08]])
    E.append(R)
    return sc, pts, nrm, whd, E


def test_cd2_parity_small(ctx):
    sc, pts, nrm, whd, E = thin_box_scene(96, 1e-3)
    kinds = compare(ctx, sc, sc.x0, 1e-3, pts, nrm, whd, E)
    assert set(kinds) >= {(1, 3), (3, 1), (2, 2)}          # every record type occurs


def test_constraint_rows_from_gpu_records(ctx):
    """Constraints::fill collision rows (Constraints.cpp:417-470) from the GPU's CD2
    records equal the restated loop applied to the oracle's records, exactly."""
    from test_constraints import restated
    sc, pts, nrm, whd, E = thin_box_scene(96, 1e-3)
    ctx.set_topology(sc.X, sc.tri, sc.hinge)
    ctx.upload_state(sc.x0, np.zeros_like(sc.x0))
    g = ctx.collide(1e-3, pts, nrm, whd, E)
    o = oracle_cd(1e-3, pts, nrm, whd, E, sc.x0, sc.tri)
    A = eolc.collision_constraints(g, sc.N)
    R = restated(o, sc.N)
    print(f"constraint rows {A.shape[0]} from {len(g)} records, nnz {A.nnz}", flush=True)
    assert A.shape == R.shape and A.shape[0] > 0
    assert (abs(A - R)).max() == 0.0


def test_cd_eol_points_parity(ctx):
    sc, pts, nrm, whd, E = thin_box_scene(64, 1e-3)
    # points right on the cloth so the EOL vertex-point branch fires
    pts = pts.copy()
    pts[:8, :2] = sc.X[[5, 77, 300, 901, 1500, 2222, 3000, 4000]]
    pts[:8, 2] = 1e-4
    compare(ctx, sc, sc.x0, 2e-3, pts, nrm, whd, E, eol=True)


def test_cd2_no_obstacles_and_far_obstacles(ctx):
    sc = scene.Scene(16, sigma_x=1e-3)
    ctx.set_topology(sc.X, sc.tri, sc.hinge)
    ctx.upload_state(sc.x0, np.zeros_like(sc.x0))
    assert ctx.collide(1e-3) == []
    pts = np.array([[0.5, 0.5, 3.0]]); nrm = np.array([[0, 0, 1.0]])
    assert ctx.collide(1e-3, pts, nrm) == []


def test_overflow_reports_required_size(ctx):
    import ctypes as C
    sc, pts, nrm, whd, E = thin_box_scene(48, 1e-3)
    ctx.set_topology(sc.X, sc.tri, sc.hinge)
    ctx.upload_state(sc.x0, np.zeros_like(sc.x0))
    full = ctx.collide(1e-3, pts, nrm, whd, E)
    assert len(full) > 1
    buf = (eolc.CollisionRec * 1)()
    cnt = C.c_int(0)
    Es = np.concatenate([np.asarray(e).T.reshape(-1) for e in E])
    rc = eolc.lib().eolc_collide(ctx._h, 1e-3, len(pts), pts.ctypes.data_as(C.c_void_p),
                                 nrm.ctypes.data_as(C.c_void_p), len(whd),
                                 np.ascontiguousarray(whd).ctypes.data_as(C.c_void_p),
                                 Es.ctypes.data_as(C.c_void_p), 0, 8, 12, buf, 1, C.byref(cnt))
    assert rc == eolc.ERR_OVERFLOW and cnt.value == len(full)


def test_cd2_parity_config2(ctx):
    """512x512 cloth with noise 1e-3, 64 points and 4 boxes at threshold 1e-3."""
    sc = scene.Scene(512, sigma_x=1e-3)
    pts, nrm, whd, E = scene.obstacles(seed=1236, P=64, B=4)
    compare(ctx, sc, sc.x0, 1e-3, pts, nrm, whd, E)


def _material_layout(verts, tri):
    """A material-space layout with every face CCW (the context requires positive
    areas): the first axis-pair projection of the positions that has one."""
    v = np.asarray(verts, dtype=np.float64)
    for i, j in ((0, 1), (1, 0), (0, 2), (2, 0), (1, 2), (2, 1)):
        X = v[:, [i, j]]
        a, b, c = X[tri[:, 0]], X[tri[:, 1]], X[tri[:, 2]]
        if np.all((b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0]) > 0):
            return X
    raise AssertionError("no CCW projection")


@pytest.mark.parametrize("name", ["point_over_triangle", "points_nearest_and_tie", "eol_points",
                                  "corner_under_triangle", "vertex_inside_box", "edge_through_box_edge"])
def test_closed_form_scenes(ctx, name):
    """The hand-worked scenes of the oracle's pins through the GPU path: records bit
    for bit equal to the oracle's, and equal to the closed-form values."""
    import collide_scenes as S
    print(f"closed-form collision scene {name}", flush=True)
    built = getattr(S, name)()
    if len(built) == 4:
        verts, tri, pts, nrm = built
        whd, E = np.zeros((0, 3)), []
    else:
        verts, tri = built
        pts, nrm = np.zeros((0, 3)), np.zeros((0, 3))
        whd, E = S.BOX_WHD[None, :], [S.box_frame(S.BOX_CENTER)]
    eol = name == "eol_points"
    X = _material_layout(verts, tri)
    ctx.set_topology(X, tri, np.zeros((0, 4), dtype=np.int32))
    x = np.asarray(verts, dtype=np.float64).reshape(-1)
    ctx.upload_state(x, np.zeros_like(x))
    g = ctx.collide(S.THR, pts, nrm, whd, E, eol_points=eol)
    o = oracle_cd(S.THR, pts, nrm, whd, E, x, tri, eol=eol)
    assert len(g) == len(o) and len(g) > 0
    for k, (a, b) in enumerate(zip(g, o)):
        assert rec_tuple(a) == rec_tuple(b), f"record {k} differs"
    exp_fn = getattr(S, "expected_" + name, None)
    if exp_fn is not None:
        v2 = S.jittered(verts, O.mt19937_uniform)
        exp = exp_fn(v2)
        S.compare(g[: len(exp)] if eol else g, exp)
