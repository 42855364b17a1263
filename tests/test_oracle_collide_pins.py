"""Pins of the oracle's proximity queries (orc_point_tri, orc_box_tri) on scenes whose
records are known in closed form (tests/collide_scenes.py): every record type of
btc::pointTriCollision and btc::boxTriCollision (boxTriCollision.cpp:602-1224), with
the fields a plausible slip would change (which face or edge, the sign of a normal,
the order of barycentric weights, the clamp branch of an edge record, the vertex /
face index mix of V2-T1)."""
import numpy as np

import collide_scenes as S
import oracle as O


def jit(verts):
    return S.jittered(verts, O.mt19937_uniform)


def test_point_over_triangle_record():
    verts, tri, pts, nrm = S.point_over_triangle()
    n, recs = O.point_tri(S.THR, pts, nrm, verts, tri)
    S.compare(recs, S.expected_point_over_triangle(jit(verts)))
    np.testing.assert_allclose(list(recs[0].weights2), [0.5, 0.2, 0.3], atol=1e-5)
    assert abs(recs[0].dist - 1.5e-3) < 1e-5


def test_point_rules_reject():
    """The side test (proj >= 0 against the normal flipped toward the point's), the
    5 thr distance, the fixed 1e-3 AABB pad and the barycentric range each reject."""
    verts, tri, pts, nrm = S.point_over_triangle()
    # the point below its own normal: the flipped face normal points down, proj < 0
    n, _ = O.point_tri(S.THR, pts, -nrm, verts, tri)
    assert n == 0
    # 1.5 mm above but the threshold 0.2 mm: 5 thr = 1 mm < 1.5 mm
    n, _ = O.point_tri(2e-4, pts, nrm, verts, tri)
    assert n == 0
    # 2.5 mm above with thr 1e-3: within 5 thr, but outside the AABB pads (1e-3 on each)
    n, _ = O.point_tri(S.THR, pts + [0, 0, 1e-3], nrm, verts, tri)
    assert n == 0
    # projected outside the triangle (x + y > 1)
    n, _ = O.point_tri(S.THR, [[0.6, 0.6, 1.5e-3]], nrm, verts, tri)
    assert n == 0


def test_point_nearest_face_and_first_of_ties():
    verts, tri, pts, nrm = S.points_nearest_and_tie()
    n, recs = O.point_tri(S.THR, pts, nrm, verts, tri)
    assert n == 2
    assert recs[0].tri2 == 1 and abs(recs[0].dist - 0.5e-3) < 1e-5
    assert recs[1].tri2 == 0 and abs(recs[1].dist - 1.5e-3) < 1e-5
    v2 = jit(verts)
    for r, x1 in zip(recs, pts):
        f = tri[r.tri2]
        S.compare([r], [S.rec(count1=1, count2=3, verts1=[int(r.verts1[0]), -1, -1], verts2=list(f),
                              tri2=r.tri2, weights1=[1.0, 0.0, 0.0],
                              weights2=S.bary(v2[f[0]], v2[f[1]], v2[f[2]], x1))])


def test_eol_point_records():
    verts, tri, pts, nrm = S.eol_points()
    n, recs = O.point_tri(S.THR, pts, nrm, verts, tri, eol=True)
    exp = S.expected_eol_points(jit(verts))
    # the EOL records come first, in cloth-vertex order; then the point-face records
    S.compare(recs[: len(exp)], exp)
    assert [r.verts1[0] for r in recs[: len(exp)]] == [1, 3]
    assert all(r.count1 == 1 and r.count2 == 3 for r in recs[len(exp):])


def test_box_corner_under_triangle():
    verts, tri = S.corner_under_triangle()
    n, recs = O.box_tri(S.THR, S.BOX_WHD, S.box_frame(S.BOX_CENTER), verts, tri)
    exp = S.expected_corner_under_triangle(jit(verts))
    S.compare(recs, exp)
    np.testing.assert_allclose(list(recs[0].weights2), [1 / 3, 1 / 3, 1 / 3], atol=1e-4)
    np.testing.assert_allclose(list(recs[0].nor1), -np.ones(3) / np.sqrt(3), atol=1e-15)
    # the edge records cross one face each: x2 on the cloth edge where it enters the
    # face of n1c, 1.5 mm above the box edge
    np.testing.assert_allclose(list(recs[1].weights2), [1 / 3, 2 / 3, 0.0], atol=1e-4)
    np.testing.assert_allclose(list(recs[2].weights2), [2 / 3, 1 / 3, 0.0], atol=1e-4)
    np.testing.assert_allclose([r.dist for r in recs], 1.5e-3, atol=1e-5)


def test_box_vertex_inside_uses_face_normal_of_vertex_index():
    verts, tri = S.vertex_inside_box()
    n, recs = O.box_tri(S.THR, S.BOX_WHD, S.box_frame(S.BOX_CENTER), verts, tri)
    S.compare(recs, S.expected_vertex_inside_box(jit(verts)))
    np.testing.assert_allclose(list(recs[0].weights1), [0.05, 0.7, 0.25], atol=1e-5)
    # face 1 (the face that holds vertex 0) has a different normal: the quirk decides
    v2 = jit(verts)
    n_own = S.face_normal(v2[0], v2[4], v2[5])
    assert np.max(np.abs(np.array(recs[0].nor2) - n_own)) > 0.1 and np.max(np.abs(np.array(recs[0].nor2) + n_own)) > 0.1


def test_box_edge_through_box_edge():
    verts, tri = S.edge_through_box_edge()
    n, recs = O.box_tri(S.THR, S.BOX_WHD, S.box_frame(S.BOX_CENTER), verts, tri)
    S.compare(recs, S.expected_edge_through_box_edge(jit(verts)))
    assert abs(recs[0].dist - 5e-4 * np.sqrt(2)) < 1e-5
    np.testing.assert_allclose([recs[0].weights1[1], recs[0].weights2[1]], [0.5, 0.5], atol=1e-4)


def test_box_eol_mode_skips_vertex_records():
    """With EOL the V2-T1 pass is skipped (conformal geometry, 671-673)."""
    verts, tri = S.vertex_inside_box()
    n, recs = O.box_tri(S.THR, S.BOX_WHD, S.box_frame(S.BOX_CENTER), verts, tri, eol=True)
    assert n == 0
