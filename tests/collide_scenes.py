"""Small proximity-query scenes whose records are known in closed form.

Each scene is a cloth (vertices, triangles) plus obstacles at threshold 1e-3, built
so that exactly the records listed by its `expected` function are reported by
btc::pointTriCollision / btc::boxTriCollision (boxTriCollision.cpp:602-1224).  The
expected values are computed here with numpy from the jittered cloth positions
(the jitter stream is pinned separately against libstdc++) by formulas independent
of the oracle's: least-squares barycentric coordinates, normal equations for the
closest points of two lines, explicit geometry of the box.  Used by the oracle's CPU
pins and by the GPU proximity tests."""
import numpy as np

THR = 1e-3


def jittered(verts, uniform):
    """verts + U(-1, 1) thr 1e-3 per coordinate, in the reference's order of
    operations (dis(gen) * threshold * 1e-3), from the mt19937 stream `uniform`."""
    v = np.asarray(verts, dtype=np.float64).reshape(-1, 3)
    u = np.asarray(uniform(1, v.size)).reshape(-1, 3)
    return v + u * THR * 1e-3


def unit(v):
    v = np.asarray(v, dtype=np.float64)
    return v / np.linalg.norm(v)


def face_normal(a, b, c):
    return unit(np.cross(b - a, c - a))


def bary(a, b, c, p):
    """Weights (of a, b, c) of the point of the plane of abc closest to p."""
    A = np.stack([b - a, c - a], axis=1)
    (beta, gamma), *_ = np.linalg.lstsq(A, p - a, rcond=None)
    return np.array([1.0 - beta - gamma, beta, gamma])


def line_params(p1, d1, p2, d2):
    """(s, t) minimising |p1 + s d1 - p2 - t d2|."""
    w = p2 - p1
    M = np.array([[d1 @ d1, -(d1 @ d2)], [d1 @ d2, -(d2 @ d2)]])
    return np.linalg.solve(M, np.array([w @ d1, w @ d2]))


def box_frame(center):
    E = np.eye(4)
    E[:3, 3] = center
    return E


def rec(**kw):
    """An expected record with the reference's defaults for unset fields."""
    r = {"count1": 0, "count2": 0, "verts1": [-1, -1, -1], "verts2": [-1, -1, -1], "tri1": -1, "tri2": -1,
         "weights1": [0.0, 0.0, 0.0], "weights2": [0.0, 0.0, 0.0], "nedge1": 0, "edge1": [], "edge2": -1}
    r.update(kw)
    return r


# ---------------------------------------------------------------------------
# point obstacles over a single triangle (pointTriCollision, 1067-1224)

def point_over_triangle():
    """One CCW triangle in z = 0 and a point 1.5 mm above it with normal +z: a
    (1, 3) record with the foot point's barycentric weights (0.5, 0.2, 0.3)."""
    verts = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    tri = np.array([[0, 1, 2]])
    pts = np.array([[0.2, 0.3, 1.5e-3]])
    nrm = np.array([[0.0, 0.0, 1.0]])
    return verts, tri, pts, nrm


def expected_point_over_triangle(v2):
    verts, tri, pts, nrm = point_over_triangle()
    a, b, c = v2[0], v2[1], v2[2]
    x1, n1 = pts[0], nrm[0]
    n2 = face_normal(a, b, c)
    if n1 @ n2 < 0:
        n2 = -n2
    proj = (x1 - a) @ n2
    x2 = x1 - proj * n2
    return [rec(dist=np.linalg.norm(x2 - x1), nor1=n1, nor2=n2, pos1=x1, pos2=x2, count1=1, count2=3,
                verts1=[0, -1, -1], verts2=[0, 1, 2], weights1=[1.0, 0.0, 0.0], weights2=bary(a, b, c, x1),
                tri2=0, pos1_=x1 - 0.1 * THR * n1)]


def points_nearest_and_tie():
    """Two stacked copies of a triangle (z = 0 and z = 1 mm) and the z = 0 triangle
    repeated: a point at z = 1.5 mm takes the nearer face (index 1, 0.5 mm); a point
    below the sheet with normal -z sees the faces at 1.5 / 2.5 mm and keeps the
    first of the two identical faces at 1.5 mm (index 0; strict <)."""
    verts = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0],
                      [0.0, 0.0, 1e-3], [1.0, 0.0, 1e-3], [0.0, 1.0, 1e-3]])
    tri = np.array([[0, 1, 2], [3, 4, 5], [0, 1, 2]])
    pts = np.array([[0.3, 0.3, 1.5e-3], [0.3, 0.3, -1.5e-3]])
    nrm = np.array([[0.0, 0.0, 1.0], [0.0, 0.0, -1.0]])
    return verts, tri, pts, nrm


def eol_points():
    """Cloth vertices against nearby points (the EOL branch, 1099-1139): vertex 0
    has points at 0.6 thr (index 0) and 0.3 thr (index 1) and keeps index 1;
    vertex 1 has one at 1.2 thr and gets no record; vertex 2 has two points at the
    same 0.5 thr and keeps the first."""
    verts = np.array([[0.0, 0.0, 0.0], [1.0, 0.0, 0.0], [0.0, 1.0, 0.0]])
    tri = np.array([[0, 1, 2]])
    pts = np.array([[0.0, 0.0, 0.6e-3], [0.0, 0.0, -0.3e-3], [1.0, 0.0, 1.2e-3],
                    [0.0, 1.0, 0.5e-3], [0.0, 1.0, -0.5e-3]])
    nrm = np.array([[0.0, 0.0, 1.0], [0.0, 0.0, -1.0], [0.0, 0.0, 1.0], [0.0, 0.0, 1.0], [0.0, 0.0, -1.0]])
    return verts, tri, pts, nrm


def expected_eol_points(v2):
    verts, tri, pts, nrm = eol_points()
    out = []
    for i2, cand in ((0, (0, 1)), (2, (3, 4))):
        d = [np.linalg.norm(v2[i2] - pts[i]) for i in cand]
        i1 = cand[int(np.argmin(d))]      # argmin keeps the first of equal distances
        out.append(rec(dist=np.linalg.norm(v2[i2] - pts[i1]), nor1=nrm[i1], nor2=nrm[i1], pos1=pts[i1],
                       pos2=v2[i2], count1=3, count2=1, verts1=[i1, -1, -1], verts2=[i2, -1, -1],
                       weights1=[1.0, 0.0, 0.0], weights2=[1.0, 0.0, 0.0], pos1_=pts[i1] - 0.1 * THR * nrm[i1]))
    return out


# ---------------------------------------------------------------------------
# boxes (boxTriCollision, 602-1063).  Box of edge 0.2 centred at (0.5, 0.5, 0):
# corner k at 0.5 + 0.1 s for the signs s of the unit corner table.

BOX_WHD = np.array([0.2, 0.2, 0.2])
BOX_CENTER = np.array([0.5, 0.5, 0.0])
_CORNERS = np.array([[-1, -1, -1], [-1, -1, 1], [-1, 1, -1], [-1, 1, 1],
                     [1, -1, -1], [1, -1, 1], [1, 1, -1], [1, 1, 1]], dtype=np.float64)


def box_corner(k):
    return BOX_CENTER + 0.5 * BOX_WHD * _CORNERS[k]


def corner_under_triangle():
    """A triangle 1.5 mm above the bottom plane of the box around the bottom corner
    (0.4, 0.4, -0.1): corner 0 lies 1.5 mm below the triangle (a (1, 3) record with
    weights 1/3 each and the corner's angle-weighted normal (-1, -1, -1)/sqrt(3)),
    and the triangle's edge 1 -> 2 runs inside the box across the corner, 1.5 mm
    above the bottom edges 1 (along -y at x = 0.4) and 4 (along +x at y = 0.4): two
    (2, 2) records, through the one-face branch of each edge."""
    zc = -0.1 + 1.5e-3
    verts = np.array([[0.39, 0.39, zc], [0.42, 0.39, zc], [0.39, 0.42, zc]])
    tri = np.array([[0, 1, 2]])
    return verts, tri


def expected_corner_under_triangle(v2):
    a, b, c = v2[0], v2[1], v2[2]
    x1 = box_corner(0)
    n1 = unit([-1.0, -1.0, -1.0])
    n2 = face_normal(a, b, c)
    if n1 @ n2 < 0:
        n2 = -n2
    proj = (x1 - a) @ n2
    x2 = x1 - proj * n2
    out = [rec(dist=np.linalg.norm(x2 - x1), nor1=n1, nor2=n2, pos1=x1, pos2=x2, count1=1, count2=3,
               verts1=[0, -1, -1], verts2=[0, 1, 2], weights1=[1.0, 0.0, 0.0], weights2=bary(a, b, c, x1),
               tri2=0, nedge1=3, edge1=[0, 1, 4], pos1_=x1 - 0.1 * THR * n1)]
    # cloth edge 2 in createEdges order is 1 -> 2 (hash order 0-1, 2-0, 1-2)
    x2a, x2b = v2[1], v2[2]
    dx2 = x2b - x2a
    for k1, ia, ib, n1c, n1d in ((1, 2, 0, [-1.0, 0, 0], [0, 0, -1.0]), (4, 0, 4, [0, -1.0, 0], [0, 0, -1.0])):
        x1a, x1b = box_corner(ia), box_corner(ib)
        dx1 = x1b - x1a
        n1c, n1d = np.array(n1c), np.array(n1d)
        # the cloth edge crosses the face of n1c where its line meets that plane
        u2c = ((x1a - x2a) @ n1c) / (dx2 @ n1c)
        s, t = line_params(x1a, dx1, x2a, dx2)
        inside = (x2a - x1a) @ n1c < 0
        u2 = max(0.0, min(u2c, t)) if inside else max(u2c, min(1.0, t))
        p2 = (1 - u2) * x2a + u2 * x2b
        u1 = ((p2 - x1a) @ dx1) / (dx1 @ dx1)
        p1 = (1 - u1) * x1a + u1 * x1b
        nor = unit(np.cross(dx1, dx2))
        nor1 = unit(n1c + n1d)
        if nor @ nor1 < 0:
            nor = -nor
        out.append(rec(dist=np.linalg.norm(p2 - p1), nor1=nor1, nor2=nor, pos1=p1, pos2=p2, count1=2, count2=2,
                       verts1=[ia, ib, -1], verts2=[1, 2, -1], weights1=[1 - u1, u1, 0.0],
                       weights2=[1 - u2, u2, 0.0], nedge1=1, edge1=[k1], edge2=2, edgeDir=unit(dx1),
                       pos1_=p1 - 0.1 * THR * nor1))
    return out


def vertex_inside_box():
    """Cloth vertex 0 inside the box 0.5 mm above its bottom face, over the bottom
    triangle 20 (corners 0, 12, 4) at unit coordinates (0.2, -0.3): a (3, 1) record
    with weights (0.05, 0.7, 0.25).  Vertex 0 belongs to face 1 only; face 0 is a
    tilted triangle elsewhere, and the record's cloth normal is face 0's
    (faceNors2.col(i2) reads a face normal by vertex index, 731)."""
    verts = np.array([[0.52, 0.47, -0.0995],
                      [0.55, 0.45, -0.12], [0.56, 0.45, -0.13], [0.55, 0.46, -0.12],
                      [0.53, 0.47, -0.11], [0.52, 0.48, -0.11]])
    tri = np.array([[1, 2, 3], [0, 4, 5]])
    return verts, tri


def expected_vertex_inside_box(v2):
    x2 = v2[0]
    n1 = np.array([0.0, 0.0, -1.0])
    x1 = x2.copy()
    x1[2] = box_corner(0)[2]
    w = bary(box_corner(0), BOX_CENTER + np.array([0, 0, -0.1]), box_corner(4), x1)
    n2 = face_normal(v2[1], v2[2], v2[3])       # face 0, not a face of vertex 0
    if n2 @ n1 < 0:
        n2 = -n2
    return [rec(dist=np.linalg.norm(x2 - x1), nor1=n1, nor2=n2, pos1=x1, pos2=x2, count1=3, count2=1,
                verts1=[0, 12, 4], verts2=[0, -1, -1], weights1=w, weights2=[1.0, 0.0, 0.0], tri1=20,
                pos1_=x1 - 0.1 * THR * n1)]


def edge_through_box_edge():
    """Two cloth triangles sharing the edge 0 -> 1, which crosses the bottom-front
    box edge 4 (y = 0.4, z = -0.1, along +x) through the box: it enters through the
    front face and leaves through the bottom face, passing 0.5 mm from the box edge
    in both y and z at x = 0.5.  The two-face branch gives the closest points of the
    two lines; the other cloth edges stay outside the box."""
    e, L = 5e-4, 0.01
    q = np.array([0.5, 0.4 + e, -0.1 + e])
    d = np.array([0.0, 1.0, -1.0])
    a, b = q - L * d, q + L * d
    c = np.array([0.51, 0.38, -0.12])
    dd = np.array([0.49, 0.38, -0.12])
    verts = np.stack([a, b, c, dd])
    tri = np.array([[0, 1, 2], [1, 0, 3]])
    return verts, tri


def expected_edge_through_box_edge(v2):
    x2a, x2b = v2[0], v2[1]
    dx2 = x2b - x2a
    x1a, x1b = box_corner(0), box_corner(4)
    dx1 = x1b - x1a
    u1, u2 = line_params(x1a, dx1, x2a, dx2)
    p1 = (1 - u1) * x1a + u1 * x1b
    p2 = (1 - u2) * x2a + u2 * x2b
    nor = unit(np.cross(dx1, dx2))
    nor1 = unit([0.0, -1.0, -1.0])
    if nor @ nor1 < 0:
        nor = -nor
    return [rec(dist=np.linalg.norm(p2 - p1), nor1=nor1, nor2=nor, pos1=p1, pos2=p2, count1=2, count2=2,
                verts1=[0, 4, -1], verts2=[0, 1, -1], weights1=[1 - u1, u1, 0.0], weights2=[1 - u2, u2, 0.0],
                nedge1=1, edge1=[4], edge2=0, edgeDir=unit(dx1), pos1_=p1 - 0.1 * THR * nor1)]


def compare(got, exp, tol=1e-9):
    """Integer fields equal, vectors within tol (absolute; lengths here are O(1))."""
    assert len(got) == len(exp), (len(got), len(exp))
    for g, e in zip(got, exp):
        for k in ("count1", "count2", "tri1", "tri2", "edge2", "nedge1"):
            if k in e:
                assert getattr(g, k) == e[k], (k, getattr(g, k), e[k])
        assert list(g.verts1) == list(e["verts1"]), (list(g.verts1), e["verts1"])
        assert list(g.verts2) == list(e["verts2"]), (list(g.verts2), e["verts2"])
        assert list(g.edge1)[: g.nedge1] == list(e["edge1"])
        for k in ("dist", "nor1", "nor2", "pos1", "pos2", "pos1_", "weights1", "weights2", "edgeDir"):
            if k in e:
                gv = np.array(getattr(g, k)) if k != "dist" else np.array([g.dist])
                ev = np.atleast_1d(np.asarray(e[k], dtype=np.float64))
                assert np.max(np.abs(gv[: ev.size] - ev)) <= tol, (k, gv, ev)
