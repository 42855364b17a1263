"""ctypes front end of the FP64 CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by ``__graft_entry__.smoke()``
and by the CPU-baseline leg of ``bench.py``.  The product library never loads
it.  Every wrapped function restates a reference routine; see the C sources for
the file:line each one follows.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None
_REF = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


class Material(C.Structure):
    _fields_ = [("density", C.c_double), ("e", C.c_double), ("nu", C.c_double),
                ("beta", C.c_double), ("dampingA", C.c_double), ("dampingB", C.c_double)]


class Csr(C.Structure):
    _fields_ = [("n", C.c_int), ("nnz", C.c_int64), ("rowptr", C.POINTER(C.c_int64)),
                ("col", C.POINTER(C.c_int)), ("val", C.POINTER(C.c_double))]


class SolveStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("relres", C.c_double), ("true_relres", C.c_double),
                ("status", C.c_int)]


class Collision(C.Structure):
    _fields_ = [("dist", C.c_double), ("nor1", C.c_double * 3), ("nor2", C.c_double * 3),
                ("pos1", C.c_double * 3), ("pos2", C.c_double * 3), ("pos1_", C.c_double * 3),
                ("count1", C.c_int), ("count2", C.c_int), ("verts1", C.c_int * 3),
                ("verts2", C.c_int * 3), ("weights1", C.c_double * 3), ("weights2", C.c_double * 3),
                ("tri1", C.c_int), ("tri2", C.c_int), ("nedge1", C.c_int), ("edge1", C.c_int * 3),
                ("edge2", C.c_int), ("edgeDir", C.c_double * 3)]


class Edge(C.Structure):
    _fields_ = [("verts", C.c_int * 4), ("faces", C.c_int * 2), ("internal", C.c_int),
                ("angle", C.c_double), ("normals", C.c_double * 6)]


INCLUDE_BENDING_FORCE = 1
PC_JACOBI, PC_BLOCK_JACOBI = 0, 1


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle/liboracle.so is not built (run __graft_entry__.build())")
        L = C.CDLL(path)
        L.orc_grid_mesh.argtypes = [C.c_int, _dp, _ip]
        L.orc_build_hinges.argtypes = [C.c_int, C.c_int, _ip, _ip, C.c_int]
        for nm in ("orc_inertial",):
            getattr(L, nm).argtypes = [_dp] * 7 + [C.c_double, _dp, _dp, _dp]
        L.orc_velocity_transfer.argtypes = [C.c_int, _dp, _dp, C.c_int, _ip, C.c_int, _dp, _dp, _ip, _ip, _dp]
        L.orc_velocity_transfer_eol.argtypes = [C.c_int, _dp, _dp, _dp, _ip, _dp, C.c_int, _ip, C.c_int, _dp, _dp,
                                                _ip, _ip, _dp]
        L.orc_membrane.argtypes = [_dp] * 6 + [C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp]
        L.orc_bending.argtypes = [_dp] * 8 + [C.c_double, _dp, _dp, _dp]
        L.orc_face_PQ.argtypes = [_dp] * 8
        L.orc_poldec.argtypes = [_dp, _dp]
        L.orc_forces_fill.argtypes = [C.c_int, _dp, _dp, C.c_int, _ip, C.c_int, _ip,
                                      C.POINTER(Material), _dp, C.c_double, C.c_int, _dp,
                                      C.POINTER(Csr), C.POINTER(Csr), C.c_int]
        L.orc_forces_fill_eol.argtypes = [C.c_int, _dp, _dp, C.c_int, _ip, C.c_int, _ip,
                                          C.POINTER(Material), _dp, C.c_double, C.c_int, C.c_int, _ip, _dp,
                                          C.POINTER(Csr), C.POINTER(Csr), C.c_int]
        L.orc_csr_free.argtypes = [C.POINTER(Csr)]
        L.orc_csr_matvec.argtypes = [C.POINTER(Csr), _dp, _dp, C.c_int]
        L.orc_velocity_solve.argtypes = [C.POINTER(Csr), _dp, C.c_int, _ip, _dp, C.c_void_p,
                                         C.c_double, C.c_int, C.c_int, _dp, C.POINTER(SolveStats),
                                         C.c_int]
        L.orc_dense_velocity_solve.argtypes = [C.POINTER(Csr), _dp, C.c_int, _ip, _dp, _dp]
        L.orc_step.argtypes = [C.c_int, _dp, _dp, _dp, C.c_int, _ip, C.c_int, _ip,
                               C.POINTER(Material), _dp, C.c_double, C.c_int, C.c_int, _ip,
                               C.c_void_p, C.c_double, C.c_int, C.c_int, C.POINTER(SolveStats),
                               C.c_int, _dp]
        L.orc_mt19937_uniform.argtypes = [C.c_uint32, C.c_int, C.c_double, C.c_double, _dp]
        L.orc_tri3_inc.argtypes = [_dp] * 6
        L.orc_tri3_inc.restype = C.c_int
        L.orc_create_edges.argtypes = [C.c_int, C.c_int, _ip, _dp, C.POINTER(Edge)]
        L.orc_point_tri.argtypes = [C.c_double, C.c_int, _dp, _dp, C.c_int, _dp, C.c_int, _ip,
                                    C.c_int, C.POINTER(Collision), C.c_int]
        L.orc_box_tri.argtypes = [C.c_double, _dp, _dp, C.c_int, _dp, C.c_int, _ip, C.c_void_p,
                                  C.c_int, C.POINTER(Collision), C.c_int]
        L.orc_gaussian.argtypes = [C.c_uint64, C.c_int64, C.c_double, _dp]
        _LIB = L
    return _LIB


def ref_lib():
    """The reference's own element kernels compiled into oracle/_ref (None if absent)."""
    global _REF
    if _REF is None:
        path = os.path.join(_HERE, "_ref", "libref_elements.so")
        if not os.path.exists(path):
            return None
        R = C.CDLL(path)
        R.ref_membrane.argtypes = [_dp] * 6 + [C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp]
        R.ref_inertial.argtypes = [_dp] * 7 + [C.c_double, _dp, _dp, _dp]
        R.ref_bending.argtypes = [_dp] * 8 + [C.c_double, _dp, _dp, _dp]
        R.ref_tri3_inc.argtypes = [_dp] * 5 + [_dp, _dp, _dp]
        R.ref_tri3_inc.restype = C.c_int
        _REF = R
    return _REF


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


# ---------------------------------------------------------------- elements
def inertial(xa, xb, xc, Xa, Xb, Xc, g, rho, impl=None):
    L = impl or lib()
    W, f, M = np.zeros(1), np.zeros(9), np.zeros(81)
    fn = L.ref_inertial if impl is not None else L.orc_inertial
    fn(_d(xa), _d(xb), _d(xc), _d(Xa), _d(Xb), _d(Xc), _d(g), float(rho), W, f, M)
    return W[0], f, M.reshape(9, 9)


def membrane(xa, xb, xc, Xa, Xb, Xc, e, nu, P, Q, impl=None):
    L = impl or lib()
    W, f, K = np.zeros(1), np.zeros(9), np.zeros(81)
    fn = L.ref_membrane if impl is not None else L.orc_membrane
    fn(_d(xa), _d(xb), _d(xc), _d(Xa), _d(Xb), _d(Xc), float(e), float(nu), _d(P), _d(Q), W, f, K)
    return W[0], f, K.reshape(9, 9)


def bending(x0, x1, x2, x3, X0, X1, X2, X3, beta, impl=None):
    L = impl or lib()
    W, f, K = np.zeros(1), np.zeros(12), np.zeros(144)
    fn = L.ref_bending if impl is not None else L.orc_bending
    fn(_d(x0), _d(x1), _d(x2), _d(x3), _d(X0), _d(X1), _d(X2), _d(X3), float(beta), W, f, K)
    return W[0], f, K.reshape(12, 12)


def face_PQ(xa, xb, xc, Xa, Xb, Xc):
    P, Q = np.zeros(6), np.zeros(4)
    lib().orc_face_PQ(_d(xa), _d(xb), _d(xc), _d(Xa), _d(Xb), _d(Xc), P, Q)
    return P, Q


def poldec(Mcm):
    Q = np.zeros(4)
    lib().orc_poldec(_d(Mcm), Q)
    return Q


# ---------------------------------------------------------------- mesh
def grid_mesh(n):
    N, F = (n + 1) ** 2, 2 * n * n
    X, tri = np.zeros(2 * N), np.zeros(3 * F, dtype=np.int32)
    lib().orc_grid_mesh(n, X, tri)
    return X.reshape(N, 2), tri.reshape(F, 3)


def build_hinges(N, tri):
    tri = _i(tri).reshape(-1)
    F = tri.size // 3
    cap = 3 * F
    h = np.zeros(4 * cap, dtype=np.int32)
    H = lib().orc_build_hinges(int(N), F, tri, h, cap)
    return h[: 4 * H].reshape(H, 4).copy()


# ---------------------------------------------------------------- assembly
def csr_to_scipy(A):
    import scipy.sparse as sp
    n, nnz = A.n, A.nnz
    rp = np.ctypeslib.as_array(A.rowptr, shape=(n + 1,)).copy()
    col = np.ctypeslib.as_array(A.col, shape=(max(nnz, 1),))[:nnz].copy()
    val = np.ctypeslib.as_array(A.val, shape=(max(nnz, 1),))[:nnz].copy()
    return sp.csr_matrix((val, col, rp), shape=(n, n))


def material(density=0.1, e=500.0, nu=0.3, beta=1e-5, dampingA=0.0, dampingB=1.0):
    return Material(density, e, nu, beta, dampingA, dampingB)


def forces_fill(x, X, tri, hinge, mat, grav, h, flags=0, nthreads=1):
    x, X = _d(x).reshape(-1), _d(X).reshape(-1)
    tri, hinge = _i(tri).reshape(-1), _i(hinge).reshape(-1)
    N = x.size // 3
    f = np.zeros(3 * N)
    M, MDK = Csr(), Csr()
    rc = lib().orc_forces_fill(N, x, X, tri.size // 3, tri, hinge.size // 4, hinge, C.byref(mat),
                               _d(grav), float(h), int(flags), f, C.byref(M), C.byref(MDK),
                               int(nthreads))
    Ms, MDKs = csr_to_scipy(M), csr_to_scipy(MDK)
    lib().orc_csr_free(C.byref(M))
    lib().orc_csr_free(C.byref(MDK))
    return rc, f, Ms, MDKs


def forces_fill_eol(x, X, tri, hinge, mat, grav, h, eol_nodes, flags=0, nthreads=1):
    """Forces::fill with EOL nodes: f (3N + 2E), M and MDK ((3N + 2E) square)."""
    x, X = _d(x).reshape(-1), _d(X).reshape(-1)
    tri, hinge = _i(tri).reshape(-1), _i(hinge).reshape(-1)
    en = _i(eol_nodes).reshape(-1)
    N = x.size // 3
    f = np.zeros(3 * N + 2 * en.size)
    M, MDK = Csr(), Csr()
    rc = lib().orc_forces_fill_eol(N, x, X, tri.size // 3, tri, hinge.size // 4, hinge, C.byref(mat),
                                   _d(grav), float(h), int(flags), en.size, en, f, C.byref(M), C.byref(MDK),
                                   int(nthreads))
    if rc == 1:
        return rc, None, None, None
    Ms, MDKs = csr_to_scipy(M), csr_to_scipy(MDK)
    lib().orc_csr_free(C.byref(M))
    lib().orc_csr_free(C.byref(MDK))
    return rc, f, Ms, MDKs


def _to_csr_struct(A):
    A = A.tocsr()
    A.sort_indices()
    keep = (np.ascontiguousarray(A.indptr, dtype=np.int64), _i(A.indices), _d(A.data))
    s = Csr(A.shape[0], int(A.nnz), keep[0].ctypes.data_as(C.POINTER(C.c_int64)),
            keep[1].ctypes.data_as(C.POINTER(C.c_int)), keep[2].ctypes.data_as(C.POINTER(C.c_double)))
    return s, keep


def velocity_solve(A, rhs, fixed_dof, fixed_val, v0=None, tol=1e-8, maxit=100000,
                   precond=PC_BLOCK_JACOBI, nthreads=1):
    s, keep = _to_csr_struct(A)
    n = A.shape[0]
    v = np.zeros(n)
    st = SolveStats()
    v0a = None if v0 is None else _d(v0)
    lib().orc_velocity_solve(C.byref(s), _d(rhs), len(fixed_dof), _i(fixed_dof), _d(fixed_val),
                             None if v0a is None else v0a.ctypes.data_as(C.c_void_p),
                             float(tol), int(maxit), int(precond), v, C.byref(st), int(nthreads))
    del keep
    return v, st


def dense_velocity_solve(A, rhs, fixed_dof, fixed_val):
    s, keep = _to_csr_struct(A)
    v = np.zeros(A.shape[0])
    rc = lib().orc_dense_velocity_solve(C.byref(s), _d(rhs), len(fixed_dof), _i(fixed_dof),
                                        _d(fixed_val), v)
    del keep
    return rc, v


def step(x, v, X, tri, hinge, mat, grav, h, fixed_nodes, flags=0, tol=1e-8, maxit=100000,
         precond=PC_BLOCK_JACOBI, nthreads=1, scripted=None):
    """One Cloth::step; x and v (float64, 3N) are updated in place."""
    tri, hinge = _i(tri).reshape(-1), _i(hinge).reshape(-1)
    fixed_nodes = _i(fixed_nodes)
    st = SolveStats()
    ph = np.zeros(5)
    sc = None if scripted is None else _d(scripted)
    rc = lib().orc_step(x.size // 3, x, v, _d(X).reshape(-1), tri.size // 3, tri, hinge.size // 4,
                        hinge, C.byref(mat), _d(grav), float(h), int(flags), fixed_nodes.size,
                        fixed_nodes, None if sc is None else sc.ctypes.data_as(C.c_void_p),
                        float(tol), int(maxit), int(precond), C.byref(st), int(nthreads), ph)
    return rc, st, ph


# ---------------------------------------------------------------- collisions
def tri3_inc(orig, dir, v0, v1, v2):
    """The oracle's intersect_triangle3_inc: (hit, t)."""
    t = np.zeros(1)
    hit = lib().orc_tri3_inc(_d(orig), _d(dir), _d(v0), _d(v1), _d(v2), t)
    return bool(hit), float(t[0])


def mt19937_uniform(seed, count, lo=-1.0, hi=1.0):
    out = np.zeros(count)
    lib().orc_mt19937_uniform(seed, count, lo, hi, out)
    return out


def create_edges(verts, tri):
    verts, tri = _d(verts).reshape(-1), _i(tri).reshape(-1)
    F = tri.size // 3
    buf = (Edge * (3 * F + 1))()
    ne = lib().orc_create_edges(verts.size // 3, F, tri, verts, buf)
    return [buf[i] for i in range(ne)]


def _collect(fn, cap, *args):
    buf = (Collision * cap)()
    n = fn(*args, buf, cap)
    if n > cap:
        buf = (Collision * n)()
        n = fn(*args, buf, n)
    return n, [buf[i] for i in range(max(n, 0))]


def velocity_transfer(u_old, v_old, tri_old, u_new):
    """Cloth::velocityTransfer for non-EOL nodes (brute force): (v_new, src, face,
    bary, missing)."""
    u_old, v_old, u_new = _d(u_old).reshape(-1), _d(v_old).reshape(-1), _d(u_new).reshape(-1)
    tri_old = _i(tri_old).reshape(-1)
    n_new = u_new.size // 2
    v_new = np.zeros(3 * n_new)
    src = np.zeros(n_new, dtype=np.int32)
    face = np.zeros(n_new, dtype=np.int32)
    bary = np.zeros(3 * n_new)
    missing = lib().orc_velocity_transfer(u_old.size // 2, u_old, v_old, tri_old.size // 3, tri_old, n_new, u_new,
                                          v_new, src, face, bary)
    return v_new, src, face, bary, missing


def velocity_transfer_eol(u_old, x_old, v_old, eidx, V_old, tri_old, u_new):
    """velocity_transfer with EOL nodes in the old mesh (eidx: EoL index per old
    vertex or -1; V_old: 2 per EOL node): (v_new, src, face, bary, missing)."""
    u_old, x_old, v_old = _d(u_old).reshape(-1), _d(x_old).reshape(-1), _d(v_old).reshape(-1)
    u_new, V_old = _d(u_new).reshape(-1), _d(V_old).reshape(-1)
    if V_old.size == 0:
        V_old = np.zeros(2)
    tri_old, eidx = _i(tri_old).reshape(-1), _i(eidx).reshape(-1)
    n_new = u_new.size // 2
    v_new = np.zeros(3 * n_new)
    src = np.zeros(n_new, dtype=np.int32)
    face = np.zeros(n_new, dtype=np.int32)
    bary = np.zeros(3 * n_new)
    missing = lib().orc_velocity_transfer_eol(u_old.size // 2, u_old, x_old, v_old, eidx, V_old, tri_old.size // 3,
                                              tri_old, n_new, u_new, v_new, src, face, bary)
    return v_new, src, face, bary, missing


def point_tri(thr, pts, nrm, verts, tri, eol=False, cap=4096):
    pts, nrm = _d(pts).reshape(-1), _d(nrm).reshape(-1)
    verts, tri = _d(verts).reshape(-1), _i(tri).reshape(-1)
    return _collect(lib().orc_point_tri, cap, float(thr), pts.size // 3, pts, nrm,
                    verts.size // 3, verts, tri.size // 3, tri, int(eol))


def box_tri(thr, whd, E, verts, tri, eol=False, cap=65536):
    """E: 4x4 box frame (row-major numpy); passed column-major as Eigen stores it."""
    Ecm = _d(np.asarray(E, dtype=np.float64).T).reshape(-1)
    verts, tri = _d(verts).reshape(-1), _i(tri).reshape(-1)
    return _collect(lib().orc_box_tri, cap, float(thr), _d(whd), Ecm, verts.size // 3, verts,
                    tri.size // 3, tri, None, int(eol))


def gaussian(seed, count, sigma):
    out = np.zeros(count)
    lib().orc_gaussian(C.c_uint64(seed), int(count), float(sigma), out)
    return out
