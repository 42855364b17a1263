"""Python binding of the eolc C ABI (include/eolc.h) over ctypes.

The library (``libeolc.so``, built in-tree by ``make -C eol-cloth_amd``) holds the
gfx950 kernels; this module only moves host arrays across the ABI.  There is no
CPU fallback: every compute call goes to the GPU, and loading fails loudly when
the library is missing.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libeolc.so")

(OK, ERR_INVALID, ERR_HIP, ERR_NOT_SPD, ERR_MAXIT, ERR_OVERFLOW, ERR_NONFINITE, ERR_DEGENERATE, ERR_NO_DEVICE,
 ERR_NO_MEMORY) = range(10)
STATUS_NAMES = {0: "OK", 1: "INVALID", 2: "HIP", 3: "NOT_SPD", 4: "MAXIT", 5: "OVERFLOW",
                6: "NONFINITE", 7: "DEGENERATE", 8: "NO_DEVICE", 9: "NO_MEMORY"}

OPT_INCLUDE_BENDING_FORCE = 0x1
OPT_SPD_PROJECT_BENDING = 0x2

SOLVE_FP64_BJ, SOLVE_MIXED_BJ, SOLVE_MIXED_LEARNED, SOLVE_FP64_JACOBI = 0, 1, 2, 3
SOLVE_MULTILEVEL = 0x10   # flag: add the aggregation levels of the multilevel preconditioner
ML_MAX_LEVELS = 8
MAT_M, MAT_MDK = 0, 1

EXPORTED_SYMBOLS = (
    "eolc_create", "eolc_destroy", "eolc_last_error", "eolc_set_topology", "eolc_set_material",
    "eolc_set_options", "eolc_upload_state", "eolc_download_state", "eolc_forces_fill",
    "eolc_download_forces", "eolc_download_csr", "eolc_velocity_solve", "eolc_step", "eolc_collide",
    "eolc_time_spmv", "eolc_learned_set_weights", "eolc_learned_get_weights", "eolc_learned_infer",
    "eolc_learned_loss_grad", "eolc_learned_train", "eolc_autotune", "eolc_set_spmv_variant",
    "eolc_set_batch_topology", "eolc_set_batch_material", "eolc_batch_step", "eolc_batch_autotune",
    "eolc_set_batch_variant", "eolc_stream_copy", "eolc_set_mixed_delta", "eolc_collision_constraints",
    "eolc_ml_info", "eolc_ml_aggregates", "eolc_precond_apply", "eolc_ml_partition", "eolc_set_batch_multilevel",
    "eolc_velocity_transfer", "eolc_stream_read", "eolc_step_dofs", "eolc_set_batch_scripted",
    "eolc_rigid_integrate", "eolc_box_step", "eolc_set_eol", "eolc_velocity_transfer_eol",
    "eolc_collide_noalloc", "eolc_learned_set_multilevel", "eolc_set_mixed_stall", "eolc_spmv_layout_info",
    "eolc_set_stiff_threshold", "eolc_set_ml_linear",
)

BATCH_EAGER = 0x1
BATCH_VARIANTS = 3

SPMV_VARIANTS = 29

LEARNED_NPARAMS = 26246


class SolveStats(C.Structure):
    _fields_ = [("iters", C.c_int), ("refinements", C.c_int), ("status", C.c_int),
                ("used_fallback", C.c_int), ("relres", C.c_double), ("true_relres", C.c_double),
                ("ms_assembly", C.c_double), ("ms_rhs", C.c_double), ("ms_solve", C.c_double),
                ("ms_update", C.c_double), ("launched", C.c_int), ("host_syncs", C.c_int)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class CollisionRec(C.Structure):
    _fields_ = [("dist", C.c_double), ("nor1", C.c_double * 3), ("nor2", C.c_double * 3),
                ("pos1", C.c_double * 3), ("pos2", C.c_double * 3), ("pos1_", C.c_double * 3),
                ("count1", C.c_int), ("count2", C.c_int), ("verts1", C.c_int * 3),
                ("verts2", C.c_int * 3), ("weights1", C.c_double * 3), ("weights2", C.c_double * 3),
                ("tri1", C.c_int), ("tri2", C.c_int), ("nedge1", C.c_int), ("edge1", C.c_int * 3),
                ("edge2", C.c_int), ("edgeDir", C.c_double * 3)]


class EolcError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"eolc error {STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


_lib = None
_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")


def lib():
    """Load libeolc.so (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is not built; run __graft_entry__.build()")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.eolc_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.eolc_destroy.argtypes = [vp]
        L.eolc_destroy.restype = None
        L.eolc_last_error.argtypes = [vp]
        L.eolc_last_error.restype = C.c_char_p
        L.eolc_set_topology.argtypes = [vp, C.c_int, _dp, C.c_int, _ip, C.c_int, _ip]
        L.eolc_set_material.argtypes = [vp] + [C.c_double] * 6
        L.eolc_set_options.argtypes = [vp, C.c_int]
        L.eolc_upload_state.argtypes = [vp, _dp, _dp]
        L.eolc_download_state.argtypes = [vp, _dp, _dp]
        L.eolc_forces_fill.argtypes = [vp, _dp, C.c_double]
        L.eolc_download_forces.argtypes = [vp, _dp]
        L.eolc_download_csr.argtypes = [vp, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                        C.POINTER(C.c_int64)]
        L.eolc_velocity_solve.argtypes = [vp, C.c_int, C.c_void_p, C.c_void_p, C.c_double, C.c_int,
                                          C.c_int, C.POINTER(SolveStats)]
        L.eolc_step.argtypes = [vp, _dp, C.c_double, C.c_int, C.c_void_p, C.c_void_p, C.c_double,
                                C.c_int, C.c_int, C.POINTER(SolveStats)]
        L.eolc_collide.argtypes = [vp, C.c_double, C.c_int, C.c_void_p, C.c_void_p, C.c_int,
                                   C.c_void_p, C.c_void_p, C.c_int, C.c_int, C.c_int,
                                   C.POINTER(CollisionRec), C.c_int, C.POINTER(C.c_int)]
        L.eolc_collide_noalloc.argtypes = L.eolc_collide.argtypes
        L.eolc_time_spmv.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_double)]
        L.eolc_spmv_layout_info.argtypes = [vp, C.POINTER(C.c_int64), C.c_int]
        L.eolc_set_stiff_threshold.argtypes = [vp, C.c_double]
        L.eolc_set_ml_linear.argtypes = [vp, C.c_int]
        L.eolc_set_batch_topology.argtypes = [vp, C.c_int] + [C.c_void_p] * 9
        L.eolc_set_batch_material.argtypes = [vp, C.c_void_p, C.c_void_p]
        L.eolc_batch_step.argtypes = [vp, _dp, C.c_double, C.c_double, C.c_int, C.c_int, C.c_void_p]
        L.eolc_batch_autotune.argtypes = [vp, _dp, C.c_double, C.c_double, C.c_int, C.POINTER(C.c_int),
                                          C.c_void_p]
        L.eolc_set_batch_variant.argtypes = [vp, C.c_int]
        L.eolc_set_batch_multilevel.argtypes = [vp, C.c_int]
        L.eolc_set_batch_scripted.argtypes = [vp, C.c_void_p]
        L.eolc_set_eol.argtypes = [vp, C.c_int, C.c_void_p]
        L.eolc_velocity_transfer_eol.argtypes = [vp, C.c_int] + [C.c_void_p] * 4 + [C.c_int, C.c_void_p, C.c_int,
                                                                                    C.c_void_p, C.c_int] + \
            [C.c_void_p] * 5 + [C.POINTER(C.c_double)]
        L.eolc_step_dofs.argtypes = [vp, _dp, C.c_double, C.c_int, C.c_void_p, C.c_void_p, C.c_double,
                                     C.c_int, C.c_int, C.POINTER(SolveStats)]
        L.eolc_rigid_integrate.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]
        L.eolc_box_step.argtypes = [C.c_void_p, C.c_void_p, C.c_double, C.c_void_p]
        L.eolc_velocity_transfer.argtypes = [vp, C.c_int, _dp, _dp, C.c_int, _ip, C.c_int, _dp, _dp, _ip, _ip, _dp,
                                             C.POINTER(C.c_double)]
        L.eolc_set_mixed_delta.argtypes = [vp, C.c_double]
        L.eolc_collision_constraints.argtypes = [vp, C.c_void_p, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                                 C.c_void_p, C.c_void_p, C.c_int, C.c_int64, C.POINTER(C.c_int),
                                                 C.POINTER(C.c_int64)]
        L.eolc_stream_copy.argtypes = [vp, C.c_int64, C.c_int, C.POINTER(C.c_double)]
        L.eolc_stream_read.argtypes = [vp, C.c_int64, C.c_int, C.POINTER(C.c_double)]
        L.eolc_autotune.argtypes = [vp, C.c_int, C.c_int, C.POINTER(C.c_int), C.c_void_p]
        L.eolc_set_spmv_variant.argtypes = [vp, C.c_int, C.c_int]
        L.eolc_ml_info.argtypes = [vp, C.POINTER(C.c_int), C.c_void_p]
        L.eolc_ml_aggregates.argtypes = [vp, C.c_int, C.c_void_p]
        L.eolc_precond_apply.argtypes = [vp, C.c_int, C.c_void_p, C.c_int, _dp, _dp]
        L.eolc_ml_partition.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.POINTER(C.c_int), C.c_void_p,
                                        C.c_void_p]
        L.eolc_learned_set_weights.argtypes = [vp, C.c_void_p, C.c_int]
        L.eolc_learned_set_multilevel.argtypes = [vp, C.c_int]
        L.eolc_set_mixed_stall.argtypes = [vp, C.c_int, C.c_int, C.c_int]
        L.eolc_learned_get_weights.argtypes = [vp, C.c_void_p, C.c_int]
        L.eolc_learned_infer.argtypes = [vp, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]
        L.eolc_learned_loss_grad.argtypes = [vp, C.c_int, C.c_void_p, C.c_void_p,
                                             C.POINTER(C.c_double), C.c_void_p]
        L.eolc_learned_train.argtypes = [vp, C.c_int, C.c_void_p, C.c_void_p, C.c_int, C.c_double,
                                         C.c_void_p]
        _lib = L
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64).reshape(-1)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32).reshape(-1)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def ml_partition(rowptr, col):
    """Multilevel aggregation hierarchy of a vertex block pattern (host code, no GPU):
    list of per-vertex aggregate maps for levels 1..L."""
    rp, cl = _i(rowptr), _i(col)
    N = rp.size - 1
    L = C.c_int(0)
    cnt = np.zeros(ML_MAX_LEVELS + 1, dtype=np.int32)
    agg = np.zeros(ML_MAX_LEVELS * N, dtype=np.int32)
    rc = lib().eolc_ml_partition(N, _ptr(rp), _ptr(cl), C.byref(L), _ptr(cnt), _ptr(agg))
    if rc != OK:
        raise EolcError(rc, "ml_partition: bad pattern")
    return [agg[l * N:(l + 1) * N].copy() for l in range(L.value)]


def _frame(E):
    """4x4 frame (array-like, row-major as numpy indexes it) -> Eigen's column-major 16."""
    return np.ascontiguousarray(np.asarray(E, dtype=np.float64).reshape(4, 4).T).reshape(-1)


def rigid_integrate(E0, phi, h):
    """Rigid::integrate on the host: E0 (4x4) times the exponential of h * phi^,
    phi = (w, v) a body-frame twist.  Returns the 4x4 frame."""
    out = np.zeros(16)
    rc = lib().eolc_rigid_integrate(_ptr(_frame(E0)), _ptr(_d(phi)), C.c_double(h), _ptr(out))
    if rc != OK:
        raise EolcError(rc, "rigid_integrate: bad arguments")
    return out.reshape(4, 4).T.copy()


def box_step(E, vel, h):
    """Box::step on the host: the frame (4x4) after one step with Box::v = vel
    (angular, linear)."""
    out = np.zeros(16)
    rc = lib().eolc_box_step(_ptr(_frame(E)), _ptr(_d(vel)), C.c_double(h), _ptr(out))
    if rc != OK:
        raise EolcError(rc, "box_step: bad arguments")
    return out.reshape(4, 4).T.copy()


class Context:
    """One eolc_ctx: device buffers for one cloth on one GPU."""

    def __init__(self, device=0):
        self._h = C.c_void_p()
        rc = lib().eolc_create(int(device), C.byref(self._h))
        if rc != OK:
            raise EolcError(rc, "eolc_create failed (no usable GPU?)")
        self.N = 0
        self.n_eol = 0

    def close(self):
        if self._h:
            lib().eolc_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc, allow=()):
        if rc != OK and rc not in allow:
            raise EolcError(rc, lib().eolc_last_error(self._h).decode())
        return rc

    def last_error(self):
        return lib().eolc_last_error(self._h).decode()

    def set_topology(self, X, tri, hinge):
        X, tri, hinge = _d(X), _i(tri), _i(hinge)
        self.N = X.size // 2
        self.n_eol = 0
        return self._check(lib().eolc_set_topology(self._h, self.N, X, tri.size // 3, tri,
                                                   hinge.size // 4, hinge))

    def set_material(self, density=0.1, e=500.0, nu=0.3, beta=1e-5, dampingA=0.0, dampingB=1.0):
        return self._check(lib().eolc_set_material(self._h, density, e, nu, beta, dampingA, dampingB))

    def set_options(self, flags):
        return self._check(lib().eolc_set_options(self._h, int(flags)))

    def upload_state(self, x, v):
        return self._check(lib().eolc_upload_state(self._h, _d(x), _d(v)))

    def download_state(self):
        x, v = np.zeros(3 * self.N), np.zeros(3 * self.N)
        self._check(lib().eolc_download_state(self._h, x, v))
        return x, v

    def forces_fill(self, grav, h):
        return self._check(lib().eolc_forces_fill(self._h, _d(grav), float(h)))

    def set_eol(self, nodes=()):
        """EOL nodes (EoL_index = position in `nodes`): f, M and MDK get 2 rows per
        node after the next forces_fill; empty clears them."""
        nd = _i(nodes).reshape(-1)
        self._check(lib().eolc_set_eol(self._h, nd.size, _ptr(nd) if nd.size else None))
        self.n_eol = int(nd.size)

    def download_forces(self):
        f = np.zeros(3 * self.N + 2 * self.n_eol)
        self._check(lib().eolc_download_forces(self._h, f))
        return f

    def csr_nnz(self, which):
        """Scalar nonzeros of M or MDK (9 per stored 3x3 block)."""
        nnz = C.c_int64(0)
        rc = lib().eolc_download_csr(self._h, int(which), None, None, None, 0, C.byref(nnz))
        self._check(rc, allow=(ERR_OVERFLOW,))
        return nnz.value

    def download_csr(self, which):
        """Scalar CSR (scipy.sparse.csr_matrix) of M or MDK."""
        import scipy.sparse as sp
        nnz = C.c_int64(0)
        rc = lib().eolc_download_csr(self._h, int(which), None, None, None, 0, C.byref(nnz))
        self._check(rc, allow=(ERR_OVERFLOW,))
        n = 3 * self.N + 2 * self.n_eol
        rp = np.zeros(n + 1, dtype=np.int64)
        col = np.zeros(max(nnz.value, 1), dtype=np.int32)
        val = np.zeros(max(nnz.value, 1))
        self._check(lib().eolc_download_csr(self._h, int(which), _ptr(rp), _ptr(col), _ptr(val),
                                            nnz.value, C.byref(nnz)))
        return sp.csr_matrix((val[: nnz.value], col[: nnz.value], rp), shape=(n, n))

    def velocity_solve(self, fixed_dof=(), fixed_val=(), tol=1e-8, maxit=100000,
                       mode=SOLVE_FP64_BJ, allow=()):
        fd, fv = _i(fixed_dof), _d(fixed_val)
        st = SolveStats()
        rc = lib().eolc_velocity_solve(self._h, fd.size, _ptr(fd) if fd.size else None,
                                       _ptr(fv) if fv.size else None, float(tol), int(maxit),
                                       int(mode), C.byref(st))
        self._check(rc, allow=allow)
        return rc, st

    def step(self, grav, h, fixed_nodes=(), scripted=None, tol=1e-8, maxit=100000,
             mode=SOLVE_FP64_BJ, allow=()):
        fn = _i(fixed_nodes)
        sc = None if scripted is None else _d(scripted)
        st = SolveStats()
        rc = lib().eolc_step(self._h, _d(grav), float(h), fn.size, _ptr(fn) if fn.size else None,
                             _ptr(sc), float(tol), int(maxit), int(mode), C.byref(st))
        self._check(rc, allow=allow)
        return rc, st

    def step_dofs(self, grav, h, fixed_dofs=(), scripted=None, tol=1e-8, maxit=100000,
                  mode=SOLVE_FP64_BJ, allow=()):
        """eolc_step with single-DOF fixed rows (scripted: one velocity per DOF)."""
        fd = _i(fixed_dofs)
        sc = None if scripted is None else _d(scripted)
        st = SolveStats()
        rc = lib().eolc_step_dofs(self._h, _d(grav), float(h), fd.size, _ptr(fd) if fd.size else None,
                                  _ptr(sc), float(tol), int(maxit), int(mode), C.byref(st))
        self._check(rc, allow=allow)
        return rc, st

    def velocity_transfer(self, u_old, v_old, tri_old, u_new, allow=()):
        """Cloth::velocityTransfer (non-EOL) on the GPU: (rc, v_new, src, face, bary, ms)."""
        u_old, v_old, u_new = _d(u_old), _d(v_old), _d(u_new)
        tri_old = _i(tri_old)
        n_new = u_new.size // 2
        v_new = np.zeros(3 * n_new)
        src = np.zeros(max(n_new, 1), dtype=np.int32)
        face = np.zeros(max(n_new, 1), dtype=np.int32)
        bary = np.zeros(max(3 * n_new, 1))
        ms = C.c_double(0)
        rc = lib().eolc_velocity_transfer(self._h, u_old.size // 2, u_old, v_old, tri_old.size // 3, tri_old, n_new,
                                          u_new, v_new, src, face, bary, C.byref(ms))
        self._check(rc, allow=allow)
        return rc, v_new, src[:n_new], face[:n_new], bary[:3 * n_new], ms.value

    def velocity_transfer_eol(self, u_old, x_old, v_old, eol_old, V_old, tri_old, u_new, allow=()):
        """velocity_transfer with EOL nodes in the old mesh (eol_old: EoL index per old
        vertex or -1; V_old: 2 per EOL node): (rc, v_new, src, face, bary, ms)."""
        u_old, x_old, v_old, u_new = _d(u_old), _d(x_old), _d(v_old), _d(u_new)
        eol_old, tri_old = _i(eol_old).reshape(-1), _i(tri_old)
        V = _d(V_old)
        n_eol = V.size // 2
        if V.size == 0:
            V = np.zeros(2)
        n_new = u_new.size // 2
        v_new = np.zeros(3 * n_new)
        src = np.zeros(max(n_new, 1), dtype=np.int32)
        face = np.zeros(max(n_new, 1), dtype=np.int32)
        bary = np.zeros(max(3 * n_new, 1))
        ms = C.c_double(0)
        rc = lib().eolc_velocity_transfer_eol(self._h, u_old.size // 2, _ptr(u_old), _ptr(x_old), _ptr(v_old),
                                              _ptr(eol_old), n_eol, _ptr(V), tri_old.size // 3, _ptr(tri_old), n_new,
                                              _ptr(u_new), _ptr(v_new), _ptr(src), _ptr(face), _ptr(bary),
                                              C.byref(ms))
        self._check(rc, allow=allow)
        return rc, v_new, src[:n_new], face[:n_new], bary[:3 * n_new], ms.value

    def ml_info(self):
        """Aggregates per level of the multilevel hierarchy: [N, n_1, ..., n_L]."""
        L = C.c_int(0)
        cnt = np.zeros(ML_MAX_LEVELS + 1, dtype=np.int32)
        self._check(lib().eolc_ml_info(self._h, C.byref(L), _ptr(cnt)))
        return [int(v) for v in cnt[: L.value + 1]]

    def ml_aggregates(self, level):
        agg = np.zeros(self.N, dtype=np.int32)
        self._check(lib().eolc_ml_aggregates(self._h, int(level), _ptr(agg)))
        return agg

    def precond_apply(self, r, fixed_dof=(), mode=SOLVE_FP64_BJ):
        """z = P r with the preconditioner of `mode` for the last forces_fill."""
        fd, r = _i(fixed_dof), _d(r)
        z = np.zeros(3 * self.N)
        self._check(lib().eolc_precond_apply(self._h, fd.size, _ptr(fd) if fd.size else None, int(mode), r, z))
        return z

    def time_spmv(self, precision, reps=20):
        ms = C.c_double(0)
        self._check(lib().eolc_time_spmv(self._h, int(precision), int(reps), C.byref(ms)))
        return ms.value

    def spmv_layout_info(self):
        """Patch layout sizes: applies, records read per product, lower entries,
        storage positions, cross blocks, LDS slots of the largest patch, window
        vertices over all patches, largest window."""
        info = (C.c_int64 * 8)()
        self._check(lib().eolc_spmv_layout_info(self._h, info, 8))
        keys = ("applies", "records", "lower_entries", "positions", "cross_blocks", "lds_slots", "window_vertices",
                "window_max")
        return {k: int(v) for k, v in zip(keys, info)}

    def collide(self, threshold, pts=None, nrm=None, whd=None, E=None, eol_points=False,
                box_npts=8, box_nedges=12, cap=1 << 16, noalloc=False):
        """CD2 (eol_points=False) or CD (True).  E: list of 4x4 row-major frames.
        noalloc: eolc_collide_noalloc (no device allocation, the library's scratch)."""
        fn = lib().eolc_collide_noalloc if noalloc else lib().eolc_collide
        pts = _d(pts if pts is not None else np.zeros((0, 3)))
        nrm = _d(nrm if nrm is not None else np.zeros((0, 3)))
        whd = _d(whd if whd is not None else np.zeros((0, 3)))
        Es = np.zeros((0,)) if E is None or len(E) == 0 else \
            np.concatenate([np.asarray(e, dtype=np.float64).T.reshape(-1) for e in E])
        Es = _d(Es)
        P, B = pts.size // 3, whd.size // 3
        while True:
            buf = (CollisionRec * cap)()
            cnt = C.c_int(0)
            rc = fn(self._h, float(threshold), P, _ptr(pts), _ptr(nrm), B, _ptr(whd),
                    _ptr(Es), int(bool(eol_points)), int(box_npts), int(box_nedges),
                    buf, cap, C.byref(cnt))
            if rc == ERR_OVERFLOW and cnt.value > cap:
                cap = cnt.value
                continue
            self._check(rc)
            return [buf[i] for i in range(cnt.value)]

    def set_mixed_stall(self, first_cycle=10000, min_cycle=1000, sticky=8):
        """Stall limits of the mixed modes (FP64 continuation after an FP32 stall)."""
        return self._check(lib().eolc_set_mixed_stall(self._h, int(first_cycle), int(min_cycle), int(sticky)))

    def set_ml_linear(self, where):
        """Linear coarse fields on the multilevel levels >= 3: 0 never, 1 in FP32
        iterations only (default), 2 in every mode (eolc_set_ml_linear)."""
        self._check(lib().eolc_set_ml_linear(self._h, int(where)))

    def set_stiff_threshold(self, rel):
        self._check(lib().eolc_set_stiff_threshold(self._h, float(rel)))

    def set_mixed_delta(self, delta):
        return self._check(lib().eolc_set_mixed_delta(self._h, float(delta)))

    def stream_read(self, nbytes, reps=10):
        """Average ms of one streaming read of nbytes."""
        ms = C.c_double(0)
        self._check(lib().eolc_stream_read(self._h, int(nbytes), int(reps), C.byref(ms)))
        return ms.value

    def stream_copy(self, nbytes, reps=10):
        """Average ms of one copy-kernel launch reading and writing nbytes each."""
        ms = C.c_double(0)
        self._check(lib().eolc_stream_copy(self._h, int(nbytes), int(reps), C.byref(ms)))
        return ms.value

    def autotune(self, precision, reps=10):
        """(best variant, average ms of every SpMV variant) on the current matrix."""
        best = C.c_int(0)
        ms = np.zeros(SPMV_VARIANTS)
        self._check(lib().eolc_autotune(self._h, int(precision), int(reps), C.byref(best), _ptr(ms)))
        return best.value, ms

    def set_spmv_variant(self, precision, variant):
        return self._check(lib().eolc_set_spmv_variant(self._h, int(precision), int(variant)))

    # ---- batched cloths
    def set_batch_topology(self, Xs, tris, hinges, fixeds):
        """Per-instance lists of arrays (local indices).  Returns vert_off (B+1)."""
        B = len(Xs)
        Nb = _i([np.asarray(x).reshape(-1, 2).shape[0] for x in Xs])
        Fb = _i([np.asarray(t).reshape(-1, 3).shape[0] for t in tris])
        Hb = _i([np.asarray(h).reshape(-1, 4).shape[0] for h in hinges])
        Pb = _i([np.asarray(f).reshape(-1).size for f in fixeds])
        X = _d(np.concatenate([np.asarray(x, dtype=np.float64).reshape(-1) for x in Xs]))
        T = _i(np.concatenate([np.asarray(t).reshape(-1) for t in tris]))
        Hh = _i(np.concatenate([np.asarray(h).reshape(-1) for h in hinges]))
        P = _i(np.concatenate([np.asarray(f).reshape(-1) for f in fixeds]))
        off = np.zeros(B + 1, dtype=np.int32)
        self._check(lib().eolc_set_batch_topology(self._h, B, _ptr(Nb), _ptr(X), _ptr(Fb), _ptr(T), _ptr(Hb),
                                                  _ptr(Hh), _ptr(Pb), _ptr(P), _ptr(off)))
        self.N = int(off[-1])
        self.n_eol = 0
        self.B = B
        return off

    def set_batch_material(self, E, beta):
        E, beta = _d(E), _d(beta)
        return self._check(lib().eolc_set_batch_material(self._h, _ptr(E), _ptr(beta)))

    def batch_step(self, grav, h, tol=1e-8, maxit=100000, eager=False, allow=()):
        st = (SolveStats * self.B)()
        rc = lib().eolc_batch_step(self._h, _d(grav), float(h), float(tol), int(maxit),
                                   BATCH_EAGER if eager else 0, st)
        self._check(rc, allow=allow)
        return rc, [st[b] for b in range(self.B)]

    def batch_autotune(self, grav, h, tol=1e-8, maxit=100000):
        best = C.c_int(0)
        ms = np.zeros(BATCH_VARIANTS)
        self._check(lib().eolc_batch_autotune(self._h, _d(grav), float(h), float(tol), int(maxit), C.byref(best),
                                              _ptr(ms)))
        return best.value, ms

    def set_batch_scripted(self, vel=None):
        """Scripted velocities of the batch's fixed nodes (3 per node, the order of
        set_batch_topology's fixed lists); None resets them to zero."""
        v = None if vel is None else _d(vel)
        return self._check(lib().eolc_set_batch_scripted(self._h, _ptr(v)))

    def set_batch_multilevel(self, on=True):
        """Multilevel coarse levels in the batched persistent solve."""
        return self._check(lib().eolc_set_batch_multilevel(self._h, int(bool(on))))

    def set_batch_variant(self, variant):
        return self._check(lib().eolc_set_batch_variant(self._h, int(variant)))

    # ---- learned block preconditioner
    def learned_set_weights(self, theta):
        th = np.ascontiguousarray(theta, dtype=np.float32).reshape(-1)
        return self._check(lib().eolc_learned_set_weights(self._h, _ptr(th), th.size))

    def learned_get_weights(self):
        th = np.zeros(LEARNED_NPARAMS, dtype=np.float32)
        self._check(lib().eolc_learned_get_weights(self._h, _ptr(th), th.size))
        return th

    def learned_infer(self, fixed_dof=()):
        """(o[N, 6], P[N, 3, 3]) of the current system."""
        fd = _i(fixed_dof)
        o = np.zeros((self.N, 6), dtype=np.float32)
        P = np.zeros((self.N, 3, 3), dtype=np.float32)
        self._check(lib().eolc_learned_infer(self._h, fd.size, _ptr(fd) if fd.size else None,
                                             _ptr(o), _ptr(P)))
        return o, P

    def learned_loss_grad(self, e, fixed_dof=()):
        fd, e = _i(fixed_dof), _d(e)
        loss = C.c_double(0)
        g = np.zeros(LEARNED_NPARAMS, dtype=np.float32)
        self._check(lib().eolc_learned_loss_grad(self._h, fd.size, _ptr(fd) if fd.size else None,
                                                 _ptr(e), C.byref(loss), _ptr(g)))
        return loss.value, g

    def learned_set_multilevel(self, on=True):
        """Train the blocks with the multilevel coarse corrections in the loss."""
        return self._check(lib().eolc_learned_set_multilevel(self._h, int(bool(on))))

    def learned_train(self, e, steps, lr, fixed_dof=()):
        fd, e = _i(fixed_dof), _d(e)
        losses = np.zeros(int(steps))
        self._check(lib().eolc_learned_train(self._h, fd.size, _ptr(fd) if fd.size else None,
                                             _ptr(e), int(steps), float(lr), _ptr(losses)))
        return losses


def records_array(records):
    """A contiguous ctypes array of CollisionRec from a list of records."""
    return (CollisionRec * max(len(records), 1))(*records)


def collision_constraints(records, N, eol=None):
    """Constraints::fill collision rows from CD2 records: scipy CSR A_ineq (rows x 3N).
    Host-only (no GPU context needed)."""
    import scipy.sparse as sp
    n = len(records)
    if isinstance(records, C.Array):   # already contiguous (e.g. records_array)
        buf = records if n else (CollisionRec * 1)()
    else:
        buf = (CollisionRec * max(n, 1))(*records) if n else (CollisionRec * 1)()
    e = None if eol is None else np.ascontiguousarray(eol, dtype=np.uint8)
    rows, nnz = C.c_int(0), C.c_int64(0)
    rc = lib().eolc_collision_constraints(None, buf, n, int(N), _ptr(e), None, None, None, 0, 0, C.byref(rows),
                                          C.byref(nnz))
    if rc not in (OK, ERR_OVERFLOW):
        raise EolcError(rc, "collision_constraints failed")
    rp = np.zeros(rows.value + 1, dtype=np.int64)
    col = np.zeros(max(nnz.value, 1), dtype=np.int32)
    val = np.zeros(max(nnz.value, 1))
    rc = lib().eolc_collision_constraints(None, buf, n, int(N), _ptr(e), _ptr(rp), _ptr(col), _ptr(val), rows.value,
                                          nnz.value, C.byref(rows), C.byref(nnz))
    if rc != OK:
        raise EolcError(rc, "collision_constraints failed")
    return sp.csr_matrix((val[: nnz.value], col[: nnz.value], rp), shape=(rows.value, 3 * int(N)))
