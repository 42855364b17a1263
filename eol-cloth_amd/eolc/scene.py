"""Synthetic cloth scenes for the benchmark configurations (host-side input generation).

* ``grid(n)``: n x n quads over the material square [0,1]^2, vertex (i, j) ->
  j*(n+1)+i, checkerboard diagonals (quad (i, j) split along v00-v11 when i+j is
  even), every face counter-clockwise in X.
* ``hinges(N, tri)``: the interior edges in the order and orientation ArcSim gives
  them (external/ArcSim/mesh.cpp:356-375: faces added in order, edge (v[i], v[i+1])
  created by the first face that contains it, adjf[0] = the face traversing
  n[0] -> n[1]), as (n[0], n[1], apex of adjf[0], apex of adjf[1]) --
  the vertex order edgeBasedF hands to ComputeBending (Forces.cpp:692-697).
* ``gaussian(seed, count, sigma)``: splitmix64 -> Box-Muller, the portable stream
  every benchmark input is drawn from.
* ``uniform(seed, count)``: splitmix64 -> 53-bit uniform in [0, 1).
"""
import numpy as np

GOLDEN = np.uint64(0x9E3779B97F4A7C15)


def grid(n):
    n = int(n)
    np1 = n + 1
    j, i = np.meshgrid(np.arange(np1), np.arange(np1), indexing="ij")
    X = np.stack([i.reshape(-1) / n, j.reshape(-1) / n], axis=1).astype(np.float64)
    jj, ii = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    ii, jj = ii.reshape(-1), jj.reshape(-1)
    v00 = jj * np1 + ii
    v10, v01 = v00 + 1, v00 + np1
    v11 = v01 + 1
    even = ((ii + jj) & 1) == 0
    t1 = np.where(even[:, None], np.stack([v00, v10, v11], 1), np.stack([v00, v10, v01], 1))
    t2 = np.where(even[:, None], np.stack([v00, v11, v01], 1), np.stack([v10, v11, v01], 1))
    tri = np.stack([t1, t2], axis=1).reshape(-1, 3).astype(np.int32)
    return X, tri


def hinges(N, tri):
    tri = np.asarray(tri, dtype=np.int64).reshape(-1, 3)
    F = tri.shape[0]
    a = tri.reshape(-1)
    b = tri[:, [1, 2, 0]].reshape(-1)
    r = np.arange(3 * F)
    lo, hi = np.minimum(a, b), np.maximum(a, b)
    key = lo * int(N) + hi
    order = np.lexsort((r, key))
    ks = key[order]
    start = np.ones(len(ks), dtype=bool)
    start[1:] = ks[1:] != ks[:-1]
    gid = np.cumsum(start) - 1
    leader = order[start]                      # creating record of each edge, sorted by key
    lead_of = leader[gid]                      # per sorted record
    n0 = a[lead_of]
    side = np.where(a[order] == n0, 0, 1)
    face = order // 3
    ng = len(leader)
    adj = np.full((ng, 2), -1, dtype=np.int64)
    # later faces overwrite (e->adjf[side] = face): take the last record per (edge, side)
    for s in (0, 1):
        m = side == s
        g, f = gid[m], face[m]
        last = np.full(ng, -1, dtype=np.int64)
        np.maximum.at(last, g, order[m])
        adj[:, s] = np.where(last >= 0, last // 3, -1)
    interior = (adj[:, 0] >= 0) & (adj[:, 1] >= 0)
    # hinge order = edge creation order = leader record order
    sel = np.nonzero(interior)[0]
    sel = sel[np.argsort(leader[sel])]
    L = leader[sel]
    e0, e1 = a[L], b[L]
    fa, fb = adj[sel, 0], adj[sel, 1]

    def apex(fc):
        t = tri[fc]
        m = (t != e0[:, None]) & (t != e1[:, None])
        return t[np.arange(len(fc)), np.argmax(m, axis=1)]

    return np.stack([e0, e1, apex(fa), apex(fb)], axis=1).astype(np.int32)


def _splitmix(seed, count):
    s = np.uint64(seed) + GOLDEN * np.arange(1, count + 1, dtype=np.uint64)
    z = s
    z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
    z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def uniform(seed, count):
    return (_splitmix(seed, count) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)


def gaussian(seed, count, sigma):
    count = int(count)
    m = (count + 1) // 2
    z = _splitmix(seed, 2 * m)
    u = ((z >> np.uint64(11)) + np.uint64(1)).astype(np.float64) * (1.0 / 9007199254740992.0)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    t = 6.283185307179586476925286766559 * u2
    out = np.empty(2 * m)
    out[0::2] = sigma * r * np.cos(t)
    out[1::2] = sigma * r * np.sin(t)
    return out[:count]


def fixed_corners(n):
    """The two y = 1 corners (all three DOFs fixed)."""
    return np.array([n * (n + 1), (n + 1) ** 2 - 1], dtype=np.int32)


class Scene:
    """A square cloth of n x n quads with the benchmark's initial state."""

    def __init__(self, n, sigma_x=0.0, sigma_v=0.0, seed_x=1234, seed_v=1235):
        self.n = int(n)
        self.X, self.tri = grid(self.n)
        self.N = self.X.shape[0]
        self.F = self.tri.shape[0]
        self.hinge = hinges(self.N, self.tri)
        self.H = self.hinge.shape[0]
        x = np.zeros((self.N, 3))
        x[:, :2] = self.X
        self.x0 = x.reshape(-1)
        if sigma_x:
            self.x0 = self.x0 + gaussian(seed_x, 3 * self.N, sigma_x)
        self.v0 = gaussian(seed_v, 3 * self.N, sigma_v) if sigma_v else np.zeros(3 * self.N)
        self.fixed = fixed_corners(self.n)


def obstacles(seed=1236, P=64, B=4):
    """Point obstacles uniform in [0,1]^2 x [-2e-3, 2e-3] with normals (0,0,+-1), and
    axis-aligned boxes with edges in [0.1, 0.3] centred in [0.15, 0.85]^2 at z = 0."""
    u = uniform(seed, 4 * P + 6 * B)
    up, ub = u[: 4 * P].reshape(P, 4), u[4 * P:].reshape(B, 6)
    pts = np.stack([up[:, 0], up[:, 1], -2e-3 + 4e-3 * up[:, 2]], axis=1)
    nrm = np.zeros((P, 3))
    nrm[:, 2] = np.where(up[:, 3] < 0.5, -1.0, 1.0)
    whd = 0.1 + 0.2 * ub[:, 0:3]
    E = []
    for k in range(B):
        M = np.eye(4)
        M[0, 3] = 0.15 + 0.7 * ub[k, 3]
        M[1, 3] = 0.15 + 0.7 * ub[k, 4]
        M[2, 3] = 0.0
        E.append(M)
    return pts, nrm, whd, E


class BatchScene:
    """Many cloths of uneven resolution and material: per instance the side is
    uniform in {side_lo..side_hi} quads, Young's modulus log-uniform in [E_lo, E_hi],
    bending stiffness log-uniform in [beta_lo, beta_hi] (splitmix64 stream `seed`),
    each with the single-cloth initial state (noise sigma_x, sigma_v; noise seeds
    seed * 1000 + 2b and seed * 1000 + 2b + 1) and its two top corners fixed.  With
    noise_per_spacing the position noise of each cloth is that fraction of its grid
    spacing instead of sigma_x."""

    def __init__(self, B=256, side_lo=16, side_hi=128, E_lo=1e2, E_hi=1e4, beta_lo=1e-6, beta_hi=1e-4,
                 sigma_x=1e-3, sigma_v=0.05, seed=1237, noise_per_spacing=None):
        u = uniform(seed, 3 * B).reshape(B, 3)
        self.B = int(B)
        self.sides = (side_lo + np.floor(u[:, 0] * (side_hi - side_lo + 1))).astype(int)
        self.sides = np.minimum(self.sides, side_hi)
        self.E = np.exp(np.log(E_lo) + u[:, 1] * (np.log(E_hi) - np.log(E_lo)))
        self.beta = np.exp(np.log(beta_lo) + u[:, 2] * (np.log(beta_hi) - np.log(beta_lo)))
        # noise_per_spacing: position noise sigma_x = noise_per_spacing / side per cloth,
        # the same noise relative to the grid spacing on every cloth
        self.sigma_x = [sigma_x if noise_per_spacing is None else noise_per_spacing / int(s) for s in self.sides]
        self.inst = [Scene(int(s), self.sigma_x[b], sigma_v, seed_x=seed * 1000 + 2 * b, seed_v=seed * 1000 + 2 * b + 1)
                     for b, s in enumerate(self.sides)]
        self.N_real = int(sum(sc.N for sc in self.inst))

    def setup(self, ctx, density=0.1, nu=0.3):
        """Topology, material and initial state on a context; returns vert_off."""
        off = ctx.set_batch_topology([sc.X for sc in self.inst], [sc.tri for sc in self.inst],
                                     [sc.hinge for sc in self.inst], [sc.fixed for sc in self.inst])
        ctx.set_material(density, 1.0, nu, 1e-5, 0.0, 1.0)
        ctx.set_batch_material(self.E, self.beta)
        self.off = off
        x, v = self.pack([sc.x0 for sc in self.inst]), self.pack([sc.v0 for sc in self.inst])
        ctx.upload_state(x, v)
        return off

    def pack(self, per_inst):
        """Concatenate per-instance 3N arrays into the padded numbering."""
        out = np.zeros(3 * int(self.off[-1]))
        for b, a in enumerate(per_inst):
            out[3 * self.off[b]: 3 * self.off[b] + a.size] = a
        return out

    def unpack(self, a):
        return [a[3 * self.off[b]: 3 * self.off[b] + 3 * sc.N] for b, sc in enumerate(self.inst)]
