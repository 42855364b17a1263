"""FP64 CPU oracle of the learned block preconditioner (message-passing network).

TEST INFRASTRUCTURE ONLY: the checker for the GPU forward/backward kernels.

The network (its definition is this project's own; the reference has no learned
component) maps per-vertex features of the assembled system A = M + h^2 K to a
per-vertex lower-triangular correction T_i of the block-Jacobi factor:

    x_i   = features(A)_i                       (16 numbers, see features())
    h0    = relu(x W_in + b_in)                 (64)
    a_l   = sum_j w_ij h_l[j]   (row-normalised off-diagonal block norms w_ij)
    h_l+1 = relu([h_l | a_l] W_l + b_l)         l = 0, 1, 2  (128 -> 64)
    o     = h3 W_out + b_out                    (6)
    T_i   = lower-triangular 3x3 from 0.5 tanh(o_i)
    P_i   = G_i G_i^T,  G_i = L_i (I + T_i),  L_i L_i^T = D_i^-1 (Cholesky)

P is SPD for every output (I + T_i has a positive diagonal) and equals block-Jacobi
when o = 0.  Training loss on a system with residual r and error e (A e = r):
z = P r and

    loss = 1 - (z.r)^2 / ((z.A z) (r.e)),

the A-norm error left after one optimally scaled step along z, relative to the
initial A-norm error.  Pinned by central finite differences in FP64
(tests/test_learned_oracle.py) and used to check the GPU gradients.

`quant=True` rounds the tensors the GPU stores in bfloat16 (weights used by the
matrix units, activations, aggregates, features and the backward pre-activation
gradients) the way the GPU does (round to nearest even), so the comparison with
the GPU isolates FP32-vs-FP64 accumulation.
"""
import numpy as np

NFEAT, HID, NOUT, NLAYER = 16, 64, 6, 3
TRI = [(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]   # output k -> T[r][c]


def bf16(a):
    a = np.asarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    return r.view(np.float32).astype(np.float64)


def param_shapes():
    s = [("W_in", (NFEAT, HID)), ("b_in", (HID,))]
    for l in range(NLAYER):
        s += [(f"W{l}", (2 * HID, HID)), (f"b{l}", (HID,))]
    s += [("W_out", (HID, NOUT)), ("b_out", (NOUT,))]
    return s


def unpack(theta):
    out, k = {}, 0
    for name, shp in param_shapes():
        n = int(np.prod(shp))
        out[name] = np.asarray(theta[k:k + n], dtype=np.float64).reshape(shp)
        k += n
    return out


def nparams():
    return sum(int(np.prod(s)) for _, s in param_shapes())


def init_params(seed):
    """Deterministic initialisation (splitmix64 stream), output layer zero; the
    same values the library's host helper produces (as FP64)."""
    from eolc import learned
    return learned.init_weights(seed).astype(np.float64)


def blocks(A, N):
    """Dense 3x3 blocks of a scalar CSR (scipy) matrix: dict of (i, j) -> 3x3."""
    A = A.tocoo()
    out = {}
    for r, c, v in zip(A.row, A.col, A.data):
        key = (int(r) // 3, int(c) // 3)
        if key not in out:
            out[key] = np.zeros((3, 3))
        out[key][r % 3, c % 3] += v
    return out


def features(A, mass, N, fixmask):
    """Per-vertex features and row-normalised edge weights.

    features (16): 6 upper entries of D_i / (tr D_i / 3); log(tr D_i / 3) - mean;
    m_ii / (tr D_i / 3); sum_j |A_ij|_F / |D_i|_F; (#blocks - 1) / 16; the 3 fixed
    flags; 3 zeros.  Edge weight w_ij = |A_ij|_F / sum_j |A_ij|_F (j != i).
    """
    B = blocks(A, N)
    X = np.zeros((N, NFEAT))
    nbr = [[] for _ in range(N)]
    logt = np.zeros(N)
    for (i, j), blk in B.items():
        if i != j:
            nbr[i].append((j, np.linalg.norm(blk)))
    for i in range(N):
        D = B[(i, i)]
        t = np.trace(D) / 3.0
        Dn = D / t
        X[i, 0:6] = [Dn[0, 0], Dn[0, 1], Dn[0, 2], Dn[1, 1], Dn[1, 2], Dn[2, 2]]
        logt[i] = np.log(t)
        X[i, 7] = mass[i] / t
        s = sum(w for _, w in nbr[i])
        X[i, 8] = s / np.linalg.norm(D)
        X[i, 9] = len(nbr[i]) / 16.0
        for k in range(3):
            X[i, 10 + k] = float((fixmask[i] >> k) & 1)
    X[:, 6] = logt - logt.mean()
    W = {}
    for i in range(N):
        s = sum(w for _, w in nbr[i])
        for j, w in nbr[i]:
            W[(i, j)] = w / s if s > 0 else 0.0
    return X, W, B


def forward(theta, X, W, N, quant=False):
    q = bf16 if quant else (lambda a: np.asarray(a, dtype=np.float64))
    p = unpack(theta)
    Wm = {k: q(v) if k.startswith("W") else v for k, v in p.items()}
    Xq = q(X)
    cache = {"X": Xq}
    h = q(np.maximum(Xq @ Wm["W_in"] + p["b_in"], 0.0))
    cache["h0"] = h
    rows = np.array([i for (i, j) in W]), np.array([j for (i, j) in W]), np.array([W[k] for k in W])
    for l in range(NLAYER):
        agg = np.zeros((N, HID))
        np.add.at(agg, rows[0], rows[2][:, None] * h[rows[1]])
        agg = q(agg)
        cache[f"a{l}"] = agg
        h = q(np.maximum(np.hstack([h, agg]) @ Wm[f"W{l}"] + p[f"b{l}"], 0.0))
        cache[f"h{l + 1}"] = h
    o = h @ Wm["W_out"] + p["b_out"]
    cache["o"] = o
    return o, cache, rows


def precond_from_outputs(o, B, N, fixmask):
    """P_i (3x3) from the network outputs."""
    P = np.zeros((N, 3, 3))
    aux = []
    for i in range(N):
        D = B[(i, i)].copy()
        for k in range(3):
            if (fixmask[i] >> k) & 1:
                D[k, :] = 0; D[:, k] = 0; D[k, k] = 1.0
        Dinv = np.linalg.inv(D)
        L = np.linalg.cholesky(Dinv)
        T = np.zeros((3, 3))
        th = np.tanh(o[i])
        for k, (r, c) in enumerate(TRI):
            T[r, c] = 0.5 * th[k]
        G = L @ (np.eye(3) + T)
        Pi = G @ G.T
        for k in range(3):
            if (fixmask[i] >> k) & 1:
                Pi[k, :] = 0; Pi[:, k] = 0
        P[i] = Pi
        aux.append((L, G, th))
    return P, aux


def loss_from_outputs(o, A, B, N, fixmask, r, e):
    """Loss for given network outputs o (N x 6); also z, A z and the factors."""
    P, aux = precond_from_outputs(o, B, N, fixmask)
    z = np.einsum("nij,nj->ni", P, r.reshape(N, 3)).reshape(-1)
    Az = A @ z
    zr, zAz, re = z @ r, z @ Az, r @ e
    return 1.0 - zr * zr / (zAz * re), (P, aux, z, Az, zr, zAz, re)


def loss_and_grad(theta, A, mass, N, fixmask, r, e, quant=False):
    """Loss and its gradient with respect to all parameters (FP64 backprop)."""
    X, W, B = features(A, mass, N, fixmask)
    o, cache, rows = forward(theta, X, W, N, quant)
    loss, (P, aux, z, Az, zr, zAz, re) = loss_from_outputs(o, A, B, N, fixmask, r, e)
    r3 = r.reshape(N, 3)
    dz = -2.0 * zr / (zAz * re) * (r - (zr / zAz) * Az)
    dz3 = dz.reshape(N, 3)
    dO = np.zeros((N, NOUT))
    for i in range(N):
        S = np.outer(dz3[i], r3[i])
        for k in range(3):
            if (fixmask[i] >> k) & 1:
                S[k, :] = 0; S[:, k] = 0
        L, G, th = aux[i]
        dG = (S + S.T) @ G
        dT = L.T @ dG
        for k, (rr, cc) in enumerate(TRI):
            dO[i, k] = dT[rr, cc] * 0.5 * (1.0 - th[k] ** 2)
    p = unpack(theta)
    q = bf16 if quant else (lambda a: np.asarray(a, dtype=np.float64))
    Wm = {k: q(v) if k.startswith("W") else v for k, v in p.items()}
    g = {k: np.zeros_like(v) for k, v in p.items()}
    if quant:
        dO = dO.astype(np.float32).astype(np.float64)
    h = cache[f"h{NLAYER}"]
    g["W_out"] = h.T @ dO
    g["b_out"] = dO.sum(0)
    dh = dO @ Wm["W_out"].T
    ri, rj, rw = rows
    for l in reversed(range(NLAYER)):
        hl1 = cache[f"h{l + 1}"]
        dpre = q(dh * (hl1 > 0))
        inp = np.hstack([cache[f"h{l}"], cache[f"a{l}"]])
        g[f"W{l}"] = inp.T @ dpre
        g[f"b{l}"] = dpre.sum(0)
        dinp = dpre @ Wm[f"W{l}"].T
        dh = dinp[:, :HID].copy()
        dagg = dinp[:, HID:]
        np.add.at(dh, rj, rw[:, None] * dagg[ri])
    dpre = q(dh * (cache["h0"] > 0))
    g["W_in"] = cache["X"].T @ dpre
    g["b_in"] = dpre.sum(0)
    grad = np.concatenate([g[name].reshape(-1) for name, _ in param_shapes()])
    return loss, grad, {"P": P, "o": o, "X": X}
