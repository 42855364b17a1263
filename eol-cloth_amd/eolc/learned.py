"""Host-side helpers of the learned block preconditioner.

The network, its training kernels and the preconditioner blocks live in the
library (k_learned.hip, entry points eolc_learned_*); this module holds the
parameter layout, the deterministic initialisation and the training-data loop
that drives the simulator to produce (MDK, r, e) systems.
"""
import numpy as np

from . import scene

NFEAT, HID, NOUT, NLAYER = 16, 64, 6, 3


def param_shapes():
    """Parameter tensors in the order of the flat vector the library takes."""
    s = [("W_in", (NFEAT, HID)), ("b_in", (HID,))]
    for l in range(NLAYER):
        s += [(f"W{l}", (2 * HID, HID)), (f"b{l}", (HID,))]
    s += [("W_out", (HID, NOUT)), ("b_out", (NOUT,))]
    return s


def nparams():
    return sum(int(np.prod(s)) for _, s in param_shapes())


def init_weights(seed=7):
    """He-scaled Gaussian weights (splitmix64 stream); biases and W_out zero, so
    the initial preconditioner is exactly block-Jacobi."""
    th = []
    for name, shp in param_shapes():
        n = int(np.prod(shp))
        if name.startswith("b") or name == "W_out":
            th.append(np.zeros(n))
        else:
            th.append(scene.gaussian(seed + len(th), n, np.sqrt(2.0 / shp[0])))
    return np.concatenate(th).astype(np.float32)


def fixed_dofs(fixed_nodes):
    return np.array([3 * int(i) + k for i in fixed_nodes for k in range(3)], dtype=np.int32)


def training_systems(ctx, sc, grav, h, n_steps, tol=1e-10, mode=0, snapshots=(0,)):
    """Yield (step, e) for the first n_steps steps of the scene sc on ctx.

    Each step solves the velocity system with block-Jacobi PCG to tol; the error
    vectors are e = v* - v_k, where v_k is the iterate after k PCG iterations for
    k in `snapshots` (k = 0 is the warm start v_old).  The context keeps the
    system of the step being yielded (its MDK and fixed set), so a caller can
    train on it before advancing.  The state ends at the step's solution.
    """
    from . import ERR_MAXIT
    ctx.upload_state(sc.x0, sc.v0)
    fd = fixed_dofs(sc.fixed)
    for s in range(n_steps):
        x0, v0 = ctx.download_state()
        ctx.forces_fill(grav, h)
        # fixed values as in eolc_step
        vals = np.repeat(0.99, fd.size) * v0[fd]
        sols = {}
        for k in sorted(set(snapshots) | {-1}):
            ctx.upload_state(x0, v0)
            ctx.forces_fill(grav, h)
            if k == 0:
                sols[0] = v0.copy()
                continue
            if k < 0:
                ctx.velocity_solve(fd, vals, tol=tol, mode=mode)
            else:
                ctx.velocity_solve(fd, vals, tol=1e-30, maxit=k, mode=mode, allow=(ERR_MAXIT,))
            sols[k] = ctx.download_state()[1]
        vstar = sols.pop(-1)
        ctx.upload_state(x0, v0)
        ctx.forces_fill(grav, h)
        for k, vk in sorted(sols.items()):
            e = vstar - vk
            e[fd] = 0.0
            yield s, k, e, fd
        # advance: x += h v*, v = v*
        ctx.upload_state(x0 + h * vstar, vstar)
