"""Train the learned block preconditioner on systems the simulator generates.

    python -m eolc.train_mpn --n 256 --steps 6 --out weights.npy

For each of `--steps` time steps of a seeded scene (the benchmark's distribution:
initial position noise sigma_x, velocity noise sigma_v, the standard material),
the velocity system MDK v = M v_old + h f is solved with block-Jacobi PCG, and the
error vectors e = v* - v_k of the PCG iterates v_k at the snapshot iterations
(k = 0 is the warm start v_old) become training problems for the loss
1 - (z.r)^2 / ((z.A z)(r.e)), r = A e, z = P r.  Every forward, backward and Adam
step runs in the library's kernels (eolc_learned_train).  Several scenes (seeds)
are cycled so the network does not fit one trajectory.

evaluate() compares iteration counts of the mixed solve with block-Jacobi and with
the learned blocks on held-out systems.
"""
import argparse
import sys
import time

import numpy as np

import eolc
from eolc import learned, scene

GRAV = np.array([0.0, 0.0, -9.8])
H = 5e-3


def setup(ctx, sc):
    ctx.set_topology(sc.X, sc.tri, sc.hinge)
    ctx.set_material(0.1, 500.0, 0.3, 1e-5, 0.0, 1.0)
    ctx.set_options(eolc.OPT_SPD_PROJECT_BENDING)


def train(ctx, n, seeds, steps_per_scene, snapshots, adam_steps, lr, theta0=None, log=print,
          sigma_x=1e-3, sigma_v=0.05):
    """Train on `len(seeds)` scenes x `steps_per_scene` steps x snapshots; returns
    the weights and the loss history (first and last loss of every problem)."""
    theta = learned.init_weights(7) if theta0 is None else theta0
    ctx.learned_set_weights(theta)
    hist = []
    for sd in seeds:
        sc = scene.Scene(n, sigma_x=sigma_x, sigma_v=sigma_v, seed_x=sd, seed_v=sd + 1)
        setup(ctx, sc)
        t0 = time.perf_counter()
        for s, k, e, fd in learned.training_systems(ctx, sc, GRAV, H, steps_per_scene, snapshots=snapshots):
            losses = ctx.learned_train(e, adam_steps, lr, fd)
            hist.append((sd, s, k, float(losses[0]), float(losses[-1])))
            log(f"  scene {sd} step {s} snapshot {k}: loss {losses[0]:.5f} -> {losses[-1]:.5f} "
                f"({time.perf_counter() - t0:.1f} s)")
    return ctx.learned_get_weights(), hist


def batch_fixed_dofs(bs):
    """Fixed DOFs of a batched context: every cloth's corners and padding vertices."""
    out = []
    for b, sc in enumerate(bs.inst):
        lo, hi = int(bs.off[b]), int(bs.off[b + 1])
        nodes = np.concatenate([lo + np.asarray(sc.fixed), np.arange(lo + sc.N, hi)])
        out.append((3 * nodes[:, None] + np.arange(3)).reshape(-1))
    return np.concatenate(out).astype(np.int32)


def batch_training_systems(ctx, bs, n_steps, snapshots, ml=True, tol=1e-10, log=print):
    """Training problems from the batched path.  bs (a BatchScene) is set up on ctx.
    Each time step, one batched step (one persistent launch over every cloth) gives
    the solution v* of every cloth's velocity system; the iterates v_k of the mixed
    PCG (with the multilevel levels when `ml`) after k = `snapshots` iterations on the
    same concatenated system, from the warm start v_old (k = 0), give the error
    vectors e = v* - v_k.  Yields (step, k, e, fd) with the context holding that
    step's MDK; the state then advances to (x + h v*, v*)."""
    from . import ERR_MAXIT
    fd = batch_fixed_dofs(bs)
    mode = eolc.SOLVE_MIXED_BJ | (eolc.SOLVE_MULTILEVEL if ml else 0)
    for s in range(n_steps):
        x0, v0 = ctx.download_state()
        rc, st = ctx.batch_step(GRAV, H, tol=tol, maxit=200000, eager=True)
        x1, vstar = ctx.download_state()
        vals = v0[fd] * 0.99
        for k in snapshots:
            ctx.upload_state(x0, v0)
            ctx.forces_fill(GRAV, H)
            if k == 0:
                vk = v0
            else:
                ctx.velocity_solve(fd, vals, tol=1e-30, maxit=k, mode=mode, allow=(ERR_MAXIT,))
                vk = ctx.download_state()[1]
                ctx.upload_state(x0, v0)
                ctx.forces_fill(GRAV, H)
            e = vstar - vk
            e[fd] = 0.0
            yield s, k, e, fd
        ctx.upload_state(x1, vstar)


def train_batched(ctx, bs, n_steps, snapshots, adam_steps, lr, theta0=None, ml=True, epochs=1, log=print):
    """Train on problems from the batched path (batch_training_systems); returns the
    weights and the loss history (first and last loss of every problem)."""
    theta = learned.init_weights(7) if theta0 is None else theta0
    ctx.learned_set_weights(theta)
    ctx.learned_set_multilevel(ml)
    hist = []
    t0 = time.perf_counter()
    x_init, v_init = ctx.download_state()
    for ep in range(epochs):
        ctx.upload_state(x_init, v_init)
        for s, k, e, fd in batch_training_systems(ctx, bs, n_steps, snapshots, ml=ml, log=log):
            losses = ctx.learned_train(e, adam_steps, lr, fd)
            hist.append((ep, s, k, float(losses[0]), float(losses[-1])))
            log(f"  epoch {ep} step {s} snapshot {k}: loss {losses[0]:.5f} -> {losses[-1]:.5f} "
                f"({time.perf_counter() - t0:.1f} s)")
    return ctx.learned_get_weights(), hist


def evaluate(ctx, n, seeds, steps, theta, tol=1e-8, log=print, sigma_x=1e-3, sigma_v=0.05, flags=0):
    """Iterations and solve time of MIXED_BJ and MIXED_LEARNED on the same systems
    (each step advanced with the block-Jacobi solution); `flags` is OR-ed into both
    modes (eolc.SOLVE_MULTILEVEL: the fine blocks inside the multilevel preconditioner)."""
    ctx.learned_set_weights(theta)
    rows = []
    for sd in seeds:
        sc = scene.Scene(n, sigma_x=sigma_x, sigma_v=sigma_v, seed_x=sd, seed_v=sd + 1)
        setup(ctx, sc)
        ctx.upload_state(sc.x0, sc.v0)
        fd = learned.fixed_dofs(sc.fixed)
        for s in range(steps):
            x0, v0 = ctx.download_state()
            vals = 0.99 * v0[fd]
            out = {}
            for name, mode in (("bj", eolc.SOLVE_MIXED_BJ | flags), ("learned", eolc.SOLVE_MIXED_LEARNED | flags)):
                ctx.upload_state(x0, v0)
                ctx.forces_fill(GRAV, H)
                t = time.perf_counter()
                rc, st = ctx.velocity_solve(fd, vals, tol=tol, mode=mode)
                out[name] = (st.iters, time.perf_counter() - t, st.true_relres, st.used_fallback)
                if name == "bj":
                    vstar = ctx.download_state()[1]
            rows.append((sd, s, out["bj"][0], out["learned"][0], out["bj"][1], out["learned"][1],
                         out["learned"][2], out["learned"][3]))
            log(f"  eval scene {sd} step {s}: iters bj {out['bj'][0]} learned {out['learned'][0]} "
                f"(relres {out['learned'][2]:.2e}, fallback {out['learned'][3]})")
            ctx.upload_state(x0 + H * vstar, vstar)
    return rows


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--batched", action="store_true",
                    help="training problems from the batched path (many cloths, one persistent launch)")
    ap.add_argument("--cloths", type=int, default=32)
    ap.add_argument("--side-lo", type=int, default=128)
    ap.add_argument("--side-hi", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=1)
    ap.add_argument("--no-ml", action="store_true", help="train the blocks without the multilevel levels")
    ap.add_argument("--eval-n", type=int, default=None)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--scenes", type=int, default=3)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--snapshots", default="0,20,80")
    ap.add_argument("--adam-steps", type=int, default=30)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--eval-scenes", type=int, default=2)
    ap.add_argument("--eval-steps", type=int, default=2)
    ap.add_argument("--out", default=None)
    a = ap.parse_args(argv)
    snaps = tuple(int(s) for s in a.snapshots.split(","))
    log = lambda m: print(m, flush=True)
    with eolc.Context(0) as ctx:
        t = time.perf_counter()
        if a.batched:
            # config-3 systems: the standard material, position noise of one 1024^2 grid
            # spacing relative to each cloth's own spacing, velocity noise 0.05
            bs = scene.BatchScene(B=a.cloths, side_lo=a.side_lo, side_hi=a.side_hi, E_lo=500.0, E_hi=500.0,
                                  beta_lo=1e-5, beta_hi=1e-5, sigma_v=0.05, noise_per_spacing=1e-3 * 1024,
                                  seed=2718)
            bs.setup(ctx)
            ctx.set_options(eolc.OPT_SPD_PROJECT_BENDING)
            ctx.set_batch_multilevel(True)
            log(f"batched training set: {bs.B} cloths, sides {int(bs.sides.min())}..{int(bs.sides.max())}, "
                f"{bs.N_real} vertices")
            theta, hist = train_batched(ctx, bs, a.steps, snaps, a.adam_steps, a.lr, ml=not a.no_ml,
                                        epochs=a.epochs, log=log)
        else:
            theta, hist = train(ctx, a.n, [1000 + 17 * i for i in range(a.scenes)], a.steps, snaps,
                                a.adam_steps, a.lr, log=log)
        log(f"training done in {time.perf_counter() - t:.1f} s")
        if a.out:
            np.save(a.out, theta)
            log(f"weights -> {a.out}")
        ne = a.eval_n or a.n
        for flags, what in ((0, "block-Jacobi"), (eolc.SOLVE_MULTILEVEL, "multilevel")):
            rows = evaluate(ctx, ne, [5000 + 17 * i for i in range(a.eval_scenes)], a.eval_steps, theta, log=log,
                            flags=flags)
            bj = sum(r[2] for r in rows)
            le = sum(r[3] for r in rows)
            log(f"held-out, {what}: without the network {bj} iterations, with it {le} "
                f"({bj / max(le, 1):.3f}x fewer)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
