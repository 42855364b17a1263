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
        theta, hist = train(ctx, a.n, [1000 + 17 * i for i in range(a.scenes)], a.steps, snaps,
                            a.adam_steps, a.lr, log=log)
        log(f"training done in {time.perf_counter() - t:.1f} s")
        if a.out:
            np.save(a.out, theta)
            log(f"weights -> {a.out}")
        for flags, what in ((0, "block-Jacobi"), (eolc.SOLVE_MULTILEVEL, "multilevel")):
            rows = evaluate(ctx, a.n, [5000 + 17 * i for i in range(a.eval_scenes)], a.eval_steps, theta, log=log,
                            flags=flags)
            bj = sum(r[2] for r in rows)
            le = sum(r[3] for r in rows)
            log(f"held-out, {what}: without the network {bj} iterations, with it {le} "
                f"({bj / max(le, 1):.3f}x fewer)")
    return 0


if __name__ == "__main__":
    sys.exit(main())
