#!/usr/bin/env python3
"""Benchmark: steps/s of the per-step cloth hot path on one MI355X.

Workload: the 1024x1024-quad cloth (1,050,625 vertices, 2,097,152 triangles,
3,143,680 hinges), x0 = grid + N(0, 1e-3^2), v0 ~ N(0, 0.05^2) (seeds 1234 and
1235), two top corners fixed, rho 0.1, E 500, nu 0.3, beta 1e-5, h 5e-3,
gravity -9.8.  One step = element evaluation + atomic-free assembly of M, MDK, f +
rhs + velocity solve to 1e-8 relative residual (FP64, reduced system) + x += h v,
all on the device with the state resident in HBM.  Each step's solve is checked
against the tolerance.

The headline path is the learned-preconditioned mixed-precision PCG (FP32 matrix
and vectors, FP64 dots and true residual with reliable updates, fine blocks P_i from
the message-passing network with weights eol-cloth_amd/eolc/weights/mpn_cloth.npy,
plus the coarse levels of the additive multilevel preconditioner
(EOLC_SOLVE_MULTILEVEL), block-Jacobi fallback).  Every other path runs from the
same initial state as its own leg, so the gain of each ingredient is stated: FP64
3x3 block-Jacobi PCG (the baseline of the metric), mixed block-Jacobi, the learned
blocks without the coarse levels, and both block-Jacobi precisions with them.  A
leg that fails is recorded with its error and the others still run; optional legs
are skipped once --budget-s seconds of measurement have passed.  Before timing, the
SpMV launch variant of each precision is chosen by the library's autotuner on the
first matrix.  The noisy initial state makes the exact bending Hessian indefinite,
so every path assembles the SPD-projected hinge Hessians
(EOLC_OPT_SPD_PROJECT_BENDING).

Prints one JSON line (rank 0) with `roofline` (the dominant kernel, the FP32 SpMV:
algorithmic bytes over its average launch time, HIP events on the library's
stream), `roofline_step` (B_step = B_asm + k B_it + r B_res + 2 v(8) over the time
of a headline step) and `cpu_baseline` (the oracle on this host's cores, a bounded
sample of the same workload).  The roofline's `traffic` (HBM bytes per SpMV
launch) is read from profiles/pmc_spmv_cloth1024.json, the calibrated rocprofv3
FETCH_SIZE/WRITE_SIZE summary of `bench.py --profile-spmv`, when it was measured
for the same kernel variant and size.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "eol-cloth_amd"))

import numpy as np  # noqa: E402

import eolc  # noqa: E402
from eolc import scene  # noqa: E402

GRAV = np.array([0.0, 0.0, -9.8])
H = 5e-3
TOL = 1e-8
PEAK_HBM = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmv_bytes(N, nb, s, diff):
    """Algorithmic bytes of one k_pcg_spmv launch: 9 scalars and a column index per
    block read (the FP32 difference form skips the N diagonal blocks and reads one
    row-mass scalar per row instead), per-row length and row index, z and p_old
    read, p_new and q written, fixed mask."""
    blocks = nb - N if diff else nb
    return (9 * s + 4) * blocks + (4 * N if diff else 0) + 8 * N + 4 * s * 3 * N + N


class MissedTolerance(RuntimeError):
    pass


def step_line(name, k, st):
    return (f"{name} step {k}: {st.iters} iterations, {st.refinements} replacements, solve {st.ms_solve:.1f} ms, "
            f"true residual {st.true_relres:.2e}, status {st.status}")


def run_path(ctx, sc, mode, steps, warmup, maxit, name="", window=12):
    """W untimed steps, then K timed steps of one solve mode along the benchmark
    trajectory: steps 0, 1, ..., window - 1 from the scene's initial state, then the
    state is reset to the initial one and the sequence repeats.  (The sheet degenerates
    as it relaxes from its noise: by step 17 triangles of 1e-11 m^2 make the system
    unsolvable to 1e-8 for every solver here, FP64 block-Jacobi included, so the
    measured window stops short of that.)  Every step must end EOLC_OK with its FP64
    true residual <= 1e-8.  Each step's line goes to stderr after the step; the state
    resets (one host-to-device copy) are inside the timed region."""
    s = 0

    def one(label, k):
        nonlocal s
        if s % window == 0:
            ctx.upload_state(sc.x0, sc.v0)
        rc, st = ctx.step(GRAV, H, sc.fixed, tol=TOL, maxit=maxit, mode=mode, allow=(eolc.ERR_MAXIT,))
        log(step_line(label, k, st) + f" (trajectory step {s % window})")
        s += 1
        return rc, st

    for w in range(warmup):
        rc, st = one(name + " warm-up", w)
        if rc != eolc.OK:
            raise MissedTolerance(f"warm-up solve failed: {st.as_dict()}")
    t0 = time.perf_counter()
    stats = []
    for k in range(steps):
        rc, st = one(name, k)
        stats.append(st.as_dict())
        if rc != eolc.OK:   # a failed step leaves the state as it was: stop here
            break
    t1 = time.perf_counter()
    bad = [s for s in stats if not (s["status"] == eolc.OK and s["true_relres"] <= TOL)]
    if bad:
        raise MissedTolerance(f"solve missed the tolerance: {bad[0]}")
    return (t1 - t0) / len(stats), stats


def leg(results, name, fn):
    """Runs one measurement; a failure is recorded under its name and never hides
    the other measurements."""
    t = time.perf_counter()
    try:
        out = fn()
        if out is None:
            out = {}
        out["status"] = "ok"
    except Exception as e:  # noqa: BLE001 - recorded in the JSON line
        out = {"status": "failed", "error": f"{type(e).__name__}: {e}"}
        log(f"{name}: FAILED {out['error']}")
    out["wall_s"] = round(time.perf_counter() - t, 2)
    results[name] = out
    return out


def bsr_bytes(N, nb, s):
    """A(s) = (9s + 4) nb + 4 (N + 1): the 3x3 block matrix with its column indices
    and row pointers."""
    return (9 * s + 4) * nb + 4 * (N + 1)


def step_roofline(sc, nb, stats, ms_step, mixed):
    """Step-level roofline: B_step = B_asm + k B_it + r B_res + 2 v(8) over the time
    of a step, against the 8 TB/s HBM peak.  B_it = A(s) + D(s) + 9 v(s) (s = 4 in
    the mixed modes), B_res = A(8) + 3 v(8), B_asm = A(8) [+ A(4) in the mixed
    modes, which write both matrices] + 8 n + 16 N + 12 F + 16 H; k = iterations and
    r = true-residual replacements per step (means over the timed steps)."""
    N, n = sc.N, 3 * sc.N
    s = 4 if mixed else 8
    A8, As = bsr_bytes(N, nb, 8), bsr_bytes(N, nb, s)
    b_it = As + 9 * s * N + 9 * s * n
    b_res = A8 + 3 * 8 * n
    b_asm = A8 + (As if mixed else 0) + 8 * n + 16 * N + 12 * sc.F + 16 * sc.H
    k = float(np.mean([st["iters"] for st in stats]))
    r = float(np.mean([st["refinements"] for st in stats]))
    b_step = b_asm + k * b_it + r * b_res + 2 * 8 * n
    ach = b_step / (ms_step * 1e-3)
    return {"bound": "hbm", "achieved": ach / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
            "frac": ach / PEAK_HBM, "bytes_per_step": b_step, "B_it": b_it, "B_asm": b_asm, "B_res": b_res,
            "iters_per_step": k, "replacements_per_step": r}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            return next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), None)
    except OSError:
        return None


def cpu_baseline(sc, iters_per_step, threads):
    """The oracle (FP64 CPU restatement of Forces::fill + the block-Jacobi PCG the dead
    ConjugateGradient branch intended) on the host cores, on a bounded sample of the
    same 1024^2 workload: one full Forces::fill (elements, triplets, sort-and-sum CSR)
    measured whole, and a few PCG iterations on its matrix, at all cores and at one
    thread; a step is the fill plus the FP64 iteration count the GPU needed per step.
    A whole 1024^2 step (about 25,000 iterations) takes the oracle tens of minutes,
    which is why the solve is sampled."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    mat = O.material()
    d = (3 * sc.fixed[:, None] + np.arange(3)).reshape(-1).astype(np.int32)
    out = {}
    for th, k in ((threads, 10), (1, 3)):
        t0 = time.perf_counter()
        rc, f, M, K = O.forces_fill(sc.x0, sc.X, sc.tri, sc.hinge, mat, GRAV, H, nthreads=th)
        fill = time.perf_counter() - t0
        rhs = M @ sc.v0 + H * f
        t0 = time.perf_counter()
        O.velocity_solve(K, rhs, d, 0.99 * sc.v0[d], v0=sc.v0, tol=1e-30, maxit=k,
                         precond=O.PC_BLOCK_JACOBI, nthreads=th)
        t_iter = (time.perf_counter() - t0) / k
        t_step = fill + iters_per_step * t_iter
        out[th] = {"value": 1.0 / t_step, "unit": "steps/s", "cores": th, "fill_s": fill,
                   "iter_ms": t_iter * 1e3,
                   "sample": f"oracle Forces::fill at {sc.n}x{sc.n} measured whole ({fill:.2f} s) + {k} FP64 "
                             f"block-Jacobi PCG iterations ({t_iter * 1e3:.1f} ms each), a step = the fill + "
                             f"{iters_per_step:.0f} iterations (the GPU's FP64 count per step)"}
        del f, M, K, rhs
    # the plain method in full (20x20 quads, 200 steps, FP64, tolerance 1e-8): whole
    # steps measured on one thread
    s1 = scene.Scene(20)
    x, v = s1.x0.copy(), s1.v0.copy()
    t0 = time.perf_counter()
    its = 0
    for _ in range(200):
        rc, st, _ = O.step(x, v, s1.X, s1.tri, s1.hinge, O.material(), GRAV, H, s1.fixed, tol=TOL, maxit=20000)
        if rc != 0:
            raise RuntimeError(f"oracle step failed ({rc})")
        its += st.iters
    t1 = time.perf_counter() - t0
    res = dict(out[threads])
    res.update({"kind": "port", "cpu_model": cpu_model(), "nproc": os.cpu_count(),
                "one_thread": out[1],
                "config1_full": {"value": 200 / t1, "unit": "steps/s", "cores": 1, "steps": 200,
                                 "iters_per_step": its / 200,
                                 "sample": "20x20 quads, 200 whole oracle steps (fill + Jacobi PCG to 1e-8)"}})
    return res


def run_config1_gpu(steps=200):
    """The plain method (20x20 quads, 200 steps, FP64 block-Jacobi to 1e-8) on the
    GPU: steps/s of whole steps, and the positions after 200 steps against the
    oracle's (the sheet buckles, so the difference reflects the solver tolerance)."""
    s1 = scene.Scene(20)
    with eolc.Context(0) as c:
        c.set_topology(s1.X, s1.tri, s1.hinge)
        c.set_material(0.1, 500.0, 0.3, 1e-5, 0.0, 1.0)
        c.upload_state(s1.x0, s1.v0)
        t0 = time.perf_counter()
        its = 0
        for _ in range(steps):
            rc, st = c.step(GRAV, H, s1.fixed, tol=TOL, maxit=20000, mode=eolc.SOLVE_FP64_BJ)
            its += st.iters
        dt = (time.perf_counter() - t0) / steps
    return {"workload": "cloth 20x20 quads, 200 steps, FP64 block-Jacobi PCG to 1e-8",
            "steps_per_s": 1.0 / dt, "iters_per_step": its / steps}


def run_collision(reps, with_cpu):
    """Proximity queries (CD2) of the 512x512 cloth with noise 1e-3 against 64 seeded
    points and 4 seeded boxes at threshold 1e-3: device-resident cloth, records
    returned to the host each call.  The oracle's brute-force restatement of the
    same queries is timed on the host (one thread) as the CPU baseline, and its
    record count checked against the GPU's."""
    sc = scene.Scene(512, sigma_x=1e-3, sigma_v=0.05)
    pts, nrm, whd, E = scene.obstacles(seed=1236, P=64, B=4)
    ctx = eolc.Context(0)
    ctx.set_topology(sc.X, sc.tri, sc.hinge)
    ctx.upload_state(sc.x0, np.zeros_like(sc.x0))
    recs = ctx.collide(1e-3, pts, nrm, whd, E)
    t0 = time.perf_counter()
    for _ in range(reps):
        recs = ctx.collide(1e-3, pts, nrm, whd, E)
    t_gpu = (time.perf_counter() - t0) / reps
    ctx.close()
    # the QP's collision rows from the records (Constraints::fill, host code): the
    # library call on a contiguous record array (size query, then the rows)
    arr = eolc.records_array(recs)
    eolc.collision_constraints(arr, sc.N)   # warm-up (imports scipy.sparse)
    t0 = time.perf_counter()
    Aineq = eolc.collision_constraints(arr, sc.N)
    t_rows = time.perf_counter() - t0
    out = {"workload": "cloth 512x512 quads, noise 1e-3 m, 64 points + 4 boxes, threshold 1e-3 (CD2)",
           "records": len(recs), "ms_per_call": t_gpu * 1e3, "calls_per_s": 1.0 / t_gpu,
           "faces_tested_per_s": (64 + 4 * 8) * sc.F / t_gpu,
           "constraint_rows": {"rows": int(Aineq.shape[0]), "nnz": int(Aineq.nnz), "ms": t_rows * 1e3}}
    if with_cpu:
        sys.path.insert(0, os.path.join(ROOT, "oracle"))
        import oracle as O
        t1 = time.perf_counter()
        n, ro = O.point_tri(1e-3, pts, nrm, sc.x0, sc.tri)
        nrec = n
        for b in range(len(whd)):
            nb, rb = O.box_tri(1e-3, whd[b], E[b], sc.x0, sc.tri)
            nrec += nb
        t_cpu = time.perf_counter() - t1
        out["cpu_baseline"] = {"ms_per_call": t_cpu * 1e3, "cores": 1, "kind": "port", "records": nrec}
        if nrec != len(recs):
            raise RuntimeError(f"collision record count differs: gpu {len(recs)} oracle {nrec}")
    log(f"collision: {out}")
    return out


WEIGHTS = os.path.join(ROOT, "eol-cloth_amd", "eolc", "weights", "mpn_cloth.npy")
TRAINED = os.path.join(ROOT, "eol-cloth_amd", "eolc", "weights", "mpn_cloth_trained.npy")
TRAFFIC = os.path.join(ROOT, "profiles", "pmc_spmv_cloth1024.json")


def load_weights(path=WEIGHTS):
    """Network weights from `path` (a float32 .npy, loaded without pickle), else the
    initial weights, whose zero output layer gives exactly the block-Jacobi blocks."""
    from eolc import learned
    if os.path.exists(path):
        th = np.load(path, allow_pickle=False).astype(np.float32)
        if th.size == eolc.LEARNED_NPARAMS:
            return th, os.path.relpath(path, ROOT)
    return learned.init_weights(7), "initial weights (zero output: the block-Jacobi blocks)"


# SpMV launch variants: (entries per trip, column prefetch, non-temporal loads,
# waves per SIMD, block layout, LDS-staged gathers), in the library's order
SPMV_VARIANTS = [(1, False, False, 8, 0, False), (1, True, False, 8, 0, False), (1, True, True, 8, 0, False),
                 (1, False, True, 8, 0, False), (1, True, True, 6, 0, False), (2, True, True, 6, 0, False),
                 (2, True, False, 6, 0, False), (2, True, True, 4, 0, False), (1, True, True, 4, 0, False),
                 (1, False, False, 8, 1, False), (1, False, False, 6, 1, False), (1, False, True, 8, 1, False),
                 (1, False, False, 8, 2, False), (1, False, True, 8, 2, False), (1, False, False, 6, 2, False),
                 (1, False, False, 8, 3, False), (1, False, True, 8, 3, False), (1, False, True, 6, 3, False),
                 (1, False, False, 8, 4, False), (1, False, True, 8, 0, True), (1, False, True, 4, 0, True),
                 (1, False, False, 8, 1, True), (1, False, False, 4, 4, True), (1, False, True, 8, 3, True),
                 (1, False, True, 4, 3, True), (1, False, False, 8, 5, False), (2, False, False, 8, 5, False),
                 (1, False, False, 4, 5, False), (2, False, False, 4, 5, False)]


def spmv_kernel_name(fp32, variant):
    u, pf, nt, w, lay, stg = SPMV_VARIANTS[variant]
    b = lambda x: "true" if x else "false"
    if lay == 5:
        t = "float" if fp32 else "double"
        return f"k_pcg_spmv_patch<{t}, {t}, {b(nt)}, {w}, {u}>"
    t = "float, true" if fp32 else ("double, true" if lay == 4 else "double, false")
    return f"k_pcg_spmv<{t}, {u}, {b(pf)}, {b(nt)}, {w}, {lay if (fp32 or lay >= 3) else 0}, {b(stg)}>"


def measured_traffic(kernel, n):
    """HBM bytes per launch of `kernel` at n x n from the committed counter summary."""
    if n != 1024:
        return None
    try:
        with open(TRAFFIC) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    # summaries made before the staged variants name the kernels without the last
    # template argument (false)
    short = kernel[: -len(", false>")] + ">" if kernel.endswith(", false>") else None
    for k in t.get("kernels", []):
        if k.get("kernel_template") in (kernel, short):
            return k.get("hbm_bytes")
    return None


def run_transfer(sc, reps=3):
    """Cloth::velocityTransfer after a remesh-like change at the benchmark's size: the
    old 1024^2 mesh, every second old vertex kept (moved by 1e-9) and 1 M new points
    at seeded random positions.  Device time per call (binning and search; the
    uploads excluded), the old O(N_new N_old) scans being what it replaces."""
    rng = np.random.default_rng(5)
    P = np.concatenate([sc.X[::2] + 1e-9, rng.uniform(0.0, 1.0, size=(1_000_000, 2))])
    with eolc.Context(0) as c:
        ms = []
        for _ in range(reps + 1):
            rc, vg, sg, fg, bg, t = c.velocity_transfer(sc.X, sc.v0, sc.tri, P)
            ms.append(t)
    out = {"workload": f"old mesh {sc.N} vertices, {len(P)} new points", "ms_device": float(np.median(ms[1:])),
           "points_per_s": len(P) / (np.median(ms[1:]) * 1e-3), "matched": int(np.sum(sg >= 0)),
           "in_faces": int(np.sum(fg >= 0))}
    log(f"velocity transfer: {out}")
    return out


def run_eol_fill(ctx, sc, n, reps=5):
    """Forces::fill with the sheet's boundary as EOL nodes (the EOL rows of
    Forces.cpp:177-329, 399-497, 580-683, 746-884) against the plain fill: wall time
    per call (each call ends with a stream synchronisation)."""
    ij = np.stack([np.arange(sc.N) % (n + 1), np.arange(sc.N) // (n + 1)], 1)
    eol = np.where((ij[:, 0] == 0) | (ij[:, 0] == n) | (ij[:, 1] == 0) | (ij[:, 1] == n))[0]

    def timed():
        ctx.forces_fill(GRAV, H)
        t = time.perf_counter()
        for _ in range(reps):
            ctx.forces_fill(GRAV, H)
        return (time.perf_counter() - t) / reps * 1e3

    ctx.upload_state(sc.x0, sc.v0)
    plain = timed()
    ctx.set_eol(eol)
    with_eol = timed()
    f = ctx.download_forces()
    ctx.set_eol([])
    out = {"workload": f"cloth {n}x{n}, the {len(eol)} boundary nodes as EOL nodes", "eol_nodes": int(len(eol)),
           "ms_fill_plain": plain, "ms_fill_eol": with_eol, "rows": int(f.size),
           "max_abs_f_X": float(np.max(np.abs(f[3 * sc.N:])))}
    log(f"EOL fill: {out}")
    return out


def run_batched(steps, warmup, maxit, multilevel=True):
    """The many-cloth workload: 256 cloths, sides uniform in {16..128} quads, E
    log-uniform in [1e2, 1e4] Pa, beta log-uniform in [1e-6, 1e-4], the single-cloth
    noise, seed 1237; one persistent launch per solve, graph-captured step, solve
    workgroup size from the batch autotuner.  Steps/s of the whole batch, with the
    block-Jacobi preconditioner and (from the same state) with the multilevel
    coarse levels added."""
    bs = scene.BatchScene()
    t = time.perf_counter()
    ctx = eolc.Context(0)
    bs.setup(ctx)
    ctx.set_options(eolc.OPT_SPD_PROJECT_BENDING)
    x0, v0 = ctx.download_state()
    setup_s = time.perf_counter() - t
    out = {"workload": f"{bs.B} cloths, sides {int(bs.sides.min())}..{int(bs.sides.max())} quads, "
                       f"{bs.N_real} vertices"}
    for name, ml in (("block_jacobi", False), ("multilevel", True)):
        if ml and not multilevel:
            continue
        ctx.set_batch_multilevel(ml)
        ctx.upload_state(x0, v0)
        best, ms = ctx.batch_autotune(GRAV, H, tol=TOL, maxit=maxit)
        log(f"batched {name}: {bs.B} cloths, {bs.N_real} vertices, set-up {setup_s:.1f} s, "
            f"autotune variant {best} ms {np.round(ms, 2).tolist()}")
        for _ in range(warmup):
            ctx.batch_step(GRAV, H, tol=TOL, maxit=maxit)
        t0 = time.perf_counter()
        its = []
        for _ in range(steps):
            rc, st = ctx.batch_step(GRAV, H, tol=TOL, maxit=maxit)
            bad = [s.as_dict() for s in st if not (s.true_relres <= TOL)]
            if rc != eolc.OK or bad:
                raise RuntimeError(f"batched solve missed the tolerance: {bad[:1]}")
            its.append([s.iters for s in st])
        dt = (time.perf_counter() - t0) / steps
        its = np.array(its)
        log(f"batched {name}: {1 / dt:.3f} steps/s, iterations per cloth min {its.min()} median "
            f"{int(np.median(its))} max {its.max()}")
        out[name] = {"steps_per_s": 1 / dt, "ms_per_step": dt * 1e3, "cloth_steps_per_s": bs.B / dt,
                     "iters_min": int(its.min()), "iters_median": int(np.median(its)),
                     "iters_max": int(its.max()), "solve_variant": best,
                     "variant_ms": [round(float(m), 3) for m in ms]}
    ctx.close()
    head = out["multilevel"] if multilevel else out["block_jacobi"]
    out.update({"steps_per_s": head["steps_per_s"], "ms_per_step": head["ms_per_step"],
                "preconditioner": "multilevel" if multilevel else "block_jacobi"})
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1024, help="quads per side")
    ap.add_argument("--maxit", type=int, default=60000)
    ap.add_argument("--window", type=int, default=12,
                    help="trajectory steps before the state is reset to the initial one")
    ap.add_argument("--aux-steps", type=int, default=3,
                    help="steps of the paths without the coarse levels (from the initial state, no warm-up)")
    ap.add_argument("--budget-s", type=float, default=420.0,
                    help="after this many seconds of measurement, the remaining optional legs are skipped")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64 block-Jacobi comparison")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-batch", action="store_true", help="skip the batched many-cloth workload")
    ap.add_argument("--no-collide", action="store_true", help="skip the proximity-query measurement")
    ap.add_argument("--no-eol", action="store_true", help="skip the EOL-rows fill measurement")
    ap.add_argument("--no-transfer", action="store_true", help="skip the velocity-transfer measurement")
    ap.add_argument("--no-ml", action="store_true",
                    help="headline without the multilevel coarse levels (and skip the multilevel lines)")
    ap.add_argument("--profile-solve", action="store_true",
                    help="kernel-trace runs: one step of each solve mode capped at 400 PCG iterations, then exit")
    ap.add_argument("--ml-linear", type=int, default=None, choices=(0, 1, 2),
                    help="linear coarse fields on the multilevel levels >= 3 for every leg "
                         "(eolc_set_ml_linear; default: the library's, 1 = in FP32 iterations)")
    ap.add_argument("--profile-fill", action="store_true",
                    help="counter runs: 10 mixed-mode element evaluations + assemblies (the step's fill) and "
                         "a 1 GiB device copy for calibration, then exit")
    ap.add_argument("--profile-spmv", action="store_true",
                    help="counter runs: assemble once, 10 launches of each SpMV kernel and a 1 GiB "
                         "device copy for calibration, then exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU: the multi-GPU layout is a design only, and no figure "
              "at several GPUs is produced.", file=sys.stderr)
        sys.exit(2)

    t = time.perf_counter()
    sc = scene.Scene(args.n, sigma_x=1e-3, sigma_v=0.05)
    log(f"scene {args.n}^2: N={sc.N} F={sc.F} H={sc.H} ({time.perf_counter() - t:.1f} s)")
    ctx = eolc.Context(0)
    t = time.perf_counter()
    ctx.set_topology(sc.X, sc.tri, sc.hinge)
    ctx.set_material(0.1, 500.0, 0.3, 1e-5, 0.0, 1.0)
    ctx.set_options(eolc.OPT_SPD_PROJECT_BENDING)
    if args.ml_linear is not None:
        ctx.set_ml_linear(args.ml_linear)
    log(f"topology set-up {time.perf_counter() - t:.1f} s")
    # what a remesh costs on a live context: the pattern, SELL layout and gather maps
    # rebuilt and every per-topology device buffer replaced
    t = time.perf_counter()
    ctx.set_topology(sc.X, sc.tri, sc.hinge)
    rebuild_s = time.perf_counter() - t
    log(f"topology rebuild {rebuild_s:.2f} s")

    ctx.upload_state(sc.x0, sc.v0)
    ctx.forces_fill(GRAV, H)
    copy_ms = ctx.stream_copy(1 << 30, 10)
    copy_gbps = 2 * (1 << 30) / (copy_ms * 1e-3) / 1e9
    read_gbps = (1 << 30) / (ctx.stream_read(1 << 30, 10) * 1e-3) / 1e9
    log(f"copy kernel: 1 GiB read + 1 GiB written in {copy_ms:.3f} ms ({copy_gbps:.0f} GB/s); "
        f"read kernel {read_gbps:.0f} GB/s")
    tuned = {}
    for prec in (32, 64):
        best, ms = ctx.autotune(prec, reps=10)
        tuned[prec] = {"variant": best, "us": [round(float(m) * 1e3, 1) for m in ms]}
        log(f"autotune fp{prec}: variant {best}, us per launch {tuned[prec]['us']}")

    t = time.perf_counter()
    ml_counts = ctx.ml_info()
    ml_build_s = time.perf_counter() - t
    log(f"multilevel hierarchy: aggregates per level {ml_counts} ({ml_build_s:.2f} s)")

    if args.profile_solve:
        # one step of each mode, PCG capped at 400 iterations
        ML = eolc.SOLVE_MULTILEVEL
        for mode, name in ((eolc.SOLVE_MIXED_BJ, "mixed-bj"), (eolc.SOLVE_FP64_BJ, "fp64-bj"),
                           (eolc.SOLVE_MIXED_BJ | ML, "mixed-bj-ml"), (eolc.SOLVE_FP64_BJ | ML, "fp64-bj-ml")):
            ctx.upload_state(sc.x0, sc.v0)
            rc, st = ctx.step(GRAV, H, sc.fixed, tol=TOL, maxit=400, mode=mode, allow=(eolc.ERR_MAXIT,))
            log(f"{name}: {st.iters} iterations, solve {st.ms_solve:.1f} ms, assembly {st.ms_assembly:.1f} ms")
        ctx.close()
        return

    if args.profile_fill:
        # the step's fill as eolc_step runs it (face and hinge evaluation, assembly of
        # the FP64 matrix and the SpMV's FP32 layout, rhs), 10 times, then the copy
        for _ in range(10):
            ctx.forces_fill(GRAV, H)
        ms = ctx.stream_copy(1 << 30, 3)
        log(f"calibration copy: 1 GiB read + 1 GiB written, {ms:.3f} ms per launch")
        ctx.close()
        return

    if args.profile_spmv:
        # every SpMV variant of both precisions (10 launches each, after the 10 of
        # the autotuner), then the calibration copy: 1 GiB read + 1 GiB written
        for prec in (32, 64):
            for var in range(eolc.SPMV_VARIANTS):
                ctx.set_spmv_variant(prec, var)
                ms = ctx.time_spmv(prec, 10)
                log(f"spmv fp{prec} variant {var}: {ms * 1e3:.1f} us per launch")
        ms = ctx.stream_copy(1 << 30, 3)
        log(f"calibration copy: 1 GiB read + 1 GiB written, {ms:.3f} ms per launch "
            f"({2 * (1 << 30) / (ms * 1e-3) / 1e9:.0f} GB/s)")
        ctx.close()
        return

    theta, wsrc = load_weights()
    ctx.learned_set_weights(theta)
    nb = ctx.csr_nnz(eolc.MAT_MDK) // 9
    t_start = time.perf_counter()

    def over_budget(name):
        spent = time.perf_counter() - t_start
        if spent > args.budget_s:
            log(f"{name}: skipped, {spent:.0f} s spent of the {args.budget_s:.0f} s budget")
            return True
        return False

    def path_leg(mode, mixed, name, steps=None, warmup=None):
        def fn():
            k_steps = args.steps if steps is None else steps
            w_steps = args.warmup if warmup is None else warmup
            tt, st = run_path(ctx, sc, mode, k_steps, w_steps, args.maxit, name, args.window)
            return {"steps_per_s": 1 / tt, "ms_per_step": tt * 1e3, "steps": k_steps, "warmup": w_steps,
                    "iters": [s["iters"] for s in st],
                    "refinements": [s["refinements"] for s in st],
                    "launched": [s["launched"] for s in st], "host_syncs": [s["host_syncs"] for s in st],
                    "used_fallback": [s["used_fallback"] for s in st],
                    "true_relres_max": max(s["true_relres"] for s in st),
                    "phase_ms": {k: float(np.mean([s[k] for s in st])) for k in
                                 ("ms_assembly", "ms_rhs", "ms_solve", "ms_update")},
                    "roofline_step": step_roofline(sc, nb, st, tt * 1e3, mixed)}
        return fn

    legs = {}
    if args.no_ml:
        head_mode, head_name = eolc.SOLVE_MIXED_LEARNED, "mixed-learned"
        bj_name, bj_mode = "mixed_bj", eolc.SOLVE_MIXED_BJ
    else:
        head_mode, head_name = eolc.SOLVE_MIXED_LEARNED | eolc.SOLVE_MULTILEVEL, "mixed-learned-multilevel"
        bj_name, bj_mode = "mixed_bj_ml", eolc.SOLVE_MIXED_BJ | eolc.SOLVE_MULTILEVEL
    # The block-Jacobi run of the headline's preconditioner family goes first: the
    # learned solve falls back to block-Jacobi after twice its iteration count.
    leg(legs, bj_name, path_leg(bj_mode, True, bj_name))
    head = leg(legs, "headline", path_leg(head_mode, True, head_name))
    if head["status"] == "ok":
        log(f"{head_name}: {head['steps_per_s']:.4f} steps/s, iters {head['iters']}, "
            f"fallback {head['used_fallback']}, host syncs {head['host_syncs']}")
    spmv32_ms = ctx.time_spmv(32, 20)
    # the network trained on batched-path systems (its own leg: what training gains)
    if os.path.exists(TRAINED) and not over_budget("trained_learned"):
        th_t, src_t = load_weights(TRAINED)
        ctx.learned_set_weights(th_t)
        r = leg(legs, "trained_learned", path_leg(head_mode, True, "trained_learned"))
        r["weights"] = src_t
        ctx.learned_set_weights(theta)
        if r["status"] == "ok":
            log(f"trained_learned: {r['steps_per_s']:.4f} steps/s, iters {r['iters']}")
    # the baseline of the metric, then the coarse levels in FP64
    others = [("fp64_bj", eolc.SOLVE_FP64_BJ, False, args.no_fp64),
              ("fp64_bj_ml", eolc.SOLVE_FP64_BJ | eolc.SOLVE_MULTILEVEL, False, args.no_ml or args.no_fp64)]
    for name, mode, mixed, skip in others:
        if skip or over_budget(name):
            continue
        r = leg(legs, name, path_leg(mode, mixed, name))
        if r["status"] == "ok":
            log(f"{name}: {r['steps_per_s']:.4f} steps/s, iters {r['iters']}, replacements {r['refinements']}")
    spmv64_ms = ctx.time_spmv(64, 20)
    if head["status"] == "ok":
        for name, r in legs.items():
            if name != "headline" and r["status"] == "ok":
                r["speedup_of_headline"] = r["ms_per_step"] / head["ms_per_step"]
    # the dominant kernel: the algorithmic bytes of one product in the survey's unit
    # (the full difference-form block matrix, every off-diagonal block read by both of
    # its rows, plus the vectors) over its average launch time; beside it the bytes
    # its own layout must stream (the patch layout reads a block inside a patch once)
    layout = ctx.spmv_layout_info()

    def kernel_roofline(fp32, variant, ms):
        s = 4 if fp32 else 8
        full = spmv_bytes(sc.N, nb, s, True) if (fp32 or SPMV_VARIANTS[variant][4] in (4, 5)) else \
            spmv_bytes(sc.N, nb, s, False)
        if SPMV_VARIANTS[variant][4] == 5:
            own = (10 * s * layout["records"] + 2 * layout["lower_entries"] + 8 * layout["positions"] +
                   4 * layout["window_vertices"] + (4 * s * 3 + s + 1) * sc.N)
        elif SPMV_VARIANTS[variant][4] in (1, 4):   # symmetric records: upper records + lower records
            own = (10 * s) * (nb - sc.N) // 2 + 8 * (nb - sc.N) // 2 + 8 * sc.N + (4 * s * 3 + s + 1) * sc.N
        else:
            own = full
        kname = spmv_kernel_name(fp32, variant)
        return {"bound": "hbm", "achieved": full / (ms * 1e-3) / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
                "frac": full / (ms * 1e-3) / PEAK_HBM, "traffic": measured_traffic(kname, args.n),
                "kernel": kname, "variant": variant, "algorithmic_bytes": full, "avg_launch_ms": ms,
                "kernel_layout_bytes": own, "kernel_layout_GBps": own / (ms * 1e-3) / 1e9,
                "kernel_layout_frac": own / (ms * 1e-3) / PEAK_HBM}

    roof = kernel_roofline(True, tuned[32]["variant"], spmv32_ms)
    roof.update({"copy_kernel_GBps": copy_gbps, "read_kernel_GBps": read_gbps,
                 "fp64": kernel_roofline(False, tuned[64]["variant"], spmv64_ms), "patch_layout": layout})
    extra = {}
    if not args.no_eol and not over_budget("eol_fill"):
        leg(extra, "eol_fill", lambda: run_eol_fill(ctx, sc, args.n))
    ctx.close()
    cpu = None
    if not args.no_cpu and not over_budget("cpu_baseline"):
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        fp = legs.get("fp64_bj", {})
        iters = np.mean(fp["iters"]) if fp.get("status") == "ok" else 25000.0
        cpu = leg(extra, "cpu_baseline_leg", lambda: cpu_baseline(sc, iters, threads))
        extra.pop("cpu_baseline_leg")
        log(f"cpu baseline: {cpu}")
    if not over_budget("config1"):
        leg(extra, "config1", run_config1_gpu)
    if not args.no_collide and not over_budget("collision"):
        leg(extra, "collision", lambda: run_collision(10, not args.no_cpu))
    if not args.no_batch and not over_budget("batched"):
        leg(extra, "batched", lambda: run_batched(args.steps, args.warmup, args.maxit, not args.no_ml))
    if not args.no_transfer and not over_budget("velocity_transfer"):
        leg(extra, "velocity_transfer", lambda: run_transfer(sc))
    # the ingredients without the coarse levels, over the first --aux-steps steps of
    # the trajectory (no warm-up; mixed_bj first: it sets the learned solve's fallback cap)
    if not args.no_ml and args.aux_steps > 0:
        ctx = eolc.Context(0)
        ctx.set_topology(sc.X, sc.tri, sc.hinge)
        ctx.set_material(0.1, 500.0, 0.3, 1e-5, 0.0, 1.0)
        ctx.set_options(eolc.OPT_SPD_PROJECT_BENDING)
        if args.ml_linear is not None:
            ctx.set_ml_linear(args.ml_linear)
        ctx.upload_state(sc.x0, sc.v0)
        ctx.forces_fill(GRAV, H)
        ctx.set_spmv_variant(32, tuned[32]["variant"])
        ctx.set_spmv_variant(64, tuned[64]["variant"])
        ctx.learned_set_weights(theta)
        aux = {}
        for name, mode in (("mixed_bj", eolc.SOLVE_MIXED_BJ), ("mixed_learned", eolc.SOLVE_MIXED_LEARNED),
                           ("fp64_bj_first_steps", eolc.SOLVE_FP64_BJ),
                           ("mixed_learned_ml_first_steps", head_mode)):
            if over_budget(name):
                continue
            r = leg(aux, name, path_leg(mode, name != "fp64_bj_first_steps", name, args.aux_steps, 0))
            if r["status"] == "ok":
                log(f"{name}: {r['steps_per_s']:.4f} steps/s over the first {args.aux_steps} steps, "
                    f"iters {r['iters']}")
        if aux.get("fp64_bj_first_steps", {}).get("status") == "ok":
            base = aux["fp64_bj_first_steps"]["ms_per_step"]
            for r in aux.values():
                if r.get("status") == "ok":
                    r["speedup_vs_fp64_bj"] = base / r["ms_per_step"]
        extra["first_steps"] = aux
        ctx.close()
    ok = head["status"] == "ok"
    res = {
        "metric": "steps/s, 1024^2 cloth at 1e-8 rel. residual: learned-PCG vs block-Jacobi PCG",
        "value": head.get("steps_per_s"), "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": head.get("ms_per_step"), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32-matrix/f64-residual",
        "data": "synthetic",
        "config": {"workload": f"cloth {args.n}x{args.n} quads, noise 1e-3 m, one step = "
                               f"assemble + solve to 1e-8 + update; trajectory steps 0..{args.window - 1} "
                               f"from the initial state, repeated",
                   "path": head_name if wsrc.endswith(".npy") else
                   f"{head_name} with zero-output weights (exactly the block-Jacobi blocks)",
                   "vertices": sc.N, "triangles": sc.F, "hinges": sc.H, "blocks": nb,
                   "bending_hessian": "spd-projected", "weights": wsrc,
                   "multilevel_aggregates": None if args.no_ml else ml_counts,
                   "multilevel_linear_fields": 1 if args.ml_linear is None else args.ml_linear},
        "speedup_vs_fp64_bj": legs.get("fp64_bj", {}).get("speedup_of_headline"),
        "roofline": roof,
        "roofline_step": head.get("roofline_step"),
        "cpu_baseline": cpu,
        "legs": legs,
        "multilevel_build_s": ml_build_s,
        "topology_rebuild_s": round(rebuild_s, 3),
        "autotune": tuned,
        **extra,
    }
    print(json.dumps(res), flush=True)
    if not ok:
        sys.exit(1)


if __name__ == "__main__":
    main()
