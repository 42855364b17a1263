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
(EOLC_SOLVE_MULTILEVEL), block-Jacobi fallback).  FP64 3x3 block-Jacobi PCG (the
baseline), mixed block-Jacobi PCG, and both block-Jacobi precisions with the coarse
levels run from the same initial state, so the gain of each ingredient is stated.  Before timing, the SpMV launch
variant of each precision is chosen by the library's autotuner on the first
matrix.  The noisy initial state makes the exact bending Hessian indefinite, so
every path assembles the SPD-projected hinge Hessians
(EOLC_OPT_SPD_PROJECT_BENDING).

Prints one JSON line (rank 0) with roofline and cpu_baseline objects.  The
roofline's `traffic` (HBM bytes per SpMV launch) is read from
profiles/pmc_spmv_cloth1024.json, the calibrated rocprofv3 FETCH_SIZE/WRITE_SIZE
summary of `bench.py --profile-spmv`, when it was measured for the same kernel
variant and size.
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


def run_path(ctx, sc, mode, steps, warmup, maxit):
    ctx.upload_state(sc.x0, sc.v0)
    for _ in range(warmup):
        ctx.step(GRAV, H, sc.fixed, tol=TOL, maxit=maxit, mode=mode)
    t0 = time.perf_counter()
    stats = []
    for _ in range(steps):
        rc, st = ctx.step(GRAV, H, sc.fixed, tol=TOL, maxit=maxit, mode=mode)
        stats.append(st.as_dict())
    t1 = time.perf_counter()
    bad = [s for s in stats if not (s["true_relres"] <= TOL)]
    if bad:
        raise RuntimeError(f"solve missed the tolerance: {bad[0]}")
    return (t1 - t0) / steps, stats


def cpu_baseline(sc, iters_per_step, threads):
    """The oracle (FP64 CPU restatement) on a bounded sample of the same workload:
    one full Forces::fill (elements + triplets + setFromTriplets-style CSR) and 10
    iterations of its block-Jacobi PCG on the same matrix, scaled to steps/s with the
    FP64 iteration count the GPU needed per step."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    mat = O.material()
    t0 = time.perf_counter()
    rc, f, M, K = O.forces_fill(sc.x0, sc.X, sc.tri, sc.hinge, mat, GRAV, H, nthreads=threads)
    t1 = time.perf_counter()
    rhs = M @ sc.v0 + H * f
    d = (3 * sc.fixed[:, None] + np.arange(3)).reshape(-1).astype(np.int32)
    k = 10
    t2 = time.perf_counter()
    O.velocity_solve(K, rhs, d, 0.99 * sc.v0[d], v0=sc.v0, tol=1e-30, maxit=k,
                     precond=O.PC_BLOCK_JACOBI, nthreads=threads)
    t3 = time.perf_counter()
    t_iter = (t3 - t2) / k
    t_step = (t1 - t0) + iters_per_step * t_iter
    # one thread, like the single-threaded reference: Forces::fill on a 256x256 sheet
    # scaled by the element count, and 2 PCG iterations on the full matrix
    small = scene.Scene(256, sigma_x=1e-3, sigma_v=0.05)
    t4 = time.perf_counter()
    O.forces_fill(small.x0, small.X, small.tri, small.hinge, mat, GRAV, H, nthreads=1)
    fill1 = (time.perf_counter() - t4) * (sc.F + sc.H) / (small.F + small.H)
    t5 = time.perf_counter()
    O.velocity_solve(K, rhs, d, 0.99 * sc.v0[d], v0=sc.v0, tol=1e-30, maxit=2, precond=O.PC_BLOCK_JACOBI,
                     nthreads=1)
    iter1 = (time.perf_counter() - t5) / 2
    step1 = fill1 + iters_per_step * iter1
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            model = next((ln.split(":", 1)[1].strip() for ln in fh if ln.startswith("model name")), None)
    except OSError:
        pass
    return {"value": 1.0 / t_step, "unit": "steps/s", "cores": threads, "kind": "port",
            "cpu_model": model, "nproc": os.cpu_count(),
            "sample": f"oracle Forces::fill once + {k} FP64 block-Jacobi PCG iterations at "
                      f"{sc.n}x{sc.n} ({(t1 - t0):.2f} s fill, {t_iter * 1e3:.1f} ms/iter), "
                      f"scaled with {iters_per_step:.0f} iterations/step",
            "fill_s": t1 - t0, "iter_ms": t_iter * 1e3,
            "one_thread": {"value": 1.0 / step1, "unit": "steps/s", "cores": 1,
                           "sample": "Forces::fill of a 256x256 sheet scaled by element count "
                                     f"({fill1:.1f} s at 1024^2) + 2 PCG iterations at 1024^2 "
                                     f"({iter1 * 1e3:.0f} ms/iter)"}}


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
TRAFFIC = os.path.join(ROOT, "profiles", "pmc_spmv_cloth1024.json")


def load_weights():
    from eolc import learned
    if os.path.exists(WEIGHTS):
        th = np.load(WEIGHTS, allow_pickle=False).astype(np.float32)
        if th.size == eolc.LEARNED_NPARAMS:
            return th, os.path.relpath(WEIGHTS, ROOT)
    return learned.init_weights(7), "initial weights (zero output = block-Jacobi)"


# SpMV launch variants: (entries per trip, column prefetch, non-temporal loads,
# waves per SIMD, symmetric layout), in the library's order
SPMV_VARIANTS = [(1, False, False, 8, 0), (1, True, False, 8, 0), (1, True, True, 8, 0),
                 (1, False, True, 8, 0), (1, True, True, 6, 0), (2, True, True, 6, 0),
                 (2, True, False, 6, 0), (2, True, True, 4, 0), (1, True, True, 4, 0),
                 (1, False, False, 8, 1), (1, False, False, 6, 1), (1, False, True, 8, 1),
                 (1, False, False, 8, 2), (1, False, True, 8, 2), (1, False, False, 6, 2),
                 (1, False, False, 8, 3), (1, False, True, 8, 3), (1, False, True, 6, 3)]


def spmv_kernel_name(fp32, variant):
    u, pf, nt, w, lay = SPMV_VARIANTS[variant]
    b = lambda x: "true" if x else "false"
    t = "float, true" if fp32 else "double, false"
    return f"k_pcg_spmv<{t}, {u}, {b(pf)}, {b(nt)}, {w}, {lay if (fp32 or lay == 3) else 0}>"


def measured_traffic(kernel, n):
    """HBM bytes per launch of `kernel` at n x n from the committed counter summary."""
    if n != 1024:
        return None
    try:
        with open(TRAFFIC) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    for k in t.get("kernels", []):
        if k.get("kernel_template") == kernel:
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
    ap.add_argument("--maxit", type=int, default=200000)
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
        for mode, name in ((eolc.SOLVE_MIXED_BJ, "mixed-bj"), (eolc.SOLVE_FP64_BJ, "fp64-bj"),
                           (eolc.SOLVE_MIXED_BJ | eolc.SOLVE_MULTILEVEL, "mixed-bj-ml")):
            ctx.upload_state(sc.x0, sc.v0)
            rc, st = ctx.step(GRAV, H, sc.fixed, tol=TOL, maxit=400, mode=mode, allow=(eolc.ERR_MAXIT,))
            log(f"{name}: {st.iters} iterations, solve {st.ms_solve:.1f} ms, assembly {st.ms_assembly:.1f} ms")
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
    # the block-Jacobi runs go first: the learned solve falls back to block-Jacobi
    # after twice the iterations of the last block-Jacobi solve on the context
    out_fp64 = None
    if not args.no_fp64:
        t64, st64 = run_path(ctx, sc, eolc.SOLVE_FP64_BJ, args.steps, args.warmup, args.maxit)
        log(f"fp64-bj: {1 / t64:.4f} steps/s, iters {[s['iters'] for s in st64]}")
        spmv64_ms = ctx.time_spmv(64, 20)
        out_fp64 = {"steps_per_s": 1 / t64, "iters": [s["iters"] for s in st64],
                    "ms_per_step": t64 * 1e3, "spmv_ms": spmv64_ms}
    tbj, stbj = run_path(ctx, sc, eolc.SOLVE_MIXED_BJ, args.steps, args.warmup, args.maxit)
    log(f"mixed-bj: {1 / tbj:.4f} steps/s, iters {[s['iters'] for s in stbj]}")
    out_mbj = {"steps_per_s": 1 / tbj, "iters": [s["iters"] for s in stbj], "ms_per_step": tbj * 1e3}
    extra = {}
    if not args.no_ml:
        # the coarse levels alone, in both precisions (the mixed one also sets the
        # learned mode's fallback cap for the multilevel case)
        for name, mode in (("fp64_bj_ml", eolc.SOLVE_FP64_BJ | eolc.SOLVE_MULTILEVEL),
                           ("mixed_bj_ml", eolc.SOLVE_MIXED_BJ | eolc.SOLVE_MULTILEVEL)):
            if name == "fp64_bj_ml" and args.no_fp64:
                continue
            tt, st = run_path(ctx, sc, mode, args.steps, args.warmup, args.maxit)
            log(f"{name}: {1 / tt:.4f} steps/s, iters {[s['iters'] for s in st]}")
            extra[name] = {"steps_per_s": 1 / tt, "iters": [s["iters"] for s in st], "ms_per_step": tt * 1e3,
                           "launched": [s["launched"] for s in st],
                           "phase_ms": {k: float(np.mean([s[k] for s in st])) for k in
                                        ("ms_assembly", "ms_solve")}}
    if args.no_ml:
        head_mode, head_name = eolc.SOLVE_MIXED_LEARNED, "mixed-learned"
    else:
        head_mode, head_name = eolc.SOLVE_MIXED_LEARNED | eolc.SOLVE_MULTILEVEL, "mixed-learned-multilevel"
    t_head, st_head = run_path(ctx, sc, head_mode, args.steps, args.warmup, args.maxit)
    log(f"{head_name}: {1 / t_head:.4f} steps/s, iters {[s['iters'] for s in st_head]}, "
        f"fallback {[s['used_fallback'] for s in st_head]}, launched {[s['launched'] for s in st_head]}, "
        f"host syncs {[s['host_syncs'] for s in st_head]}")
    spmv32_ms = ctx.time_spmv(32, 20)
    if out_fp64:
        out_fp64["speedup_of_headline"] = t64 / t_head
    out_mbj["speedup_of_headline"] = tbj / t_head
    for v in extra.values():
        v["speedup_of_headline"] = (1 / v["steps_per_s"]) / t_head
    nb = None
    # nb from the matrix pattern size: 9 scalars per block
    import ctypes as C
    nnz = C.c_int64(0)
    eolc.lib().eolc_download_csr(ctx._h, eolc.MAT_MDK, None, None, None, 0, C.byref(nnz))
    nb = nnz.value // 9
    # algorithmic bytes over the SELL storage actually streamed (padding excluded)
    alg = spmv_bytes(sc.N, nb, 4, True)
    achieved = alg / (spmv32_ms * 1e-3)
    kname = spmv_kernel_name(True, tuned[32]["variant"])
    roof = {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
            "frac": achieved / PEAK_HBM, "traffic": measured_traffic(kname, args.n),
            "kernel": kname, "variant": tuned[32]["variant"], "algorithmic_bytes": alg,
            "avg_launch_ms": spmv32_ms, "copy_kernel_GBps": copy_gbps, "read_kernel_GBps": read_gbps}
    eol_out = None if args.no_eol else run_eol_fill(ctx, sc, args.n)
    ctx.close()
    batched = None if args.no_batch else run_batched(args.steps, args.warmup, args.maxit, not args.no_ml)
    collision = None if args.no_collide else run_collision(10, not args.no_cpu)
    transfer = None if args.no_transfer else run_transfer(sc)
    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or min(16, os.cpu_count() or 1)
        iters = np.mean(out_fp64["iters"]) if out_fp64 else np.mean([s["iters"] for s in st_head])
        cpu = cpu_baseline(sc, iters, threads)
        log(f"cpu baseline: {cpu}")
    res = {
        "metric": "steps/s, 1024^2 cloth at 1e-8 rel. residual: learned-PCG vs block-Jacobi PCG",
        "value": 1.0 / t_head, "unit": "steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": t_head * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32-matrix/f64-residual",
        "data": "synthetic",
        "config": {"workload": f"cloth {args.n}x{args.n} quads, noise 1e-3 m, one step = "
                               f"assemble + solve to 1e-8 + update", "path": head_name,
                   "vertices": sc.N, "triangles": sc.F, "hinges": sc.H, "blocks": nb,
                   "bending_hessian": "spd-projected", "weights": wsrc,
                   "multilevel_aggregates": None if args.no_ml else ml_counts},
        "iters": [s["iters"] for s in st_head],
        "launched": [s["launched"] for s in st_head],
        "host_syncs": [s["host_syncs"] for s in st_head],
        "used_fallback": [s["used_fallback"] for s in st_head],
        "true_relres": [s["true_relres"] for s in st_head],
        "phase_ms": {k: float(np.mean([s[k] for s in st_head])) for k in
                     ("ms_assembly", "ms_rhs", "ms_solve", "ms_update")},
        "fp64_bj": out_fp64,
        "mixed_bj": out_mbj,
        **extra,
        "multilevel_build_s": ml_build_s,
        "topology_rebuild_s": round(rebuild_s, 3),
        "eol_fill": eol_out,
        "velocity_transfer": transfer,
        "autotune": tuned,
        "roofline": roof,
        "cpu_baseline": cpu,
        "batched": batched,
        "collision": collision,
    }
    print(json.dumps(res), flush=True)


if __name__ == "__main__":
    main()
