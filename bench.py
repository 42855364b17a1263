#!/usr/bin/env python3
"""Benchmark: steps/s of the per-step cloth hot path on one MI355X.

Workload: the 1024x1024-quad cloth (1,050,625 vertices, 2,097,152 triangles,
3,143,680 hinges), x0 = grid + N(0, 1e-3^2), v0 ~ N(0, 0.05^2) (seeds 1234 and
1235), two top corners fixed, rho 0.1, E 500, nu 0.3, beta 1e-5, h 5e-3,
gravity -9.8.  One step = element evaluation + atomic-free assembly of M, MDK, f +
rhs + velocity solve to 1e-8 relative residual (FP64, reduced system) + x += h v,
all on the device with the state resident in HBM.  Each step's solve is checked
against the tolerance.

The headline path is the mixed-precision PCG (FP32 matrix and vectors, FP64 dots
and true residual with reliable updates); FP64 3x3 block-Jacobi PCG runs on the
same initial state for comparison.  The noisy initial state makes the exact
bending Hessian indefinite, so both paths assemble the SPD-projected hinge
Hessians (EOLC_OPT_SPD_PROJECT_BENDING).

Prints one JSON line (rank 0) with roofline and cpu_baseline objects.
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
    """Algorithmic bytes of one k_pcg_spmv launch over the nb stored blocks: block values
    (9 scalars, plus the mass scalar in the FP32 difference form), column indices,
    per-row length and row index, z and p_old read, p_new and q written, fixed mask."""
    per_block = (10 if diff else 9) * s + 4
    return per_block * nb + 8 * N + 4 * s * 3 * N + N


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
    return {"value": 1.0 / t_step, "unit": "steps/s", "cores": threads, "kind": "port",
            "sample": f"oracle Forces::fill once + {k} FP64 block-Jacobi PCG iterations at "
                      f"{sc.n}x{sc.n} ({(t1 - t0):.2f} s fill, {t_iter * 1e3:.1f} ms/iter), "
                      f"scaled with {iters_per_step:.0f} iterations/step",
            "fill_s": t1 - t0, "iter_ms": t_iter * 1e3}


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
    ap.add_argument("--profile-spmv", action="store_true",
                    help="counter runs: assemble once, 10 launches of each SpMV kernel and a 1 GiB "
                         "device copy for calibration, then exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU; the multi-GPU layout is described in "
              "DESIGN.md and no figure at several GPUs is produced.", file=sys.stderr)
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

    if args.profile_spmv:
        ctx.upload_state(sc.x0, sc.v0)
        ctx.forces_fill(GRAV, H)
        for prec in (32, 64):
            ms = ctx.time_spmv(prec, 10)
            log(f"spmv fp{prec}: {ms * 1e3:.1f} us per launch")
        import torch
        a = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        b = torch.empty_like(a)
        torch.mul(a, 2.0, out=b)   # one elementwise kernel: reads 1 GiB, writes 1 GiB
        torch.cuda.synchronize()
        log("calibration kernel: 1 GiB read + 1 GiB written")
        ctx.close()
        return

    head_mode, head_name = eolc.SOLVE_MIXED_BJ, "mixed-bj"
    t_head, st_head = run_path(ctx, sc, head_mode, args.steps, args.warmup, args.maxit)
    log(f"{head_name}: {1 / t_head:.4f} steps/s, iters {[s['iters'] for s in st_head]}")
    spmv32_ms = ctx.time_spmv(32, 20)
    out_fp64 = None
    if not args.no_fp64:
        t64, st64 = run_path(ctx, sc, eolc.SOLVE_FP64_BJ, args.steps, args.warmup, args.maxit)
        log(f"fp64-bj: {1 / t64:.4f} steps/s, iters {[s['iters'] for s in st64]}")
        spmv64_ms = ctx.time_spmv(64, 20)
        out_fp64 = {"steps_per_s": 1 / t64, "iters": [s["iters"] for s in st64],
                    "ms_per_step": t64 * 1e3, "spmv_ms": spmv64_ms,
                    "speedup_of_headline": t64 / t_head}
    nb = None
    # nb from the matrix pattern size: 9 scalars per block
    import ctypes as C
    nnz = C.c_int64(0)
    eolc.lib().eolc_download_csr(ctx._h, eolc.MAT_MDK, None, None, None, 0, C.byref(nnz))
    nb = nnz.value // 9
    traffic = spmv_bytes(sc.N, nb, 4, True)
    achieved = traffic / (spmv32_ms * 1e-3)
    roof = {"bound": "hbm", "achieved": achieved / 1e9, "peak": PEAK_HBM / 1e9, "unit": "GB/s",
            "frac": achieved / PEAK_HBM, "traffic": None, "kernel": "k_pcg_spmv<float>",
            "bytes_per_launch": traffic, "avg_launch_ms": spmv32_ms}
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
                   "bending_hessian": "spd-projected"},
        "iters": [s["iters"] for s in st_head],
        "true_relres": [s["true_relres"] for s in st_head],
        "phase_ms": {k: float(np.mean([s[k] for s in st_head])) for k in
                     ("ms_assembly", "ms_rhs", "ms_solve", "ms_update")},
        "fp64_bj": out_fp64,
        "roofline": roof,
        "cpu_baseline": cpu,
    }
    print(json.dumps(res), flush=True)
    ctx.close()


if __name__ == "__main__":
    main()
