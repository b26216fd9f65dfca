"""Benchmark: 3D Poisson 128^3 (7-point, n = 2,097,152, nnz = 14,581,760).

One step = one solve to a relative residual of 1e-8 from x0 = 0 with the
FP32-inner / FP64-refinement CG (Jacobi preconditioned), inputs resident in
HBM.  Reported beside it, in the same run: FP64 Jacobi-PCG time-to-solution,
standalone SpMV GB/s (f64 and f32, Infinity Cache flushed between reps), 50
weighted-Jacobi sweeps (omega = 2/3), and the compiled reference CPU path
(oracle/_ref) solving the same system on one host core.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N_SIDE = 128


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def spmv_bytes(n: int, nnz: int, s: int) -> int:
    # (s+4) nnz + 4 (n+1) + 2 s n : values + u32 columns + row pointers + x, y
    return (s + 4) * nnz + 4 * (n + 1) + 2 * s * n


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--spmv-reps", type=int, default=100)
    ap.add_argument("--spmv-warmup", type=int, default=20)
    ap.add_argument("--spmv-only", action="store_true",
                    help="standalone SpMV and a 256 MiB copy only (for profiler counter runs)")
    ap.add_argument("--inner-rtol", type=float, default=1e-4)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        return 2

    import torch
    from faspxx_amd import fx
    from faspxx_amd import problems as P

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    # a dedicated stream (default priority): the legacy null stream cannot
    # be captured into the solvers' replay graphs
    stream = torch.cuda.Stream(device=dev)
    torch.cuda.set_stream(stream)
    ctx = fx.Context(0, stream=stream.cuda_stream)

    t0 = time.time()
    pr = P.config(2)
    A64 = pr.A
    n, nnz = A64.nrow, A64.nnz
    log(f"generated {pr.name}: n={n} nnz={nnz} in {time.time() - t0:.1f}s")
    t0 = time.time()
    A = fx.Csr.from_problem(ctx, A64)
    A.make_f32()
    b = fx.Vec.from_numpy(ctx, pr.b)
    x = fx.Vec(ctx, n)
    ctx.sync()
    setup_s = time.time() - t0
    log(f"uploaded + row-blocked in {setup_s:.2f}s")

    def ev():
        return torch.cuda.Event(enable_timing=True)

    # ---- standalone SpMV, Infinity Cache flushed between reps ------------
    # Flush by READING 512 MB (twice the 256 MB Infinity Cache): a write-based
    # flush would leave dirty lines whose write-back lands inside the timed SpMV.
    flush = torch.ones(512 * 1024 * 1024 // 4, dtype=torch.float32, device=dev)
    flush_sink = torch.empty((), dtype=torch.float32, device=dev)
    res_spmv = {}
    for dt, s in ((fx.F64, 8), (fx.F32, 4)):
        xv = fx.Vec.from_numpy(ctx, np.random.default_rng(1).uniform(-1, 1, n),
                               dt)
        yv = fx.Vec(ctx, n, dt)
        # every launch of this kernel is flushed (warm-up ones included), so
        # a profiler's average over the kernel matches the mean below
        times = []
        for i in range(args.spmv_warmup + args.spmv_reps):
            torch.sum(flush, dim=0, out=flush_sink)
            e0, e1 = ev(), ev()
            e0.record(stream)
            A.Apply(xv, yv)
            e1.record(stream)
            e1.synchronize()
            if i >= args.spmv_warmup:
                times.append(e0.elapsed_time(e1))
        med, mean = statistics.median(times), statistics.mean(times)
        byts = spmv_bytes(n, nnz, s)
        res_spmv["f64" if dt == fx.F64 else "f32"] = {
            "ms_median_flushed": med, "ms_mean_flushed": mean,
            "GBs_flushed": byts / mean / 1e6, "GBs_median_flushed": byts / med / 1e6,
            "GFLOPs_flushed": 2 * nnz / mean / 1e6, "bytes": byts, "reps": args.spmv_reps,
        }
        log(f"SpMV {'f64' if dt == fx.F64 else 'f32'}: {mean * 1e3:.1f} us mean flushed "
            f"({byts / mean / 1e6:.0f} GB/s), median {med * 1e3:.1f} us")
        del xv, yv
    if args.spmv_only:
        # a plain copy of known size: the scale for profiler byte counters
        src = torch.ones(64 * 1024 * 1024, dtype=torch.float32, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.sum(flush, dim=0, out=flush_sink)
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"spmv": res_spmv, "copy_bytes_each_way": src.numel() * 4}), flush=True)
        return 0
    del flush
    torch.cuda.empty_cache()

    # ---- 50 weighted-Jacobi sweeps, omega = 2/3 --------------------------
    pj = fx.Precond(ctx, "jacobi", A, 2.0 / 3.0, 1)
    sj = fx.Solver(ctx, fx.JACOBI)
    sj.Setup(A, pj, fx.params(maxIter=50, relTol=1e-30, absTol=0.0))
    jac_ms = []
    for _ in range(2):
        x.set(0.0)
        r = sj.Solve(b, x)
        jac_ms.append(r.solve_ms)
    sweep_bytes = 12 * nnz + 4 * (n + 1) + 4 * 8 * n
    jac = {"ms_50_sweeps": min(jac_ms), "iters": r.numIter, "final_norm2": r.norm2,
           "GBs_per_sweep": sweep_bytes * 50 / min(jac_ms) / 1e6}
    log(f"Jacobi 50 sweeps: {min(jac_ms):.2f} ms ({jac['GBs_per_sweep']:.0f} GB/s)")

    # ---- FP64 Jacobi-PCG to rtol 1e-8 ------------------------------------
    pc = fx.Precond(ctx, "jacobi", A, 1.0, 1)
    scg = fx.Solver(ctx, fx.CG)
    scg.Setup(A, pc, fx.params(maxIter=20000, relTol=1e-8, absTol=0.0))
    cg_ms = []
    for _ in range(max(2, args.warmup + 1)):
        x.set(0.0)
        rcg = scg.Solve(b, x)
        cg_ms.append(rcg.solve_ms)
    xr = x.numpy()
    cg_relres = float(np.linalg.norm(pr.b - A64.matvec(xr)) / np.linalg.norm(pr.b))
    cg_iter_bytes = 12 * nnz + 4 * (n + 1) + 13 * 8 * n
    cg = {"ms": min(cg_ms), "iters": rcg.numIter, "relres_fp64": cg_relres,
          "GBs_per_iter": cg_iter_bytes * rcg.numIter / min(cg_ms) / 1e6}
    log(f"FP64 PCG: {min(cg_ms):.2f} ms, {rcg.numIter} iters, relres {cg_relres:.2e}")

    # ---- timed steps: FP32-inner / FP64-refined CG ------------------------
    outer = fx.params(maxIter=50, relTol=1e-8, absTol=0.0)
    inner = fx.params(maxIter=20000, relTol=args.inner_rtol, absTol=0.0)
    ir = fx.IR(ctx, A, 1.0, outer, inner)
    for _ in range(args.warmup):
        x.set(0.0)
        ir.Solve(b, x)
    torch.cuda.synchronize()
    step_ms, inner_its, outer_its = [], 0, 0
    t_start = time.perf_counter()
    for _ in range(args.steps):
        x.set(0.0)
        rir, inner_its = ir.Solve(b, x)
        step_ms.append(rir.solve_ms)
        outer_its = rir.numIter
    torch.cuda.synchronize()
    wall_ms = (time.perf_counter() - t_start) * 1e3 / args.steps
    xr = x.numpy()
    ir_relres = float(np.linalg.norm(pr.b - A64.matvec(xr)) / np.linalg.norm(pr.b))
    ir_ms = statistics.median(step_ms)
    log(f"IR: {ir_ms:.2f} ms/solve ({outer_its} outer, {inner_its} inner), relres {ir_relres:.2e}")

    # ---- CPU baseline: the compiled reference's Jacobi-PCG, full solve -----
    cpu = None
    if not args.no_cpu:
        try:
            sys.path.insert(0, os.path.join(ROOT, "oracle"))
            from oracle import CG as RCG, Matrix, Reference
            Rf = Reference()
            M = Matrix.of(A64)
            out = Rf.solve(RCG, M, pr.b, np.zeros(n), pc="jacobi", weight=1.0, maxIter=20000,
                           relTol=1e-8, absTol=0.0)
            cpu = {"value": round(out["solve_ms"], 1), "unit": "ms", "cores": 1,
                   "kind": "reference",
                   "sample": (f"reference CG::Solve with a Jacobi preconditioner (oracle/_ref, "
                              f"-O2 -march=x86-64-v3, OpenMP off, 1 thread) on the same 128^3 "
                              f"system to rtol 1e-8: {out['numIter']} iterations, "
                              f"steady_clock around Solve only; retcode {out['retcode']}"),
                   "iters": out["numIter"], "speedup_gpu_fp64_cg": round(out["solve_ms"] / cg["ms"], 1)}
            log(f"CPU reference: {out['solve_ms']:.0f} ms, {out['numIter']} iters")
        except Exception as e:  # reference library absent on this host
            cpu = {"value": None, "unit": "ms", "cores": 1, "kind": "reference",
                   "sample": f"unavailable: {e}"}

    # HBM bytes per launch of the f64 SpMV, from the committed profiler
    # counter summary (FETCH_SIZE / WRITE_SIZE runs of `bench.py --spmv-only`,
    # scaled by a copy of known size); null when it was taken on another
    # column-index format
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "k_spmv_f64_poisson3d_128_traffic.json")
    if os.path.exists(tpath):
        with open(tpath) as fh:
            tr = json.load(fh)
        if tr.get("index_bytes") == A.index_bytes and tr.get("nnz") == nnz:
            traffic = tr["traffic_bytes"]

    f64 = res_spmv["f64"]
    line = {
        "metric": "CG time-to-rtol 1e-8 + SpMV GB/s, 3D Poisson 128³; FP32-refined vs FP64",
        "value": round(ir_ms, 4),
        "unit": "ms",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ir_ms, 4),
        "wall_ms_per_step": round(wall_ms, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32-inner/f64-outer",
        "data": "synthetic (3D Poisson 7-point, x* ~ U[-1,1] seed 1, b = A x*)",
        "config": {"workload": "poisson3d_128 Jacobi-PCG to rtol 1e-8", "n": n, "nnz": nnz,
                   "inner_rtol": args.inner_rtol, "outer_iters": outer_its,
                   "inner_iters": inner_its},
        "relres_fp64": ir_relres,
        "fp64_cg": cg,
        "speedup_ir_vs_fp64": round(cg["ms"] / ir_ms, 3),
        "spmv": res_spmv,
        "jacobi": jac,
        "setup_s": round(setup_s, 3),
        "roofline": {"bound": "hbm", "kernel": "k_spmv f64 (MAT::Apply)",
                     "achieved": round(f64["GBs_flushed"], 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(f64["GBs_flushed"] / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "algorithmic_bytes": f64["bytes"],
                     "index_bytes": A.index_bytes},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
