#!/usr/bin/env python3
"""Benchmark of the HiFlow3 sparse hot path on one MI355X.

Default workload (configuration 2): plain FP64 CG with the Jacobi
preconditioner on the 3D Q1 trilinear FE Laplacian of a 128^3 hexahedral mesh
(127^3 = 2,048,383 interior DOFs, 27-point CSR, 54,439,939 nonzeros, int32
indices), seeded uniform[-1,1] right-hand side, x0 = 0, relative tolerance
1e-8.  One step = one full solve with the matrix, vectors and preconditioner
resident in HBM.  `value` is the CSR SpMV bandwidth inside the timed solves:
algorithmic bytes per launch (12 nnz + 20 n + 4) over the mean SpMV launch
time, measured with HIP events on the context stream.

Prints one JSON line.  --gpus N > 1 is not provided: this build runs on one GPU.
"""
import argparse
import importlib.util
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG_DIR = os.path.join(ROOT, "hiflow3_2.0_amd")
HBM_PEAK_GBS = 8000.0


def load_pkg():
    if "hiflow3_amd" in sys.modules:
        return sys.modules["hiflow3_amd"]
    spec = importlib.util.spec_from_file_location("hiflow3_amd", os.path.join(PKG_DIR, "__init__.py"),
                                                  submodule_search_locations=[PKG_DIR])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hiflow3_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def spmv_bytes(n, nnz, val_bytes=8, idx_bytes=4):
    return (val_bytes + idx_bytes) * nnz + (idx_bytes + 2 * val_bytes) * n + idx_bytes


def read_traffic(name):
    """HBM bytes per launch measured by rocprofv3 PMC runs (profiles/<name>), or None."""
    path = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if "=" in line and not line.lstrip().startswith("#"):
            k, v = line.split("=", 1)
            vals[k.strip()] = v.strip()
    try:
        return float(vals["traffic_bytes_per_launch"])
    except (KeyError, ValueError):
        return None


def cpu_spmv_baseline(row, col, val, x, budget_s=12.0):
    """The oracle's OpenMP CSR SpMV (CPUopenmp_CSR_lMatrix::VectorMult) on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle as O
    A = O.Csr(row, col, val)
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    y = O.spmv(A, x, omp=True)  # warm-up
    times = []
    t_end = time.perf_counter() + budget_s
    while time.perf_counter() < t_end and len(times) < 50:
        t0 = time.perf_counter()
        y = O.spmv(A, x, omp=True)
        times.append(time.perf_counter() - t0)
    med = float(np.median(times))
    return y, med, len(times), threads


def bench_c2(P, args):
    nx = 127
    t0 = time.perf_counter()
    row, col, val, rhs = P.problem(P.Q1_3D, nx, seed=args.seed)
    n, nnz = len(row) - 1, len(col)
    log(f"[bench] generated 3D Q1 {nx + 1}^3 cells: n={n} nnz={nnz} ({time.perf_counter() - t0:.1f}s)")
    ctx = P.Context(0)
    A = P.CsrMatrix(ctx, row, col, val)
    M = P.Preconditioner.jacobi(A)
    B = P.Vector(ctx, rhs)
    X = P.Vector(ctx, n)
    log(f"[bench] uploaded; device {ctx.info()}")

    # SpMV parity at full size against the oracle (bitwise for rows within one tile)
    xr = np.random.default_rng(1).uniform(-1, 1, n)
    Xr, Yr = P.Vector(ctx, xr), P.Vector(ctx, n)
    A.spmv(Xr, Yr)
    y_gpu = Yr.numpy()
    cpu = None
    if not args.no_cpu:
        y_cpu, med, reps, threads = cpu_spmv_baseline(row, col, val, xr, args.cpu_budget)
        cpu = {"value": spmv_bytes(n, nnz) / med / 1e9, "unit": "GB/s", "cores": threads, "kind": "port",
               "sample": f"oracle OpenMP CSR SpMV (CPUopenmp_CSR_lMatrix order) on the full 128^3 Q1 matrix, "
                         f"median of {reps} calls, {med * 1e3:.1f} ms per call"}
        rel = float(np.linalg.norm(y_gpu - y_cpu) / np.linalg.norm(y_cpu))
        log(f"[bench] cpu spmv {med * 1e3:.1f} ms ({threads} threads); gpu vs oracle rel {rel:.2e}, "
            f"bitwise {np.array_equal(y_gpu, y_cpu)}")
    else:
        rel = None
    del Xr, Yr

    kw = dict(maxits=5000, atol=0.0, rtol=1e-8, dtol=1e6)
    for w in range(args.warmup):
        X.zeros()
        info = P.pcg(A, B, X, M, **kw)
        log(f"[bench] warmup {w}: iters={info['iters']} {info['time_ms']:.1f} ms")
    ctx.set_timing(True)
    ctx.timing_reset()
    ctx.sync()
    t0 = time.perf_counter()
    infos = []
    for s in range(args.steps):
        X.zeros()
        infos.append(P.pcg(A, B, X, M, **kw))
    ctx.sync()
    wall = time.perf_counter() - t0
    launches, spmv_ms = ctx.timing(0)
    ctx.set_timing(False)
    iters = infos[-1]["iters"]
    x = X.numpy()
    import scipy.sparse as sp
    true_rel = float(np.linalg.norm(rhs - sp.csr_matrix((val, col, row)) @ x) / np.linalg.norm(rhs))
    avg_spmv_ms = spmv_ms / max(launches, 1)
    b_spmv = spmv_bytes(n, nnz)
    gbs = b_spmv / (avg_spmv_ms * 1e-3) / 1e9
    ms_step = wall * 1e3 / args.steps
    ms_iter = ms_step / max(iters, 1)
    it_bytes = 12 * nnz + 108 * n
    traffic = read_traffic("spmv_q1_3d_128_traffic.txt")
    log(f"[bench] {args.steps} solves: {ms_step:.2f} ms/solve, {iters} iters, {ms_iter * 1e3:.1f} us/iter, "
        f"spmv {avg_spmv_ms * 1e3:.1f} us ({gbs:.0f} GB/s), true rel residual {true_rel:.2e}")
    return {
        "metric": "SpMV GB/s (% HBM peak); CG time to 1e-10 rel; served solves/s + p99",
        "value": round(gbs, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded 3D Q1 FE Laplacian, uniform[-1,1] rhs)",
        "config": {"workload": "cfg2: Jacobi-PCG FP64, 3D Q1 Laplacian 128^3 cells, rtol 1e-8",
                   "n": n, "nnz": nnz, "index": "int32"},
        "roofline": {"bound": "hbm", "achieved": round(gbs, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(gbs / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "kernel": "k_spmv_sell<double>", "bytes_per_launch": b_spmv,
                     "avg_launch_us": round(avg_spmv_ms * 1e3, 2), "launches": launches},
        "cpu_baseline": cpu,
        "cg": {"iters": iters, "ms_per_iter": round(ms_iter, 4), "solve_ms": round(ms_step, 3),
               "iter_GBps_compulsory": round(it_bytes / (ms_iter * 1e-3) / 1e9, 1),
               "true_rel_residual": true_rel, "state": infos[-1]["state"]},
        "parity": {"spmv_rel_vs_oracle": rel},
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", default="c2", choices=["c2"])
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=12.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus N > 1 is not provided",
                          "n_gpus": args.gpus}), flush=True)
        sys.exit(2)
    P = load_pkg()
    out = bench_c2(P, args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
