#!/usr/bin/env python3
"""Local solver benchmark on the HIP platform, with the command line of the
reference's examples/benchmark_local_solver/benchmark_local_solver.cc:226-311.

    benchmark_local_solver.py <A> <b> <x_result> <x_start> <solver> <interface>
                              <number of threads> <number of repetitions>

A is a MatrixMarket file, b / x_result / x_start are text vectors
(lMatrix::ReadFile, lVector::ReadFile).  solver: cg or gmres (basis size 5).
interface: hip (the device-resident drivers) or lmp (the unchanged
LinearSolver templates over HIP vectors and matrices).  The stopping test is
the reference's: relative 1e-8, absolute 1e-14, at most 1000 iterations.  The
last of the repetitions is timed; its result is compared with x_result.  The
number of threads is accepted for the reference's command line; the device
does not use it.
"""
import importlib.util
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REL_EPS, ABS_EPS, MAX_ITER, BASIS = 1e-8, 1e-14, 1000, 5


def _pkg():
    if "hiflow3_amd" in sys.modules:
        return sys.modules["hiflow3_amd"]
    spec = importlib.util.spec_from_file_location("hiflow3_amd", os.path.join(HERE, "__init__.py"),
                                                  submodule_search_locations=[HERE])
    mod = importlib.util.module_from_spec(spec)
    sys.modules["hiflow3_amd"] = mod
    spec.loader.exec_module(mod)
    return mod


class _Timer:
    """Prints the wall time of a block, as the reference's ScopedTimer does."""

    def __init__(self, what, out):
        self.what, self.out = what, out

    def __enter__(self):
        self.t0 = time.perf_counter()

    def __exit__(self, *exc):
        self.out.write(f"{self.what}: {time.perf_counter() - self.t0:.6f} s\n")


def run(argv, out=sys.stdout):
    """Returns the process exit status."""
    if len(argv) != 9:
        out.write(f"Usage: {argv[0] if argv else 'benchmark_local_solver.py'} <A> <b> <x_result> <x_start> "
                  "<solver> <interface> <number of threads> <number of repetitions>\n")
        return 1
    file_a, file_b, file_xr, file_x0, solver, interface = argv[1:7]
    reps = max(1, int(argv[8]))
    if solver not in ("cg", "gmres"):
        out.write(f"Solver {solver} not supported.\n")
        return 1
    if interface not in ("hip", "lmp"):
        out.write(f"Algo {interface} not supported.\n")
        return 1
    P = _pkg()
    with _Timer("Read vector b", out):
        b = P.vec_read(file_b)
    with _Timer("Read vector x", out):
        x0 = P.vec_read(file_x0)
    with _Timer("Read matrix A", out):
        row, col, val, ncol = P.mm_read(file_a)
    n = len(row) - 1
    if ncol != n or len(b) != n or len(x0) != n:
        out.write(f"Sizes do not match: A {n}x{ncol}, b {len(b)}, x {len(x0)}.\n")
        return 1

    if interface == "hip":
        ctx = P.Context(0)
        A = P.CsrMatrix(ctx, row, col, val, ncol=ncol)
        B = P.Vector(ctx, b)
        X = P.Vector(ctx, x0)

        def solve():
            X.upload(x0)
            if solver == "cg":
                return P.pcg(A, B, X, None, maxits=MAX_ITER, atol=ABS_EPS, rtol=REL_EPS, check=False)
            return P.gmres(A, B, X, None, m=BASIS, maxits=MAX_ITER, atol=ABS_EPS, rtol=REL_EPS, check=False)

        def result():
            return X.numpy()
    else:
        holder = {}

        def solve():
            kind = P.SOLVER_CG if solver == "cg" else P.SOLVER_GMRES
            holder["x"], info = P.lmp_solve(row, col, val, b, x0, solver=kind, precond=P.PREC_NONE, m=BASIS,
                                            maxits=MAX_ITER, atol=ABS_EPS, rtol=REL_EPS)
            return info

        def result():
            return holder["x"]

    for _ in range(reps - 1):
        solve()
    with _Timer(f"Solve ({solver}, {interface})", out):
        info = solve()
    out.write(f"iterations: {info['iters']}, state: {info['state']}\n")
    if info["state"] not in (P.SUCCESS_ABS, P.SUCCESS_REL):
        out.write("Error in solver.\n")
        return 1
    x = result()
    xr = P.vec_read(file_xr)
    d = x - xr
    dot = float(d @ d)
    out.write(f"dot(x - x_result) = {dot:g}\n")
    if np.isnan(dot):
        out.write("WARNING: Resulting vector is not a number.\n")
    if abs(dot) > 1e-6:
        out.write("WARNING: Deviation between resulting and expected vector x larger than 1e-6.\n")
    out.write("Done successfully.\n")
    return 0


if __name__ == "__main__":
    sys.exit(run(sys.argv))
