"""GPU parity of the stationary Jacobi solver / smoother against the oracle.

Every Jacobi step is a fixed sequence of separately rounded operations per row
(no reductions), so iterates agree with the oracle bit for bit in FP64 and
FP32.  The solve modes' residual norm is a reduction, summed in another order
on the device: the checked residuals agree to 1e-12 (FP64) / 1e-5 (FP32)
relative, and the step count at which the control stops is the oracle's.
"""
import numpy as np
import pytest

from conftest import relnorm

pytestmark = pytest.mark.gpu

LAYOUTS = [(4, 4), (4, 8), (6, 1), (7, 0), (5, 1), (5, 2), (5, 1, 1), (5, 2, 1), (1, 2048)]
# SELL, SELL, pairs, LDS-staged, dictionary codes (the last two on one
# workgroup, whose waves walk several slices each), row blocks (CSR step)


def matrices(pkg, case):
    if case == "fv_3d":
        row, col, val, rhs = pkg.problem(pkg.FV_DIFFUSION, 10, p0=1.0, seed=3)
    elif case == "q1_2d":
        row, col, val, rhs = pkg.problem(pkg.Q1_2D, 40)
    else:
        row, col, val, rhs = pkg.problem(pkg.UPWIND_CONVDIFF, 12, p0=1e-1, p1=2.0, seed=4)
    return row, col, val, rhs


def no_diag_rows(rng, n=300, band=None):
    """Diagonally dominant random rows; every fifth row has no diagonal entry
    and row 7 is empty.  band: columns within +-band of the row (at most
    2 band + 1 distinct offsets)."""
    row, cols, vals = [0], [], []
    for i in range(n):
        c = set()
        if i != 7:
            lo, hi = (0, n) if band is None else (max(0, i - band), min(n, i + band + 1))
            c = set((lo + rng.choice(hi - lo, size=min(hi - lo, int(rng.integers(1, 9))), replace=False)).tolist())
            if i % 5 == 0:
                c.discard(i)
            else:
                c.add(i)
        c = sorted(c)
        v = rng.uniform(-0.2, 0.2, len(c))
        v[np.array(c, dtype=np.int64) == i] = 4.0 + rng.uniform()
        cols.extend(c)
        vals.extend(v.tolist())
        row.append(row[-1] + len(c))
    return np.array(row, np.int32), np.array(cols, np.int32), np.array(vals)


@pytest.mark.parametrize("case", ["fv_3d", "q1_2d", "upwind"])
def test_swap_diag_front_bitwise(pkg, ctx, oracle, case):
    row, col, val, _ = matrices(pkg, case)
    A = pkg.CsrMatrix(ctx, row, col, val)
    assert A.is_sorted()
    A.swap_diag_front()
    assert not A.is_sorted()
    r2, c2, v2 = A.download()
    S = oracle.swap_diag_front(oracle.Csr(row, col, val))
    assert np.array_equal(r2, S.row) and np.array_equal(c2, S.col) and np.array_equal(v2, S.val)
    # every row with a diagonal now starts with it
    has = [i for i in range(len(row) - 1) if i in col[row[i]:row[i + 1]]]
    assert all(c2[r2[i]] == i for i in has)


@pytest.mark.parametrize("algo", LAYOUTS)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_spmv_of_swapped_matrix_in_stored_order(pkg, ctx, oracle, algo, dtype):
    row, col, val, _ = matrices(pkg, "q1_2d")
    A = pkg.CsrMatrix(ctx, row, col, val.astype(dtype))
    A.set_spmv(*algo)
    A.swap_diag_front()
    S = oracle.swap_diag_front(oracle.Csr(row, col, val.astype(dtype)))
    x = np.random.default_rng(2).uniform(-1, 1, len(row) - 1).astype(dtype)
    X, Y = pkg.Vector(ctx, x, dtype), pkg.Vector(ctx, len(x), dtype)
    A.spmv(X, Y)
    assert np.array_equal(Y.numpy(), oracle.spmv(S, x))
    A.spmv_nodiag(X, Y)
    assert np.array_equal(Y.numpy(), oracle.spmv_nodiag(S, x))


def test_spmv_nodiag_rows_without_diagonal(pkg, ctx, oracle):
    rng = np.random.default_rng(11)
    row, col, val = no_diag_rows(rng)
    A = pkg.CsrMatrix(ctx, row, col, val)
    x = rng.uniform(-1, 1, len(row) - 1)
    Y = pkg.Vector(ctx, len(x))
    A.spmv_nodiag(pkg.Vector(ctx, x), Y)
    assert np.array_equal(Y.numpy(), oracle.spmv_nodiag(oracle.Csr(row, col, val), x))


@pytest.mark.parametrize("algo", LAYOUTS)
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("mode,w", [(2, 1.0), (3, 0.7)])
def test_smoother_bitwise(pkg, ctx, oracle, algo, dtype, mode, w):
    row, col, val, rhs = matrices(pkg, "fv_3d")
    A = pkg.CsrMatrix(ctx, row, col, val.astype(dtype))
    A.set_spmv(*algo)
    A.swap_diag_front()
    S = oracle.swap_diag_front(oracle.Csr(row, col, val.astype(dtype)))
    x0 = np.random.default_rng(3).uniform(-1, 1, len(rhs)).astype(dtype)
    X = pkg.Vector(ctx, x0, dtype)
    info = pkg.jacobi(A, pkg.Vector(ctx, rhs.astype(dtype), dtype), X, mode=mode, w=w, inner_iter=3, smooth_iter=4)
    xo, oi = oracle.jacobi(S, rhs.astype(dtype), x0, mode=mode, w=w, inner_iter=3, smooth_iter=4)
    assert info["iters"] == oi["iters"] == 12
    assert np.array_equal(X.numpy(), xo)


@pytest.mark.parametrize("case", ["fv_3d", "upwind"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("mode,w,inner", [(0, 1.0, 1), (1, 0.8, 1), (0, 1.0, 4), (1, 0.6, 3)])
def test_solve_matches_oracle(pkg, ctx, oracle, case, dtype, mode, w, inner):
    row, col, val, rhs = matrices(pkg, case)
    A = pkg.CsrMatrix(ctx, row, col, val.astype(dtype))
    A.swap_diag_front()
    S = oracle.swap_diag_front(oracle.Csr(row, col, val.astype(dtype)))
    rtol = 1e-8 if dtype == np.float64 else 1e-4
    X = pkg.Vector(ctx, len(rhs), dtype)
    info = pkg.jacobi(A, pkg.Vector(ctx, rhs.astype(dtype), dtype), X, mode=mode, w=w, inner_iter=inner,
                      maxits=4000, atol=0.0, rtol=rtol, history=True)
    xo, oi = oracle.jacobi(S, rhs.astype(dtype), None, mode=mode, w=w, inner_iter=inner,
                           ctrl=(4000, 0.0, rtol, 1e6))
    assert info["state"] == oi["status"] == pkg.SUCCESS_REL
    assert info["iters"] == oi["iters"]
    assert np.array_equal(X.numpy(), xo)
    tol = 1e-12 if dtype == np.float64 else 1e-5
    assert relnorm(info["hist"], oi["hist"]) < tol


@pytest.mark.parametrize("algo,band", [((0, 0), None), ((7, 0), None), ((5, 1), 100)])
def test_solve_rows_without_diagonal_and_maxits(pkg, ctx, oracle, algo, band):
    """Rows without a diagonal keep inv_diag 0 (their iterate is 0 after a
    step); the control stops on maxits, as the oracle's does."""
    rng = np.random.default_rng(12)
    row, col, val = no_diag_rows(rng, band=band)
    b = rng.uniform(-1, 1, len(row) - 1)
    A = pkg.CsrMatrix(ctx, row, col, val)
    A.set_spmv(*algo)
    A.swap_diag_front()
    assert algo[0] == 0 or A.get_spmv()[0] == algo[0]
    S = oracle.swap_diag_front(oracle.Csr(row, col, val))
    X = pkg.Vector(ctx, len(b))
    info = pkg.jacobi(A, pkg.Vector(ctx, b), X, mode=1, w=0.9, inner_iter=2, maxits=9, atol=0.0, rtol=1e-30,
                      check=False)
    xo, oi = oracle.jacobi(S, b, None, mode=1, w=0.9, inner_iter=2, ctrl=(9, 0.0, 1e-30, 1e6))
    assert info["rc"] == pkg.ERR_NOT_CONVERGED
    assert info["state"] == oi["status"] == pkg.FAIL_MAXITS
    assert info["iters"] == oi["iters"] == 10
    assert np.array_equal(X.numpy(), xo)


def test_swapped_matrix_refused_by_ordered_setups(pkg, ctx):
    row, col, val, _ = matrices(pkg, "fv_3d")
    A = pkg.CsrMatrix(ctx, row, col, val)
    A.swap_diag_front()
    nc, color, sizes, perm = pkg.multicoloring(row, col)
    with pytest.raises(pkg.HfhError):
        pkg.Preconditioner.mcsgs(A, [len(row) - 1])
    with pytest.raises(pkg.HfhError):
        A.ilu0(row, col)
    # Jacobi preconditioner and SpMV still work
    P = pkg.Preconditioner.jacobi(A)
    X, Y = pkg.Vector(ctx, np.ones(len(row) - 1)), pkg.Vector(ctx, len(row) - 1)
    P.apply(X, Y)
    assert np.all(np.isfinite(Y.numpy()))
