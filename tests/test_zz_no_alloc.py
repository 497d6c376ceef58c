"""The allocation-free CG (hfh_pcg_solve_noalloc) against hfh_pcg_solve.

Both run the same kernels in the same order; the allocation-free one keeps
its work space in the library's static device array and launches the loop
directly instead of through a graph.  So iterates, histories, states and
solutions are bit-identical.  Two calls in a row, the second on a smaller
system, check that nothing of the first call's contents survives into the
second; the size limit and the preconditioner restriction are checked as
errors."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _both(pkg, ctx, row, col, val, rhs, dtype, prec, maxits=3000, rtol=1e-8):
    out = []
    for noalloc in (False, True):
        A = pkg.CsrMatrix(ctx, row, col, val, dtype=dtype)
        P = pkg.Preconditioner.jacobi(A) if prec else None
        X = pkg.Vector(ctx, len(rhs), dtype)
        info = pkg.pcg(A, pkg.Vector(ctx, rhs, dtype), X, P, maxits=maxits, atol=0.0, rtol=rtol, history=True,
                       noalloc=noalloc)
        out.append((info, X.numpy()))
    return out


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("prec", [True, False])
def test_noalloc_equals_allocating_entry(pkg, ctx, dtype, prec):
    # first a 3D Q1 system, then a smaller 2D one on the same static array
    for kind, m in ((pkg.Q1_3D, 31), (pkg.Q1_2D, 40)):
        row, col, val, rhs = pkg.problem(kind, m)
        (ia, xa), (ib, xb) = _both(pkg, ctx, row, col, val, rhs, dtype, prec)
        print(f"noalloc case dtype={np.dtype(dtype).name} jacobi={prec} n={len(rhs)}: iters {ia['iters']} / "
              f"{ib['iters']}", flush=True)
        assert ia["state"] == pkg.SUCCESS_REL
        assert ib["state"] == ia["state"]
        assert ib["iters"] == ia["iters"]
        assert ib["res"] == ia["res"] and ib["res0"] == ia["res0"]
        assert np.array_equal(ib["hist"], ia["hist"])
        assert np.array_equal(xb, xa)


def test_noalloc_maxits_state(pkg, ctx):
    """A solve stopped by maxits reports the same state and iterate."""
    row, col, val, rhs = pkg.problem(pkg.FV_DIFFUSION, 12, p0=1.0, seed=2)
    res = []
    for noalloc in (False, True):
        A = pkg.CsrMatrix(ctx, row, col, val)
        X = pkg.Vector(ctx, len(rhs))
        info = pkg.pcg(A, pkg.Vector(ctx, rhs), X, pkg.Preconditioner.jacobi(A), maxits=7, atol=0.0, rtol=1e-12,
                       check=False, noalloc=noalloc)
        res.append((info, X.numpy()))
    (ia, xa), (ib, xb) = res
    assert ia["state"] == pkg.FAIL_MAXITS == ib["state"]
    assert ia["iters"] == ib["iters"] == 7
    assert np.array_equal(xa, xb)


def test_noalloc_limits(pkg, ctx):
    """Beyond the static array: HFH_ERR_INVALID before any launch; an
    MC-SGS preconditioner: HFH_ERR_UNSUPPORTED.  The same diagonal system in
    FP32 fits and is solved in one step."""
    n = 2_200_000   # FP64: 4 vectors of 17.6 MB exceed the 64 MiB array
    row = np.arange(n + 1, dtype=np.int32)
    col = np.arange(n, dtype=np.int32)
    val = np.full(n, 2.0)
    rhs = np.ones(n)
    A = pkg.CsrMatrix(ctx, row, col, val)
    X = pkg.Vector(ctx, n)
    ctl = pkg.Control(10, 0.0, 1e-8, 1e6)
    inf = pkg.SolveInfo()
    B = pkg.Vector(ctx, rhs)   # kept alive through the call
    rc = pkg.hip().hfh_pcg_solve_noalloc(A.h, None, B.h, X.h, C.byref(ctl), C.byref(inf), None)
    assert rc == pkg.ERR_INVALID
    assert "static" in pkg.hip().hfh_last_error().decode()
    Af = pkg.CsrMatrix(ctx, row, col, val, dtype=np.float32)
    Xf = pkg.Vector(ctx, n, np.float32)
    info = pkg.pcg(Af, pkg.Vector(ctx, rhs, np.float32), Xf, None, maxits=10, atol=0.0, rtol=1e-6, noalloc=True)
    assert info["iters"] == 1
    assert np.array_equal(Xf.numpy(), np.full(n, 0.5, np.float32))
    # MC-SGS is refused
    row2, col2, val2, rhs2 = pkg.problem(pkg.FV_DIFFUSION, 8, p0=1.0, seed=1)
    nc, color, sizes, perm = pkg.multicoloring(row2, col2)
    prow, pcol, pval = pkg.csr_permute(row2, col2, val2, perm)
    A2 = pkg.CsrMatrix(ctx, prow, pcol, pval)
    P2 = pkg.Preconditioner.mcsgs(A2, sizes)
    with pytest.raises(pkg.HfhError):
        pkg.pcg(A2, pkg.Vector(ctx, rhs2[perm]), pkg.Vector(ctx, len(rhs2)), P2, noalloc=True)
