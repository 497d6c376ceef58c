"""Empty and degenerate inputs through the C ABI: zero-length vectors,
matrices without entries, a 0 x 0 system, and empty rows under every SpMV
layout."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LAYOUTS = [(0, 0), (1, 2048), (2, 0), (3, 512), (4, 4), (4, 8), (5, 1), (6, 1), (6, 2)]


def test_zero_length_vectors(pkg, ctx):
    for dt in (np.float64, np.float32):
        x, y = pkg.Vector(ctx, 0, dt), pkg.Vector(ctx, 0, dt)
        assert x.dot(y) == 0.0 and x.nrm2() == 0.0
        y.axpy(x, 2.0)
        y.scaleadd(3.0, x)
        y.scale(0.5)
        y.ewmult(x)
        y.copy_from(x)
        x.zeros()
        assert x.numpy().shape == (0,)


@pytest.mark.parametrize("algo", LAYOUTS)
def test_matrix_without_entries(pkg, ctx, algo):
    n = 130                                    # three slices, the last one partial
    row = np.zeros(n + 1, np.int32)
    A = pkg.CsrMatrix(ctx, row, np.zeros(0, np.int32), np.zeros(0), ncol=n)
    A.set_spmv(*algo)
    x = np.random.default_rng(0).uniform(-1, 1, n)
    Y = pkg.Vector(ctx, np.full(n, 7.0))
    A.spmv(pkg.Vector(ctx, x), Y)
    assert np.array_equal(Y.numpy(), np.zeros(n))
    Y.upload(np.full(n, 7.0))
    A.spmv_add(pkg.Vector(ctx, x), Y)
    assert np.array_equal(Y.numpy(), np.full(n, 7.0))


def test_empty_system(pkg, ctx):
    A = pkg.CsrMatrix(ctx, np.zeros(1, np.int32), np.zeros(0, np.int32), np.zeros(0), ncol=0)
    X, B = pkg.Vector(ctx, 0), pkg.Vector(ctx, 0)
    A.spmv(X, B)
    for info in (pkg.pcg(A, B, X), pkg.gmres(A, B, X, m=5),
                 pkg.jacobi(A, B, X, mode=pkg.JACOBI_SOLVE_DAMPED, w=0.5)):
        assert info["iters"] == 0 and info["state"] == pkg.SUCCESS_ABS
    Al = pkg.CsrMatrix.cast_of(A, np.float32)
    info = pkg.dc(A, Al, B, X)
    assert info["outer_iters"] == 0 and info["state"] == pkg.SUCCESS_ABS
    A.swap_diag_front()


@pytest.mark.parametrize("algo", LAYOUTS)
def test_trailing_and_leading_empty_row_runs(pkg, ctx, oracle, algo):
    """Runs of more than 256 empty rows form row blocks with no entries at the
    start and at the end of the matrix (the last one starts at nnz)."""
    rng = np.random.default_rng(21)
    n = 1200
    lens = np.zeros(n, np.int64)
    lens[300:800] = rng.integers(1, 12, 500)
    row = np.zeros(n + 1, np.int32)
    row[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(n, size=l, replace=False)) for l in lens if l]).astype(np.int32)
    val = rng.uniform(-1, 1, row[-1])
    x = rng.uniform(-1, 1, n)
    A = pkg.CsrMatrix(ctx, row, col, val)
    A.set_spmv(*algo)
    Y = pkg.Vector(ctx, np.full(n, 3.0))
    A.spmv(pkg.Vector(ctx, x), Y)
    ref = oracle.spmv(oracle.Csr(row, col, val), x)
    if algo[0] == 2:        # lane-group tree sums: not in the reference order
        assert np.allclose(Y.numpy(), ref, rtol=1e-13, atol=1e-14)
    else:
        assert np.array_equal(Y.numpy(), ref)
