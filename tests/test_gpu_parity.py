"""GPU parity of the HIP platform against the CPU oracle, through the C ABI.

Tolerances: SpMV bit-identical for the kernels that sum rows in stored
order (row blocks, wave tiles, SELL forms), norm-wise relative 1e-13 for the
lane-group kernel; element-wise BLAS-1 (Axpy, ScaleAdd, Scale, EWMult)
bit-identical (products and sums rounded separately, as the reference CPU
build does); Dot / Norm2 1e-14 relative (another summation order); Jacobi,
SGS and ILU(0) applies bit-identical; MC-SGS 1e-13; Krylov solvers: iteration
count within +-2 of the oracle and the true residual at the requested
tolerance.
"""
import numpy as np
import pytest

from conftest import relnorm

pytestmark = pytest.mark.gpu

SPMV_TOL = 1e-13


def rand_csr(rng, n, ncol, lens):
    row = np.zeros(n + 1, np.int32)
    row[1:] = np.cumsum(lens)
    col = np.concatenate([np.sort(rng.choice(ncol, size=l, replace=False)) for l in lens]).astype(np.int32) \
        if row[-1] else np.zeros(0, np.int32)
    val = rng.uniform(-1, 1, row[-1])
    return row, col, val


ALGOS = [(1, 2048), (1, 1024), (1, 512), (1, 4096), (3, 512), (3, 1024), (3, 256), (2, 0), (4, 4), (4, 2), (4, 8), (4, 16),
         (5, 1), (5, 2), (6, 1), (6, 2), (7, 0)]
BITWISE = {1, 3, 4, 5, 6, 7}  # row-block, wave-tile and SELL kernels sum rows in the reference order


def spmv_case(pkg, ctx, oracle, row, col, val, ncol=None, dtype=np.float64, seed=0, algo=(1, 2048)):
    ncol = len(row) - 1 if ncol is None else ncol
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, ncol).astype(dtype)
    A = pkg.CsrMatrix(ctx, row, col, val.astype(dtype), ncol=ncol)
    A.set_spmv(*algo)
    X = pkg.Vector(ctx, x, dtype)
    Y = pkg.Vector(ctx, len(row) - 1, dtype)
    A.spmv(X, Y)
    ref = oracle.spmv(oracle.Csr(row, col, val.astype(dtype), ncol), x)
    return Y.numpy(), ref, A, X, Y


def test_spmv_reference_known_answers(pkg, ctx, oracle):
    # test-legacy/lmp_test.cc:80-82 and test-legacy/cuda_test.cc:75-77
    row = np.array([0, 2, 5, 7], np.int32)
    col = np.array([0, 1, 0, 1, 2, 1, 2], np.int32)
    val = np.array([2.0, -1.0, -1.0, 2.0, -1.0, -1.0, 2.0])
    A = pkg.CsrMatrix(ctx, row, col, val)
    Y = pkg.Vector(ctx, 3)
    A.spmv(pkg.Vector(ctx, [1.0, -2.0, 0.0]), Y)
    assert list(Y.numpy()) == [4.0, -5.0, 2.0]
    A.spmv(pkg.Vector(ctx, [1.2, -2.3, 0.7]), Y)
    assert np.allclose(Y.numpy(), [4.7, -6.5, 3.7], atol=1e-12)


@pytest.mark.parametrize("algo", ALGOS)
@pytest.mark.parametrize("case", ["q1_2d_256", "q1_3d_32", "fv_24", "upwind_20"])
def test_spmv_bitwise_on_configs(pkg, ctx, oracle, case, algo):
    if case == "q1_2d_256":
        row, col, val, _ = pkg.problem(pkg.Q1_2D, 255)             # config 1 at full size
    elif case == "q1_3d_32":
        row, col, val, _ = pkg.problem(pkg.Q1_3D, 31)
    elif case == "fv_24":
        row, col, val, _ = pkg.problem(pkg.FV_DIFFUSION, 24, p0=1.0, seed=1)
    else:
        row, col, val, _ = pkg.problem(pkg.UPWIND_CONVDIFF, 20, p0=1e-2, p1=5.0)
    y, ref, *_ = spmv_case(pkg, ctx, oracle, row, col, val, algo=algo)
    y32, ref32, *_ = spmv_case(pkg, ctx, oracle, row, col, val, dtype=np.float32, algo=algo)
    if algo[0] in BITWISE:
        # products rounded then summed in the reference order: bit-identical rows
        assert np.array_equal(y, ref)
        assert np.array_equal(y32, ref32)
    else:
        assert relnorm(y, ref) < SPMV_TOL
        assert relnorm(y32, ref32) < 1e-6


@pytest.mark.parametrize("algo", ALGOS)
def test_spmv_ragged_and_long_rows(pkg, ctx, oracle, algo):
    rng = np.random.default_rng(3)
    n, ncol = 600, 5000
    lens = rng.integers(0, 40, n)
    lens[[0, 17, 300]] = 0                          # empty rows
    lens[[5, 299, 599, 400]] = [1500, 4000, 1025, 300]  # rows longer than a tile
    lens[100:180] = 1                               # many short rows (more than 63 per tile)
    row, col, val = rand_csr(rng, n, ncol, lens)
    y, ref, A, X, Y = spmv_case(pkg, ctx, oracle, row, col, val, ncol=ncol, algo=algo)
    assert relnorm(y, ref) < SPMV_TOL
    if algo[0] in (4, 5, 6, 7):
        assert np.array_equal(y, ref)   # SELL sums every row sequentially, long rows included
    elif algo[0] in BITWISE:
        short = np.diff(row) <= min(algo[1], 256)
        assert np.array_equal(y[short], ref[short])
    # y += A x
    y0 = rng.uniform(-1, 1, n)
    Y.upload(y0)
    A.spmv_add(X, Y)
    assert relnorm(Y.numpy(), oracle.spmv_add(oracle.Csr(row, col, val, ncol), X.numpy(), y0)) < SPMV_TOL


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("words", [1, 2])
def test_spmv_dictionary_columns(pkg, ctx, oracle, dtype, words):
    """Dictionary-coded columns on a ragged, rectangular banded matrix with 20
    distinct offsets (empty rows, rows of 1..20 entries, partial last slice)."""
    rng = np.random.default_rng(7)
    n, ncol = 1000, 1100
    offs = np.sort(rng.choice(np.arange(-300, 400), 20, replace=False))
    rows, cols = [0], []
    for r in range(n):
        k = rng.integers(0, 21) if r % 97 else 0
        c = np.unique(np.clip(r + rng.choice(offs, k, replace=False), 0, ncol - 1)) if k else np.array([], int)
        c = c[(c - r >= offs[0]) & np.isin(c - r, offs)]
        cols.extend(c.tolist())
        rows.append(len(cols))
    row, col = np.array(rows, np.int32), np.array(cols, np.int32)
    val = rng.uniform(-1, 1, len(col))
    y, ref, A, X, Y = spmv_case(pkg, ctx, oracle, row, col, val, ncol=ncol, dtype=dtype, algo=(5, words))
    assert np.array_equal(y, ref)
    y0 = rng.uniform(-1, 1, n).astype(dtype)
    Y.upload(y0)
    A.spmv_add(X, Y)
    assert np.array_equal(Y.numpy(), oracle.spmv_add(oracle.Csr(row, col, val.astype(dtype), ncol), X.numpy(), y0))


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("words", [1, 2])
def test_spmv_dictionary_waves_walk_slices(pkg, ctx, oracle, dtype, words):
    """Dictionary-coded columns on one workgroup, whose 4 waves walk 6 slices
    each (FP32: each wave loads its next slice ahead of its current row sums):
    an all-empty slice, slices up to 20 slots wide (several steps), a partial
    last slice; y = A x and y += A x bit for bit."""
    rng = np.random.default_rng(11)
    n, ncol = 1500, 1500
    offs = np.sort(rng.choice(np.arange(-200, 200), 20, replace=False))
    rows, cols = [0], []
    for r in range(n):
        k = 0 if 128 <= r < 192 or r % 97 == 0 else int(rng.integers(0, 21))
        c = np.unique(np.clip(r + rng.choice(offs, k, replace=False), 0, ncol - 1)) if k else np.array([], int)
        c = c[np.isin(c - r, offs)]
        cols.extend(c.tolist())
        rows.append(len(cols))
    row, col = np.array(rows, np.int32), np.array(cols, np.int32)
    assert row[192] == row[128]   # slice 2 holds no entry
    val = rng.uniform(-1, 1, len(col))
    y, ref, A, X, Y = spmv_case(pkg, ctx, oracle, row, col, val, ncol=ncol, dtype=dtype, algo=(5, words, 1))
    assert np.array_equal(y, ref)
    y0 = rng.uniform(-1, 1, n).astype(dtype)
    Y.upload(y0)
    A.spmv_add(X, Y)
    assert np.array_equal(Y.numpy(), oracle.spmv_add(oracle.Csr(row, col, val.astype(dtype), ncol), X.numpy(), y0))


def banded_ragged(rng, n, ncol, halfwidth, maxlen):
    """Rows of 0..maxlen random columns within +-halfwidth of the row (every
    97th row empty), clipped to [0, ncol)."""
    rows, cols = [0], []
    for r in range(n):
        k = int(rng.integers(0, maxlen + 1)) if r % 97 else 0
        lo, hi = max(0, r - halfwidth), min(ncol, r + halfwidth + 1)
        c = np.sort(rng.choice(np.arange(lo, hi), min(k, hi - lo), replace=False)) if k and hi > lo else []
        cols.extend(int(v) for v in c)
        rows.append(len(cols))
    return np.array(rows, np.int32), np.array(cols, np.int32)


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("n,ncol,halfwidth,maxlen", [(1000, 1037, 300, 25), (131, 9, 5, 6), (4099, 4099, 40, 31),
                                                     (70, 2000, 480, 3)])
def test_spmv_lds_staged_x(pkg, ctx, oracle, dtype, n, ncol, halfwidth, maxlen):
    """x staged in LDS per slice: ragged banded rows (empty rows, odd and even
    slice widths, a partial last slice, ncol not a multiple of the 16-element
    chunk, fewer columns than one chunk, windows up to 61 chunks)."""
    rng = np.random.default_rng(n + ncol)
    row, col = banded_ragged(rng, n, ncol, halfwidth, maxlen)
    val = rng.uniform(-1, 1, len(col))
    y, ref, A, X, Y = spmv_case(pkg, ctx, oracle, row, col, val, ncol=ncol, dtype=dtype, algo=(7, 0))
    assert A.get_spmv() == (pkg.SPMV_SELL_LDS, 0)
    assert np.array_equal(y, ref)
    y0 = rng.uniform(-1, 1, n).astype(dtype)
    Y.upload(y0)
    A.spmv_add(X, Y)
    assert np.array_equal(Y.numpy(), oracle.spmv_add(oracle.Csr(row, col, val.astype(dtype), ncol), X.numpy(), y0))


def test_spmv_lds_falls_back_when_a_window_is_too_wide(pkg, ctx, oracle):
    # one slice touching 65 chunks of 16 elements of x (1040 > 1024): the pair
    # layout instead
    chunk, count = 16, 65
    rng = np.random.default_rng(11)
    n, ncol = 64, count * chunk
    lens = np.full(n, 2)
    lens[0] = count
    row = np.zeros(n + 1, np.int32)
    row[1:] = np.cumsum(lens)
    col = np.concatenate([np.arange(count) * chunk] +
                         [np.sort(rng.choice(ncol, 2, replace=False)) for _ in range(n - 1)])
    val = rng.uniform(-1, 1, len(col))
    y, ref, A, *_ = spmv_case(pkg, ctx, oracle, row, col.astype(np.int32), val, ncol=ncol, algo=(7, 0))
    assert A.get_spmv() == (pkg.SPMV_SELL_PAIRS, 1)
    assert np.array_equal(y, ref)


@pytest.mark.parametrize("solver", ["pcg", "gmres"])
def test_spmv_lds_solvers_identical_to_pairs(pkg, ctx, solver):
    """Every SpMV of the LDS-staged layout equals the pair layout's bit for
    bit, so whole solves do too: PCG (the SpMV with its fused dot) and GMRES
    (an odd n puts every other basis vector at an 8-byte offset, which the
    staging reads element by element)."""
    if solver == "pcg":
        row, col, val, rhs = pkg.problem(pkg.Q1_3D, 17)
    else:
        row, col, val, rhs = pkg.problem(pkg.UPWIND_CONVDIFF, 15, p0=1e-2, p1=5.0, seed=2)
    assert len(rhs) % 2 == 1
    out = []
    for algo in ((6, 1), (7, 0)):
        A = pkg.CsrMatrix(ctx, row, col, val)
        A.set_spmv(*algo)
        assert A.get_spmv()[0] == algo[0]
        P = pkg.Preconditioner.jacobi(A)
        X = pkg.Vector(ctx, len(rhs))
        if solver == "pcg":
            info = pkg.pcg(A, pkg.Vector(ctx, rhs), X, P, maxits=2000, atol=0.0, rtol=1e-10)
        else:
            info = pkg.gmres(A, pkg.Vector(ctx, rhs), X, P, m=30, maxits=2000, atol=0.0, rtol=1e-10)
        out.append((info["iters"], X.numpy()))
    for it, x in out[1:]:
        assert it == out[0][0]
        assert np.array_equal(x, out[0][1])


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("algo", [(4, 8), (5, 1), (6, 1), (7, 0), (1, 2048)])
def test_spmv_new_values_every_layout(pkg, ctx, oracle, algo, dtype):
    """New values for the same structure (upload_values) reach each layout's
    copy (values refilled from the CSR arrays; columns, codes and chunk lists
    kept): the SpMV stays bit-identical."""
    row, col, val, _ = pkg.problem(pkg.Q1_3D, 13)
    n = len(row) - 1
    rng = np.random.default_rng(8)
    x = rng.uniform(-1, 1, n).astype(dtype)
    A = pkg.CsrMatrix(ctx, row, col, val.astype(dtype))
    A.set_spmv(*algo)
    assert A.get_spmv()[0] == algo[0]
    X, Y = pkg.Vector(ctx, x, dtype), pkg.Vector(ctx, n, dtype)
    A.spmv(X, Y)
    assert np.array_equal(Y.numpy(), oracle.spmv(oracle.Csr(row, col, val.astype(dtype)), x))
    val2 = (val * rng.uniform(0.5, 1.5, len(val))).astype(dtype)
    A.upload_values(val2)
    A.spmv(X, Y)
    assert np.array_equal(Y.numpy(), oracle.spmv(oracle.Csr(row, col, val2), x))


def test_spmv_layout_choice(pkg, ctx):
    """The automatic SpMV layout and the LDS-window and dictionary fallbacks."""
    row, col, val, _ = pkg.problem(pkg.Q1_3D, 12)
    assert pkg.CsrMatrix(ctx, row, col, val).get_spmv() == (pkg.SPMV_SELL_LDS, 0)        # x windows fit
    assert pkg.CsrMatrix(ctx, row, col, val, dtype=np.float32).get_spmv() == (pkg.SPMV_SELL_LDS, 0)
    row, col, val, _ = pkg.problem(pkg.FV_DIFFUSION, 12)
    A = pkg.CsrMatrix(ctx, row, col, val)
    assert A.get_spmv() == (pkg.SPMV_SELL_DICT, 1)                                      # stencil rows, 7 offsets
    # a cast takes the automatic choice of its own value type, or the requested layout
    assert pkg.CsrMatrix.cast_of(A, np.float32).get_spmv() == (pkg.SPMV_SELL_DICT, 2)
    A.set_spmv(pkg.SPMV_SELL, 4)
    assert pkg.CsrMatrix.cast_of(A, np.float32).get_spmv() == (pkg.SPMV_SELL, 4)
    A = pkg.CsrMatrix(ctx, row, col, val)
    # wide x windows (random columns): pairs / SELL
    rng = np.random.default_rng(4)
    row, col, val = rand_csr(rng, 640, 40000, np.full(640, 20))
    assert pkg.CsrMatrix(ctx, row, col, val, ncol=40000).get_spmv() == (pkg.SPMV_SELL_PAIRS, 1)  # long FP64 rows
    assert pkg.CsrMatrix(ctx, row, col, val, ncol=40000, dtype=np.float32).get_spmv() == (pkg.SPMV_SELL, 4)
    row, col, val = rand_csr(rng, 640, 40000, np.full(640, 6))
    assert pkg.CsrMatrix(ctx, row, col, val, ncol=40000).get_spmv() == (pkg.SPMV_SELL, 8)  # > 256 offsets
    row, col, val, _ = pkg.problem(pkg.FV_DIFFUSION, 12)
    A = pkg.CsrMatrix(ctx, row, col, val)
    A.set_spmv(pkg.SPMV_SELL_DICT, 2)
    assert A.get_spmv() == (pkg.SPMV_SELL_DICT, 2)                                      # 7 distinct offsets
    rng = np.random.default_rng(5)
    row, col, val = rand_csr(rng, 400, 3000, rng.integers(0, 30, 400))
    B = pkg.CsrMatrix(ctx, row, col, val, ncol=3000)
    B.set_spmv(pkg.SPMV_SELL_DICT, 1)
    assert B.get_spmv()[0] == pkg.SPMV_SELL                                             # > 256 offsets
    lens = np.ones(640, int)
    lens[::64] = 200                                                                    # one long row per slice
    row, col, val = rand_csr(rng, 640, 1000, lens)
    assert pkg.CsrMatrix(ctx, row, col, val, ncol=1000).get_spmv()[0] == pkg.SPMV_STREAM  # padding too large


def test_spmv_single_row_and_rect(pkg, ctx, oracle):
    rng = np.random.default_rng(4)
    row, col, val = rand_csr(rng, 1, 7, [5])
    y, ref, *_ = spmv_case(pkg, ctx, oracle, row, col, val, ncol=7)
    assert np.array_equal(y, ref)
    row, col, val = rand_csr(rng, 50, 3, rng.integers(0, 4, 50))
    y, ref, *_ = spmv_case(pkg, ctx, oracle, row, col, val, ncol=3)
    assert np.array_equal(y, ref)


def test_csr_upload_validation(pkg, ctx):
    with pytest.raises(pkg.HfhError) as e:
        pkg.CsrMatrix(ctx, [0, 2], [1, 0], [1.0, 2.0])       # unsorted columns
    assert e.value.code == pkg.ERR_INVALID
    with pytest.raises(pkg.HfhError):
        pkg.CsrMatrix(ctx, [0, 1], [3], [1.0])               # column out of range
    with pytest.raises(pkg.HfhError):
        pkg.CsrMatrix(ctx, [0, 2, 1], [0, 1], [1.0, 2.0])    # decreasing row pointer


@pytest.mark.parametrize("n", [1, 7, 1000, 4099, (1 << 20) + 3])
def test_blas1(pkg, ctx, oracle, n):
    rng = np.random.default_rng(n)
    x = rng.uniform(-1, 1, n)
    y = rng.uniform(-1, 1, n)
    X, Y = pkg.Vector(ctx, x), pkg.Vector(ctx, y)
    assert abs(X.dot(Y) - oracle.dot(x, y)) <= 1e-14 * np.dot(np.abs(x), np.abs(y))
    assert X.nrm2() == pytest.approx(oracle.nrm2(x), rel=1e-14)
    a = 0.37
    # element-wise updates round every product and sum separately, as the
    # reference's CPU loops do: bit-identical
    Y.axpy(X, a)
    assert np.array_equal(Y.numpy(), oracle.axpy(y, x, a))
    y2 = Y.numpy()
    Y.scaleadd(-1.7, X)
    assert np.array_equal(Y.numpy(), oracle.scaleadd(y2, -1.7, x))
    y3 = Y.numpy()
    Y.scale(0.5)
    assert np.array_equal(Y.numpy(), oracle.scale(y3, 0.5))
    y4 = Y.numpy()
    Y.ewmult(X)
    assert np.array_equal(Y.numpy(), oracle.ewmult(y4, x))
    # FP32 vectors
    Xf, Yf = pkg.Vector(ctx, x, np.float32), pkg.Vector(ctx, y, np.float32)
    ref = float(np.dot(x.astype(np.float32).astype(np.float64), y.astype(np.float32).astype(np.float64)))
    assert abs(Xf.dot(Yf) - ref) <= 1e-6 * np.dot(np.abs(x), np.abs(y)) + 1e-30
    xf, yf = x.astype(np.float32), y.astype(np.float32)
    Yf.axpy(Xf, a)
    yf = oracle.axpy(yf, xf, a)
    assert np.array_equal(Yf.numpy(), yf)
    Yf.scaleadd(-1.7, Xf)
    yf = oracle.scaleadd(yf, -1.7, xf)
    assert np.array_equal(Yf.numpy(), yf)
    Yf.scale(0.3)
    yf = oracle.scale(yf, 0.3)
    assert np.array_equal(Yf.numpy(), yf)
    Yf.ewmult(Xf)
    assert np.array_equal(Yf.numpy(), oracle.ewmult(yf, xf))
    # an operand that is also the target
    Yf.axpy(Yf, 0.5)
    assert np.array_equal(Yf.numpy(), oracle.axpy(oracle.ewmult(yf, xf), oracle.ewmult(yf, xf), 0.5))


@pytest.mark.parametrize("n", [0, 5, 1 << 16])
def test_dot_dev(pkg, ctx, n):
    """The device-result Dot stores the same value hfh_dot returns, in stream order."""
    rng = np.random.default_rng(3)
    x, y = rng.uniform(-1, 1, n), rng.uniform(-1, 1, n)
    X, Y = pkg.Vector(ctx, x), pkg.Vector(ctx, y)
    D = pkg.Vector(ctx, np.full(3, 7.0))
    X.dot_dev(Y, D, 1)
    X.dot_dev(X, D, 2)
    d = D.numpy()
    assert d[0] == 7.0
    assert d[1] == (X.dot(Y) if n else 0.0)
    assert d[2] == (X.dot(X) if n else 0.0)


def test_vector_ops(pkg, ctx):
    rng = np.random.default_rng(9)
    x = rng.uniform(-1, 1, 1000)
    X = pkg.Vector(ctx, x)
    assert np.array_equal(X.numpy(), x)
    assert np.array_equal(X.numpy(10, 5), x[10:15])
    idx = np.array([5, 999, 0, 17])
    assert np.array_equal(X.get_values(idx), x[idx])
    X.set_values(idx, [1.0, 2.0, 3.0, 4.0])
    x[idx] = [1.0, 2.0, 3.0, 4.0]
    assert np.array_equal(X.numpy(), x)
    Y = pkg.Vector(ctx, 1000)
    Y.copy_range(100, X, 200, 50)
    ref = np.zeros(1000)
    ref[100:150] = x[200:250]
    assert np.array_equal(Y.numpy(), ref)
    F = pkg.Vector(ctx, 1000, np.float32)
    F.cast_from(X)
    assert np.array_equal(F.numpy(), x.astype(np.float32))
    D = pkg.Vector(ctx, 1000)
    D.cast_from(F)
    assert np.array_equal(D.numpy(), x.astype(np.float32).astype(np.float64))
    perm = rng.permutation(1000)
    D.reorder_from(X, perm)
    assert np.array_equal(D.numpy(), x[perm])
    Y.copy_from(X)
    Y.zeros()
    assert not Y.numpy().any()
    with pytest.raises(pkg.HfhError):
        X.get_values([1000])


# ---------------------------------------------------------------- preconditioners

def fv3(pkg, n=16, seed=1, sigma=1.0):
    return pkg.problem(pkg.FV_DIFFUSION, n, p0=sigma, seed=seed)


def test_jacobi_apply(pkg, ctx, oracle):
    row, col, val, rhs = fv3(pkg)
    A = pkg.CsrMatrix(ctx, row, col, val)
    P = pkg.Preconditioner.jacobi(A)
    Z = pkg.Vector(ctx, len(rhs))
    P.apply(pkg.Vector(ctx, rhs), Z)
    ref = rhs * oracle.invdiag(oracle.Csr(row, col, val))
    assert np.array_equal(Z.numpy(), ref)


def test_jacobi_zero_diagonal(pkg, ctx):
    A = pkg.CsrMatrix(ctx, [0, 1, 2], [0, 1], [1.0, 0.0])
    with pytest.raises(pkg.HfhError) as e:
        pkg.Preconditioner.jacobi(A)
    assert e.value.code == pkg.ERR_ZERO_PIVOT


def random_spd_pattern(rng, n, m):
    """Symmetric random pattern (m random pairs), diagonally dominant."""
    import scipy.sparse as sp
    i, j = rng.integers(0, n, m), rng.integers(0, n, m)
    M = sp.coo_matrix((rng.uniform(-1, 0, m), (i, j)), shape=(n, n)).tocsr()
    M = M + M.T
    M.setdiag(0)
    M.eliminate_zeros()
    M = (M + sp.diags(np.asarray(abs(M).sum(axis=1)).ravel() + 1.0)).tocsr()
    M.sort_indices()
    return M.indptr.astype(np.int32), M.indices.astype(np.int32), M.data.copy(), rng.uniform(-1, 1, n)


@pytest.mark.parametrize("omega", [1.0, 1.3])
@pytest.mark.parametrize("kind", ["fv3", "q1_3d", "q1_2d", "random"])
def test_mcsgs_apply(pkg, ctx, oracle, omega, kind):
    """Stencil matrices have few distinct offsets col - row, so their sweep
    copies hold one-byte offset codes; the random pattern has hundreds and
    keeps 4-byte columns."""
    if kind == "fv3":
        row, col, val, rhs = fv3(pkg)
    elif kind == "q1_3d":
        row, col, val, rhs = pkg.problem(pkg.Q1_3D, 15)
    elif kind == "q1_2d":
        row, col, val, rhs = pkg.problem(pkg.Q1_2D, 40, 33)
    else:
        row, col, val, rhs = random_spd_pattern(np.random.default_rng(5), 2000, 6000)
        r_of = np.repeat(np.arange(len(row) - 1), np.diff(row))
        assert len(np.unique(col - r_of)) > 256
    nc, color, sizes, perm = pkg.multicoloring(row, col)
    prow, pcol, pval = pkg.csr_permute(row, col, val, perm)
    A = pkg.CsrMatrix(ctx, prow, pcol, pval)
    P = pkg.Preconditioner.mcsgs(A, sizes, omega)
    r = rhs[perm]
    Z = pkg.Vector(ctx, len(r))
    P.apply(pkg.Vector(ctx, r), Z)
    ref = oracle.mcsgs(oracle.Csr(prow, pcol, pval), sizes, r, omega=omega)
    assert relnorm(Z.numpy(), ref) < 1e-13


def test_mcsgs_fp32_waves_walk_slices(pkg, ctx, oracle):
    """FP32 multicolour SGS colour steps run on 8 workgroups per CU; at FV
    128^3 (16,384 slices per colour) their waves walk several slices each,
    loading the next one ahead of the current row sums.  The apply follows
    the oracle's FP32 apply; 20 CG + MC-SGS steps (the steps that store r.z
    partials) follow the oracle's FP64 solve of the same FP32-rounded system
    (the oracle's FP32 CG sums its 2-million-term dot products sequentially,
    the device in FP64 partials)."""
    n_cu = ctx.info()["num_cu"]
    row, col, val, rhs = pkg.problem(pkg.FV_DIFFUSION, 128, p0=1.0, seed=2)
    nc, color, sizes, perm = pkg.multicoloring(row, col)
    assert nc == 2 and min(sizes) // 64 > 8 * n_cu * 4   # more slices than waves
    prow, pcol, pval = pkg.csr_permute(row, col, val, perm)
    pval, b = pval.astype(np.float32), rhs[perm].astype(np.float32)
    A = pkg.CsrMatrix(ctx, prow, pcol, pval)
    P = pkg.Preconditioner.mcsgs(A, sizes)
    OA = oracle.Csr(prow, pcol, pval)
    Z = pkg.Vector(ctx, len(b), np.float32)
    P.apply(pkg.Vector(ctx, b, np.float32), Z)
    ref = oracle.mcsgs(OA, sizes, b)
    assert relnorm(Z.numpy(), ref) < 1e-5
    X = pkg.Vector(ctx, len(b), np.float32)
    info = pkg.pcg(A, pkg.Vector(ctx, b, np.float32), X, P, maxits=20, atol=0.0, rtol=1e-30, history=True,
                   check=False)
    OA64 = oracle.Csr(prow, pcol, pval.astype(np.float64))
    xo, oi = oracle.cg(OA64, b.astype(np.float64), P=oracle.Prec(oracle.PREC_MCSGS, OA64, sizes=sizes),
                       ctrl=(20, 0.0, 1e-30, 1e6))
    assert info["iters"] == oi["iters"] == 20
    assert np.allclose(info["hist"], oi["hist"], rtol=1e-3)
    assert relnorm(X.numpy(), xo) < 1e-3


def test_mcsgs_fp32_empty_slices_walked(pkg, ctx, oracle):
    """A chain of 2.2 million rows (two colours, about 17,000 slices each, so
    the FP32 colour steps' waves walk several slices) with 400 isolated rows:
    first fit puts them in colour 0 as one run, so colour 0's backward step
    meets whole slices without entries, after which its waves still load
    their next slice ahead.  Apply and 12 CG steps as in the FV case."""
    n, lo, hi = 2_200_000, 1000, 1400
    link = np.ones(n - 1, bool)                    # link i -- i+1
    link[lo - 1:hi] = False                        # rows lo..hi-1 touch nothing
    iso = np.zeros(n, bool)
    iso[lo:hi] = True
    up = np.append(link, False)                    # row i has i+1
    dn = np.insert(link, 0, False)                 # row i has i-1
    cnt = 1 + up.astype(np.int32) + dn.astype(np.int32)
    row = np.zeros(n + 1, np.int32)
    row[1:] = np.cumsum(cnt)
    col = np.empty(row[-1], np.int32)
    val = np.empty(row[-1])
    i = np.arange(n)
    pos = row[:-1].copy()
    col[pos[dn]] = i[dn] - 1
    val[pos[dn]] = -1.0
    pos[dn] += 1
    col[pos] = i
    val[pos] = np.where(iso, 1.0, 2.001)
    pos += 1
    col[pos[up]] = i[up] + 1
    val[pos[up]] = -1.0
    nc, color, sizes, perm = pkg.multicoloring(row, col)
    # colours are numbered from 1, as the reference's Multicoloring numbers them
    assert nc == 2 and np.all(color[lo:hi] == 1) and min(sizes) // 64 > 8 * ctx.info()["num_cu"] * 4
    prow, pcol, pval = pkg.csr_permute(row, col, val, perm)
    pval = pval.astype(np.float32)
    b = np.random.default_rng(4).uniform(-1, 1, n).astype(np.float32)
    A = pkg.CsrMatrix(ctx, prow, pcol, pval)
    P = pkg.Preconditioner.mcsgs(A, sizes)
    OA = oracle.Csr(prow, pcol, pval)
    Z = pkg.Vector(ctx, n, np.float32)
    P.apply(pkg.Vector(ctx, b, np.float32), Z)
    assert relnorm(Z.numpy(), oracle.mcsgs(OA, sizes, b)) < 1e-5
    X = pkg.Vector(ctx, n, np.float32)
    info = pkg.pcg(A, pkg.Vector(ctx, b, np.float32), X, P, maxits=12, atol=0.0, rtol=1e-30, history=True,
                   check=False)
    OA64 = oracle.Csr(prow, pcol, pval.astype(np.float64))
    xo, oi = oracle.cg(OA64, b.astype(np.float64), P=oracle.Prec(oracle.PREC_MCSGS, OA64, sizes=sizes),
                       ctrl=(12, 0.0, 1e-30, 1e6))
    assert info["iters"] == oi["iters"] == 12
    assert np.allclose(info["hist"], oi["hist"], rtol=1e-3)
    assert relnorm(X.numpy(), xo) < 1e-3


@pytest.mark.parametrize("case", ["fv12", "upwind20"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_sgs_natural_apply(pkg, ctx, oracle, dtype, case):
    """Level-scheduled SGS: every product and difference rounded separately in
    stored order, as the CPU loop: bit-identical; also after new values, by
    level slices (12^3, below the blocking size) and by geometric blocks
    (20^3)."""
    if case == "fv12":
        row, col, val, rhs = fv3(pkg, 12)
    else:
        row, col, val, rhs = pkg.problem(pkg.UPWIND_CONVDIFF, 20, p0=1e-2, p1=5.0, seed=8)
    A = pkg.CsrMatrix(ctx, row, col, val.astype(dtype))
    P = pkg.Preconditioner.sgs(A, row, col)
    assert (P.sweep_blocks()[0] > 0) == (case == "upwind20")
    Z = pkg.Vector(ctx, len(rhs), dtype)
    P.apply(pkg.Vector(ctx, rhs.astype(dtype), dtype), Z)
    assert np.array_equal(Z.numpy(), oracle.sgs(oracle.Csr(row, col, val.astype(dtype)), rhs.astype(dtype)))
    # new values for the same structure reach the preconditioner (it reads the
    # operator's current values, as the reference's SGS does)
    val2 = val * np.random.default_rng(3).uniform(0.5, 1.5, len(val))
    A.upload_values(val2.astype(dtype))
    P.apply(pkg.Vector(ctx, rhs.astype(dtype), dtype), Z)
    assert np.array_equal(Z.numpy(), oracle.sgs(oracle.Csr(row, col, val2.astype(dtype)), rhs.astype(dtype)))


@pytest.mark.parametrize("case", ["upwind", "fv3"])
def test_ilu0_apply(pkg, ctx, oracle, case):
    if case == "upwind":
        row, col, val, rhs = pkg.problem(pkg.UPWIND_CONVDIFF, 16, p0=1e-2, p1=5.0, seed=4)
    else:
        row, col, val, rhs = fv3(pkg, 14)
    lu = pkg.ilu0_factor(row, col, val)
    LU = pkg.CsrMatrix(ctx, row, col, lu)
    P = pkg.Preconditioner.ilu0(LU, row, col)
    assert (P.sweep_blocks()[0] > 0) == (case == "upwind")   # 16^3: geometric blocks; 14^3: slices
    Z = pkg.Vector(ctx, len(rhs))
    P.apply(pkg.Vector(ctx, rhs), Z)
    assert np.array_equal(Z.numpy(), oracle.ilu_solve(oracle.Csr(row, col, lu), rhs))
    # the factors recomputed in place on the device reach the preconditioner
    val2 = val * np.random.default_rng(5).uniform(0.9, 1.1, len(val))
    LU.upload_values(val2)
    assert LU.ilu0(row, col) == 0
    P.apply(pkg.Vector(ctx, rhs), Z)
    assert np.array_equal(Z.numpy(), oracle.ilu_solve(oracle.Csr(row, col, pkg.ilu0_factor(row, col, val2)), rhs))


@pytest.mark.parametrize("case", ["upwind40", "fv14"])
def test_ilu0_factor_on_device_bitwise(pkg, ctx, oracle, case):
    """Level-scheduled ILU(0) factorisation on the device: the same factors as
    the sequential host loop, bit for bit."""
    if case == "upwind40":
        row, col, val, _ = pkg.problem(pkg.UPWIND_CONVDIFF, 40, p0=1e-2, p1=5.0, seed=6)
    else:
        row, col, val, _ = pkg.problem(pkg.FV_DIFFUSION, 14, p0=1.0, seed=2)
    A = pkg.CsrMatrix(ctx, row, col, val)
    assert A.ilu0(row, col) == 0
    r2, c2, lu_dev = A.download()
    lu_host = pkg.ilu0_factor(row, col, val)
    assert np.array_equal(lu_dev, lu_host)
    LU, miss = oracle.ilu0(oracle.Csr(row, col, val))
    assert miss == 0 and np.array_equal(lu_dev, LU.val)
    # the SpMV copy follows the new values
    x = np.random.default_rng(1).uniform(-1, 1, len(row) - 1)
    Y = pkg.Vector(ctx, len(x))
    A.spmv(pkg.Vector(ctx, x), Y)
    assert np.array_equal(Y.numpy(), oracle.spmv(oracle.Csr(row, col, lu_host), x))


@pytest.mark.parametrize("prec", ["ilu0", "sgs"])
def test_level_sweeps_merged_and_single(pkg, ctx, oracle, prec):
    """The one-launch-per-level schedule: levels of at most 1024 rows run
    merged in one workgroup, larger levels one launch each; a 40^3 grid has
    both (level sizes 1 .. 1141)."""
    row, col, val, rhs = pkg.problem(pkg.UPWIND_CONVDIFF, 40, p0=1e-2, p1=5.0, seed=6)
    lo, _ = pkg.levels(row, col)
    sizes = np.diff(lo)
    assert sizes.max() > 1024 and (sizes <= 1024).sum() > 2
    if prec == "ilu0":
        lu = pkg.ilu0_factor(row, col, val)
        P = pkg.Preconditioner.ilu0(pkg.CsrMatrix(ctx, row, col, lu), row, col)
        ref = oracle.ilu_solve(oracle.Csr(row, col, lu), rhs)
    else:
        P = pkg.Preconditioner.sgs(pkg.CsrMatrix(ctx, row, col, val), row, col)
        ref = oracle.sgs(oracle.Csr(row, col, val), rhs)
    P.set_sweep(pkg.SWEEP_LEVELS)
    Z = pkg.Vector(ctx, len(rhs))
    P.apply(pkg.Vector(ctx, rhs), Z)
    assert np.array_equal(Z.numpy(), ref)


def _ragged_lower_heavy(n, seed):
    """A nonsymmetric pattern with rows of 1 .. 40 entries, long chains and
    wide levels, diagonally dominant (ILU(0) and SGS well defined)."""
    rng = np.random.default_rng(seed)
    rows = []
    for i in range(n):
        k = int(rng.integers(0, 40))
        cols = set(rng.integers(0, n, k).tolist()) | {i}
        if i > 0 and rng.random() < 0.3:
            cols.add(i - 1)   # chains
        rows.append(sorted(cols))
    row = np.zeros(n + 1, np.int32)
    row[1:] = np.cumsum([len(c) for c in rows])
    col = np.concatenate([np.array(c, np.int32) for c in rows])
    val = rng.uniform(-1, 1, len(col))
    for i in range(n):
        for j in range(row[i], row[i + 1]):
            if col[j] == i:
                val[j] = 50.0 + rng.random()
    return row, col, val


def _grid_stencil(nx, ny, nz, lower, seed):
    """A natural-order nx x ny x nz grid matrix whose rows couple to the
    neighbours at the offsets `lower` (dx, dy, dz) and their mirrors,
    nonsymmetric values, strictly diagonally dominant."""
    rng = np.random.default_rng(seed)
    offs = sorted({o for d in lower for o in (d, tuple(-v for v in d))} | {(0, 0, 0)},
                  key=lambda d: (d[2], d[1], d[0]))
    n = nx * ny * nz
    idx = np.arange(n)
    x, y, z = idx % nx, (idx // nx) % ny, idx // (nx * ny)
    rows, cols = [], []
    for dx, dy, dz in offs:   # ascending column order within each row
        ok = (x + dx >= 0) & (x + dx < nx) & (y + dy >= 0) & (y + dy < ny) & (z + dz >= 0) & (z + dz < nz)
        rows.append(idx[ok])
        cols.append(idx[ok] + dx + nx * dy + nx * ny * dz)
    r, c = np.concatenate(rows), np.concatenate(cols)
    order = np.lexsort((c, r))
    r, c = r[order], c[order]
    val = rng.uniform(-1, 1, len(c))
    val[r == c] = 2.0 * len(offs) + rng.random(n)
    row = np.zeros(n + 1, np.int32)
    np.add.at(row, r + 1, 1)
    return np.cumsum(row).astype(np.int32), c.astype(np.int32), val


@pytest.mark.parametrize("prec", ["ilu0", "sgs"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("case", ["upwind40", "upwind37", "q1_2d", "q1_3d", "edge11", "ragged"])
def test_sweep_schedules_bitwise(pkg, ctx, oracle, prec, dtype, case):
    """The sync-free sweeps (the default: geometric blocks whose own
    dependencies resolve between lanes on a structured grid, else slices in
    level order; waves poll the results they need from other blocks or
    slices) give the level launches' results bit for bit, for any number of
    polling waves, and match the oracle's sequential loops bit for bit.
    7-point (40^3, and 37^3 whose blocks are cut at the far faces), 9-point
    2D and 27-point 3D grids take the blocked form; the ragged pattern has no
    grid and takes the slices.  The 11-point 3D stencil (7-point plus the
    neighbours at (+-1, -1, 0) and (+-1, 0, -1)) has 5 lower entries per row
    (the K = 8 kernel) and needs the sheared blocks."""
    if case == "upwind40":
        row, col, val, rhs = pkg.problem(pkg.UPWIND_CONVDIFF, 40, p0=1e-2, p1=5.0, seed=6)
    elif case == "upwind37":
        row, col, val, rhs = pkg.problem(pkg.UPWIND_CONVDIFF, 37, p0=1e-2, p1=5.0, seed=7)
    elif case == "q1_2d":
        row, col, val, rhs = pkg.problem(pkg.Q1_2D, 120, seed=8)
    elif case == "q1_3d":
        row, col, val, rhs = pkg.problem(pkg.Q1_3D, 24, seed=9)
    elif case == "edge11":
        row, col, val = _grid_stencil(24, 20, 18, [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 0), (-1, 1, 0),
                                                   (1, 0, 1), (-1, 0, 1)], 13)
        rhs = np.random.default_rng(14).uniform(-1, 1, len(row) - 1)
    else:
        row, col, val = _ragged_lower_heavy(6000, 11)
        rhs = np.random.default_rng(12).uniform(-1, 1, len(row) - 1)
    if prec == "ilu0":
        lu = pkg.ilu0_factor(row, col, val)
        A = pkg.CsrMatrix(ctx, row, col, lu.astype(dtype))
        P = pkg.Preconditioner.ilu0(A, row, col)
        ref = oracle.ilu_solve(oracle.Csr(row, col, lu.astype(dtype)), rhs.astype(dtype))
    else:
        A = pkg.CsrMatrix(ctx, row, col, val.astype(dtype))
        P = pkg.Preconditioner.sgs(A, row, col)
        ref = oracle.sgs(oracle.Csr(row, col, val.astype(dtype)), rhs.astype(dtype))
    fb, fl, bb, bl = P.sweep_blocks()
    lo, up = pkg.levels(row, col)[0], pkg.levels(row, col, upper=True)[0]
    print(f"sweep case {case}: n {len(rhs)} row levels {len(lo) - 1}/{len(up) - 1} "
          f"blocks {fb}/{bb} block levels {fl}/{bl}", flush=True)
    if case == "ragged":
        assert (fb, bb) == (0, 0)
    else:
        assert fb > 0 and bb > 0 and 0 < 2 * fl <= len(lo) - 1 and 0 < 2 * bl <= len(up) - 1
    Rv = pkg.Vector(ctx, rhs.astype(dtype), dtype)
    results = {}
    for mode, waves in ((pkg.SWEEP_LEVELS, 0), (pkg.SWEEP_SYNCFREE, 0), (pkg.SWEEP_SYNCFREE, 1),
                        (pkg.SWEEP_SYNCFREE, 4), (pkg.SWEEP_SYNCFREE, 32), (pkg.SWEEP_SLICES, 0),
                        (pkg.SWEEP_SLICES, 4)):
        P.set_sweep(mode, waves)
        Z = pkg.Vector(ctx, len(rhs), dtype)
        for _ in range(2):   # a second apply reuses the counters and buffers
            P.apply(Rv, Z)
        P.sweep_status()
        results[(mode, waves)] = Z.numpy()
        print(f"sweep case {case} {prec} {np.dtype(dtype).name} mode {mode} waves/CU {waves}: ok", flush=True)
    for k, z in results.items():
        assert np.array_equal(z, ref), k


def test_gmres_ilu0_sweep_schedules_same_iterates(pkg, ctx):
    """GMRES + ILU(0) with either schedule: the same iterates, bit for bit."""
    row, col, val, rhs = pkg.problem(pkg.UPWIND_CONVDIFF, 24, p0=1e-2, p1=5.0, seed=9)
    lu = pkg.ilu0_factor(row, col, val)
    A = pkg.CsrMatrix(ctx, row, col, val)
    P = pkg.Preconditioner.ilu0(pkg.CsrMatrix(ctx, row, col, lu), row, col)
    out = []
    for mode in (pkg.SWEEP_LEVELS, pkg.SWEEP_SYNCFREE, pkg.SWEEP_SLICES):
        P.set_sweep(mode)
        X = pkg.Vector(ctx, len(rhs))
        info = pkg.gmres(A, pkg.Vector(ctx, rhs), X, P, m=30, maxits=1000, atol=0.0, rtol=1e-8, history=True)
        out.append((info, X.numpy()))
    for o in out[1:]:
        assert out[0][0]["iters"] == o[0]["iters"]
        assert np.array_equal(out[0][0]["hist"], o[0]["hist"])
        assert np.array_equal(out[0][1], o[1])


# ---------------------------------------------------------------- Krylov drivers

def true_res(row, col, val, x, b):
    import scipy.sparse as sp
    return np.linalg.norm(b - sp.csr_matrix((val, col, row)) @ x) / np.linalg.norm(b)


@pytest.mark.parametrize("case,prec", [("q1_2d", "jacobi"), ("fv3", "jacobi"), ("fv3", "none"),
                                       ("q1_3d", "jacobi")])
def test_pcg_matches_oracle(pkg, ctx, oracle, case, prec):
    if case == "q1_2d":
        row, col, val, rhs = pkg.problem(pkg.Q1_2D, 63)
    elif case == "q1_3d":
        row, col, val, rhs = pkg.problem(pkg.Q1_3D, 23)
    else:
        row, col, val, rhs = fv3(pkg, 20)
    A = pkg.CsrMatrix(ctx, row, col, val)
    P = pkg.Preconditioner.jacobi(A) if prec == "jacobi" else None
    X = pkg.Vector(ctx, len(rhs))
    info = pkg.pcg(A, pkg.Vector(ctx, rhs), X, P, maxits=3000, atol=0.0, rtol=1e-8, history=True)
    OA = oracle.Csr(row, col, val)
    xo, oi = oracle.cg(OA, rhs, P=oracle.Prec.jacobi(OA) if prec == "jacobi" else None,
                       ctrl=(3000, 0.0, 1e-8, 1e6))
    assert info["state"] == pkg.SUCCESS_REL and oi["status"] == oracle.SUCCESS_REL
    assert abs(info["iters"] - oi["iters"]) <= 2
    k = min(20, len(oi["hist"]), len(info["hist"]))
    assert np.allclose(info["hist"][:k], oi["hist"][:k], rtol=1e-9)
    x = X.numpy()
    assert true_res(row, col, val, x, rhs) < 2e-8
    assert relnorm(x, xo) < 1e-6


def test_pcg_large_grid_matches_oracle(pkg, ctx, oracle):
    """3D Q1, 89^3 interior nodes: the SpMV's 2,760 partials times the update
    kernel's 689 workgroups exceed what the update folds in, so the pAp
    finisher runs (small systems fold it into the update kernel)."""
    row, col, val, rhs = pkg.problem(pkg.Q1_3D, 89)
    A = pkg.CsrMatrix(ctx, row, col, val)
    P = pkg.Preconditioner.jacobi(A)
    X = pkg.Vector(ctx, len(rhs))
    info = pkg.pcg(A, pkg.Vector(ctx, rhs), X, P, maxits=2000, atol=0.0, rtol=1e-8, history=True)
    OA = oracle.Csr(row, col, val)
    xo, oi = oracle.cg(OA, rhs, P=oracle.Prec.jacobi(OA), ctrl=(2000, 0.0, 1e-8, 1e6))
    assert info["state"] == pkg.SUCCESS_REL and oi["status"] == oracle.SUCCESS_REL
    assert abs(info["iters"] - oi["iters"]) <= 2
    k = min(20, len(oi["hist"]), len(info["hist"]))
    assert np.allclose(info["hist"][:k], oi["hist"][:k], rtol=1e-9)
    assert true_res(row, col, val, X.numpy(), rhs) < 2e-8


def test_pcg_mcsgs_matches_oracle(pkg, ctx, oracle):
    row, col, val, rhs = fv3(pkg, 18, sigma=1.0, seed=3)
    nc, color, sizes, perm = pkg.multicoloring(row, col)
    prow, pcol, pval = pkg.csr_permute(row, col, val, perm)
    b = rhs[perm]
    A = pkg.CsrMatrix(ctx, prow, pcol, pval)
    P = pkg.Preconditioner.mcsgs(A, sizes)
    X = pkg.Vector(ctx, len(b))
    info = pkg.pcg(A, pkg.Vector(ctx, b), X, P, maxits=2000, atol=0.0, rtol=1e-10)
    OA = oracle.Csr(prow, pcol, pval)
    xo, oi = oracle.cg(OA, b, P=oracle.Prec(oracle.PREC_MCSGS, OA, sizes=sizes), ctrl=(2000, 0.0, 1e-10, 1e6))
    assert abs(info["iters"] - oi["iters"]) <= 2
    assert true_res(prow, pcol, pval, X.numpy(), b) < 3e-10
    assert relnorm(X.numpy(), xo) < 1e-8


@pytest.mark.parametrize("case,ncol", [("q1_2d", 4), ("q1_3d", 8)])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_pcg_mcsgs_many_colours_matches_oracle(pkg, ctx, oracle, case, ncol, dtype):
    """CG + multicolour SGS with 4 and 8 colours: r.z comes from partials of
    the last colour's forward sweep and of every backward sweep."""
    if case == "q1_2d":
        row, col, val, rhs = pkg.problem(pkg.Q1_2D, 200)
    else:
        row, col, val, rhs = pkg.problem(pkg.Q1_3D, 40)
    nc, color, sizes, perm = pkg.multicoloring(row, col)
    assert nc == ncol
    prow, pcol, pval = pkg.csr_permute(row, col, val, perm)
    pval, b = pval.astype(dtype), rhs[perm].astype(dtype)
    A = pkg.CsrMatrix(ctx, prow, pcol, pval)
    P = pkg.Preconditioner.mcsgs(A, sizes)
    X = pkg.Vector(ctx, len(b), dtype)
    rtol = 1e-10 if dtype == np.float64 else 1e-5
    info = pkg.pcg(A, pkg.Vector(ctx, b, dtype), X, P, maxits=2000, atol=0.0, rtol=rtol, history=True)
    OA = oracle.Csr(prow, pcol, pval)
    xo, oi = oracle.cg(OA, b, P=oracle.Prec(oracle.PREC_MCSGS, OA, sizes=sizes), ctrl=(2000, 0.0, rtol, 1e6))
    assert info["state"] == oi["status"] == pkg.SUCCESS_REL
    assert abs(info["iters"] - oi["iters"]) <= 2
    k = min(10, len(oi["hist"]), len(info["hist"]))
    assert np.allclose(info["hist"][:k], oi["hist"][:k], rtol=1e-9 if dtype == np.float64 else 1e-4)
    assert relnorm(X.numpy(), xo) < (1e-8 if dtype == np.float64 else 1e-3)


@pytest.mark.parametrize("flexible", [False, True])
def test_gmres_ilu0_matches_oracle(pkg, ctx, oracle, flexible):
    row, col, val, rhs = pkg.problem(pkg.UPWIND_CONVDIFF, 16, p0=1e-2, p1=5.0, seed=7)
    lu = pkg.ilu0_factor(row, col, val)
    A = pkg.CsrMatrix(ctx, row, col, val)
    LU = pkg.CsrMatrix(ctx, row, col, lu)
    P = pkg.Preconditioner.ilu0(LU, row, col)
    X = pkg.Vector(ctx, len(rhs))
    info = pkg.gmres(A, pkg.Vector(ctx, rhs), X, P, m=30, flexible=flexible, maxits=1000, atol=0.0, rtol=1e-8,
                     history=True)
    OA = oracle.Csr(row, col, val)
    fn = oracle.fgmres if flexible else oracle.gmres
    xo, oi = fn(OA, rhs, P=oracle.Prec(oracle.PREC_ILU0, oracle.Csr(row, col, lu)), m=30,
                ctrl=(1000, 0.0, 1e-8, 1e6))
    assert abs(info["iters"] - oi["iters"]) <= 2
    k = min(10, len(oi["hist"]), len(info["hist"]))
    assert np.allclose(info["hist"][:k], oi["hist"][:k], rtol=1e-8)
    assert true_res(row, col, val, X.numpy(), rhs) < 2e-8


def test_gmres_restarts(pkg, ctx, oracle):
    row, col, val, rhs = pkg.problem(pkg.UPWIND_CONVDIFF, 12, p0=1e-2, p1=5.0, seed=8)
    A = pkg.CsrMatrix(ctx, row, col, val)
    P = pkg.Preconditioner.jacobi(A)
    X = pkg.Vector(ctx, len(rhs))
    info = pkg.gmres(A, pkg.Vector(ctx, rhs), X, P, m=5, maxits=3000, atol=0.0, rtol=1e-8)
    OA = oracle.Csr(row, col, val)
    xo, oi = oracle.gmres(OA, rhs, P=oracle.Prec.jacobi(OA), m=5, ctrl=(3000, 0.0, 1e-8, 1e6))
    assert info["iters"] > 5 and abs(info["iters"] - oi["iters"]) <= 2
    assert true_res(row, col, val, X.numpy(), rhs) < 3e-8


@pytest.mark.parametrize("maxits", [7, 10, 11])
def test_gmres_stops_at_maxits_mid_cycle(pkg, ctx, oracle, maxits):
    # The device runs Arnoldi step j+1 while the host handles step j; a solve
    # cut by maxits inside a cycle (7, 11) or at its end (10) must still stop
    # at that step with the iterates of the sequential loop: the history is a
    # prefix of an uncut solve's, bit for bit, and x matches the oracle's
    row, col, val, rhs = pkg.problem(pkg.UPWIND_CONVDIFF, 12, p0=1e-2, p1=5.0, seed=8)
    A = pkg.CsrMatrix(ctx, row, col, val)
    P = pkg.Preconditioner.jacobi(A)
    B = pkg.Vector(ctx, rhs)
    print(f"gmres maxits case {maxits}", flush=True)
    X = pkg.Vector(ctx, len(rhs))
    info = pkg.gmres(A, B, X, P, m=5, maxits=maxits, atol=0.0, rtol=1e-14, history=True, check=False)
    assert info["iters"] == maxits and info["state"] == pkg.FAIL_MAXITS
    Xl = pkg.Vector(ctx, len(rhs))
    full = pkg.gmres(A, B, Xl, P, m=5, maxits=3000, atol=0.0, rtol=1e-8, history=True)
    assert full["iters"] > maxits
    assert np.array_equal(info["hist"][: maxits + 1], full["hist"][: maxits + 1])
    OA = oracle.Csr(row, col, val)
    xo, oi = oracle.gmres(OA, rhs, P=oracle.Prec.jacobi(OA), m=5, ctrl=(maxits, 0.0, 1e-14, 1e6))
    assert oi["iters"] == maxits
    assert relnorm(X.numpy(), xo) < 1e-10


@pytest.mark.parametrize("dtype", [np.float64, np.float32])
@pytest.mark.parametrize("where", ["two_per_thread", "at_limit", "above_limit"])
def test_gmres_mgs_cooperative_and_multi_launch(pkg, ctx, oracle, where, dtype):
    """The MGS sweep of a column runs as one cooperative launch (w kept in
    registers: 2 elements per thread at 2 x 1024 x CUs rows, 8 at the limit
    of 8 x 1024 x CUs) or, above the limit (and at most 1024 x CUs rows), as
    one launch per step: each follows the oracle's GMRES for 10 iterations on
    a nonsymmetric tridiagonal system of that size, in FP64 and in FP32.  The
    FP32 solve is checked against the oracle in FP64 on the same FP32-rounded
    matrix and right-hand side: the oracle's own FP32 run sums its dot products
    (half a million terms and more) sequentially and drifts 2.3 % from exact
    arithmetic by the 8th iteration at 524,288 rows, while the device's tree
    reductions stay within 1e-4."""
    per = 1024 * ctx.info()["num_cu"]
    n = {"two_per_thread": 2 * per, "at_limit": 8 * per, "above_limit": 8 * per + 1000}[where]
    rng = np.random.default_rng(21)
    lo, up = -1.0 - 0.3 * rng.random(n), -1.0 + 0.3 * rng.random(n)
    r = np.repeat(np.arange(n), 3).reshape(n, 3)
    c = r + np.array([-1, 0, 1])
    v = np.stack([lo, 4.0 + rng.random(n), up], axis=1)
    keep = (c >= 0) & (c < n)
    row = np.zeros(n + 1, np.int32)
    row[1:] = np.cumsum(keep.sum(axis=1))
    col, val = c[keep].astype(np.int32), v[keep].astype(dtype)
    rhs = rng.uniform(-1, 1, n).astype(dtype)
    A = pkg.CsrMatrix(ctx, row, col, val)
    P = pkg.Preconditioner.jacobi(A)
    print(f"gmres mgs case {where} {np.dtype(dtype).name}: n {n}", flush=True)
    X = pkg.Vector(ctx, n, dtype)
    info = pkg.gmres(A, pkg.Vector(ctx, rhs, dtype), X, P, m=8, maxits=10, atol=0.0, rtol=1e-30, history=True,
                     check=False)
    OA = oracle.Csr(row, col, val.astype(np.float64))
    xo, oi = oracle.gmres(OA, rhs.astype(np.float64), P=oracle.Prec.jacobi(OA), m=8, ctrl=(10, 0.0, 1e-30, 1e6))
    assert info["iters"] == oi["iters"] == 10
    f64 = dtype == np.float64
    assert np.allclose(info["hist"], oi["hist"], rtol=1e-9 if f64 else 1e-4)
    assert relnorm(X.numpy(), xo) < (1e-10 if f64 else 1e-4)


def test_defect_correction_matches_fp64(pkg, ctx, oracle):
    row, col, val, rhs = fv3(pkg, 20, sigma=1.0, seed=5)
    nc, color, sizes, perm = pkg.multicoloring(row, col)
    prow, pcol, pval = pkg.csr_permute(row, col, val, perm)
    b = rhs[perm]
    Ah = pkg.CsrMatrix(ctx, prow, pcol, pval)
    Al = pkg.CsrMatrix.cast_of(Ah, np.float32)
    Pl = pkg.Preconditioner.mcsgs(Al, sizes)
    X = pkg.Vector(ctx, len(b))
    nb = np.linalg.norm(b)
    info = pkg.dc(Ah, Al, pkg.Vector(ctx, b), X, Pl, outer=(30, 1e-10 * nb, 0.0, 1e30),
                  inner=(1000, 0.0, 1e-4, 1e30))
    assert info["state"] == pkg.SUCCESS_ABS
    assert true_res(prow, pcol, pval, X.numpy(), b) < 1e-10
    OA = oracle.Csr(prow, pcol, pval)
    xo, oi = oracle.mpir(OA, b, Pl=oracle.Prec(oracle.PREC_MCSGS, OA, sizes=sizes), max_ecl=30,
                         abs_tol=1e-10 * nb, rel_tol=0.0, div_tol=1e30, ecl=(1000, 0.0, 1e-4, 1e30))
    assert abs(info["outer_iters"] - oi["n_ecl"]) <= 1
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    xd = spla.spsolve(sp.csr_matrix((pval, pcol, prow)).tocsc(), b)
    assert relnorm(X.numpy(), xd) < 1e-8


def test_defect_correction_res0_reset(pkg, ctx, oracle):
    """mpir.cc:275 resets res0_ after every correction, so a relative tolerance
    below 1 never ends the loop: exactly maxits corrections and no success,
    as the oracle's MPIR, although the first correction already halves the
    residual.  An inner absolute gate above the residual stops before any."""
    row, col, val, rhs = fv3(pkg, 12, sigma=0.5, seed=3)
    A = pkg.CsrMatrix(ctx, row, col, val)
    Al = pkg.CsrMatrix.cast_of(A, np.float32)
    Pl = pkg.Preconditioner.jacobi(Al)
    OA = oracle.Csr(row, col, val)
    for k in (1, 3):
        X = pkg.Vector(ctx, len(rhs))
        info = pkg.dc(A, Al, pkg.Vector(ctx, rhs), X, Pl, outer=(k, 0.0, 0.5, 1e30),
                      inner=(1000, 0.0, 0.125, 1e30), check=False)
        xo, oi = oracle.mpir(OA, rhs, Pl=oracle.Prec.jacobi(OA), max_ecl=k, abs_tol=0.0, rel_tol=0.5,
                             div_tol=1e30, ecl=(1000, 0.0, 0.125, 1e30))
        assert oi["n_ecl"] == k and oi["ret"] == 1
        assert info["outer_iters"] == k and info["state"] == pkg.FAIL_MAXITS and info["rc"] == pkg.ERR_NOT_CONVERGED
        assert info["hist"][1] < 0.5 * info["hist"][0]
        assert relnorm(info["hist"], oi["hist"]) < 1e-5
    X = pkg.Vector(ctx, len(rhs))
    info = pkg.dc(A, Al, pkg.Vector(ctx, rhs), X, Pl, outer=(4, 0.0, 0.0, 1e30),
                  inner=(1000, 2.0 * np.linalg.norm(rhs), 0.125, 1e30), check=False)
    assert info["outer_iters"] == 0 and info["state"] == pkg.FAIL_MAXITS and np.all(X.numpy() == 0.0)


def test_defect_correction_argument_checks(pkg, ctx):
    row, col, val, rhs = fv3(pkg, 6)
    A = pkg.CsrMatrix(ctx, row, col, val)
    Al = pkg.CsrMatrix.cast_of(A, np.float32)
    B = pkg.Vector(ctx, rhs)
    for args in (dict(x=B), dict(outer=(5, -1.0, 0.0, 1e6)), dict(inner=(100, 0.0, 0.1, 0.5)),
                 dict(outer=(5, 0.0, -0.1, 1e6))):
        kw = dict(x=pkg.Vector(ctx, len(rhs)), outer=(5, 0.0, 1e-8, 1e6), inner=(100, 0.0, 0.1, 1e6))
        kw.update(args)
        with pytest.raises(pkg.HfhError) as e:
            pkg.dc(A, Al, B, kw["x"], None, outer=kw["outer"], inner=kw["inner"])
        assert e.value.code == pkg.ERR_INVALID


@pytest.mark.parametrize("delta", [1e-1, 1e-3])
def test_defect_correction_keep_direction(pkg, ctx, delta):
    """Reliable-update form (no reference counterpart): the inner CG keeps its
    direction across FP64 corrections.  Pinned by what the mathematics fixes:
    the FP64 true residual meets the tolerance, the solution matches a sparse
    direct solve, and keeping the direction needs no more inner iterations
    than restarting at every correction."""
    import scipy.sparse as sp
    import scipy.sparse.linalg as spla
    row, col, val, rhs = fv3(pkg, 20, sigma=1.0, seed=5)
    nc, color, sizes, perm = pkg.multicoloring(row, col)
    prow, pcol, pval = pkg.csr_permute(row, col, val, perm)
    b = rhs[perm]
    Ah = pkg.CsrMatrix(ctx, prow, pcol, pval)
    Al = pkg.CsrMatrix.cast_of(Ah, np.float32)
    Pl = pkg.Preconditioner.mcsgs(Al, sizes)
    nb = np.linalg.norm(b)
    runs = {}
    for keep in (False, True):
        X = pkg.Vector(ctx, len(b))
        info = pkg.dc(Ah, Al, pkg.Vector(ctx, b), X, Pl, outer=(200, 1e-10 * nb, 0.0, 1e30),
                      inner=(1000, 0.0, delta, 1e30), keep_direction=keep)
        assert info["state"] == pkg.SUCCESS_ABS
        assert true_res(prow, pcol, pval, X.numpy(), b) < 1e-10
        runs[keep] = (info, X.numpy())
    assert runs[True][0]["inner_iters"] <= runs[False][0]["inner_iters"]
    xd = spla.spsolve(sp.csr_matrix((pval, pcol, prow)).tocsc(), b)
    assert relnorm(runs[True][1], xd) < 1e-8


def test_solver_failure_codes(pkg, ctx):
    row, col, val, rhs = fv3(pkg, 10)
    A = pkg.CsrMatrix(ctx, row, col, val)
    X = pkg.Vector(ctx, len(rhs))
    info = pkg.pcg(A, pkg.Vector(ctx, rhs), X, None, maxits=1, atol=0.0, rtol=1e-12, check=False)
    assert info["rc"] == pkg.ERR_NOT_CONVERGED and info["state"] == pkg.FAIL_MAXITS and info["iters"] == 1
    bn = rhs.copy()
    bn[3] = np.nan
    X.zeros()
    info = pkg.pcg(A, pkg.Vector(ctx, bn), X, pkg.Preconditioner.jacobi(A), maxits=7, atol=0.0, rtol=1e-8,
                   check=False)
    assert info["state"] == pkg.FAIL_MAXITS and info["iters"] == 7
    with pytest.raises(pkg.HfhError) as e:
        pkg.pcg(A, pkg.Vector(ctx, rhs), X, None, maxits=0)
    assert e.value.code == pkg.ERR_INVALID


def test_solver_3x3_known_answer(pkg, ctx):
    # test/LinearAlgebra/cg.cc:89-91 through the device drivers
    row = np.array([0, 3, 6, 9], np.int32)
    col = np.array([0, 1, 2] * 3, np.int32)
    val = np.array([2.0, -1.0, 0.0, -1.0, 2.0, -1.0, 0.0, -1.0, 2.0])
    A = pkg.CsrMatrix(ctx, row, col, val)
    B = pkg.Vector(ctx, [1.0, -2.0, 0.0])
    for solve in (lambda X: pkg.pcg(A, B, X, None, maxits=100, atol=1e-15, rtol=1e-12),
                  lambda X: pkg.pcg(A, B, X, pkg.Preconditioner.jacobi(A), maxits=100, atol=1e-15, rtol=1e-12),
                  lambda X: pkg.gmres(A, B, X, pkg.Preconditioner.jacobi(A), m=3, maxits=100, atol=1e-15,
                                      rtol=1e-12)):
        X = pkg.Vector(ctx, 3)
        solve(X)
        assert np.allclose(X.numpy(), [-0.25, -1.5, -0.75], atol=1e-12)


def test_blas1_beyond_cache_bitwise(pkg, ctx):
    """Axpy / ScaleAdd / Scale on vectors larger than the 256 MB Infinity Cache
    run with non-temporal streams from more workgroups: the same bits as the
    separately rounded CPU loops."""
    n = (1 << 25) + 5                      # 2 x 256 MiB + 80 B of FP64
    rng = np.random.default_rng(21)
    x = rng.uniform(-1, 1, n)
    y = rng.uniform(-1, 1, n)
    X, Y = pkg.Vector(ctx, x), pkg.Vector(ctx, y)
    a = 0.37
    Y.axpy(X, a)
    ref = y + a * x                         # product rounded, then the sum: no contraction
    assert np.array_equal(Y.numpy(), ref)
    Y.scaleadd(-1.7, X)
    ref = -1.7 * ref + x
    assert np.array_equal(Y.numpy(), ref)
    Y.scale(0.5)
    assert np.array_equal(Y.numpy(), ref * 0.5)
