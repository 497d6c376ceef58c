"""The rest of lVector<T> / lMatrix<T> on the HIP platform against the oracle:
Norm1, ArgMin / ArgMax / NormMax, add_value(s), the index set, partial
add, Rot / Rotm (CPUsimple_lvector.cc, lvector_cpu.cc), and the matrix value
operations Zeros / Scale / ScaleOffdiag / ZeroRows / ZeroCols / Pjacobi /
extract_(inv)diagelements (lmatrix_csr_cpu.cc).  Element-wise and integer
results are bit-identical; Norm1 is summed in another order and is checked to
1e-14 relative (FP64) / 1e-6 (FP32), as Dot is."""
import numpy as np
import pytest

from conftest import relnorm

pytestmark = pytest.mark.gpu

DTYPES = [np.float64, np.float32]


def fv3(pkg, n=10, seed=1, sigma=1.0):
    return pkg.problem(pkg.FV_DIFFUSION, n, p0=sigma, seed=seed)


@pytest.mark.parametrize("dt", DTYPES)
@pytest.mark.parametrize("n", [1, 5, 7, 1000, (1 << 20) + 3])
def test_norm1(pkg, ctx, oracle, dt, n):
    x = np.random.default_rng(n).normal(size=n).astype(dt)
    got = pkg.Vector(ctx, x, dt).norm1()
    ref = oracle.norm1(x)
    if dt == np.float64:
        assert abs(got - ref) <= 1e-14 * abs(ref)
    else:
        # the device sums FP32 magnitudes in FP64 partials and rounds the
        # total to FP32 once (the reference returns T); the reference sums
        # them in FP32, one add after another, whose error grows to about
        # sqrt(n) * 2^-24 (1.6e-6 relative at 2^20 entries): both are
        # checked against the exact sum, the device within one FP32
        # rounding, the oracle within the sequential bound n * 2^-24
        exact = float(np.sum(np.abs(x.astype(np.float64))))
        assert abs(got - exact) <= (2.0 ** -24 + 1e-12) * exact
        assert abs(ref - exact) <= max(n, 1) * 2.0 ** -24 * exact


@pytest.mark.parametrize("dt", DTYPES)
def test_argmax_argmin_normmax(pkg, ctx, oracle, dt):
    rng = np.random.default_rng(7)
    for n in (1, 2, 255, 256, 257, 100003):
        # few distinct values: many ties, which keep the first index
        x = rng.integers(-40, 41, n).astype(dt)
        X = pkg.Vector(ctx, x, dt)
        assert X.argmax() == oracle.argmax(x)
        assert X.argmin() == oracle.argmin(x)
        assert X.normmax() == abs(x[oracle.argmax(x)])
    cases = [[np.nan, 5.0, 9.0], [1.0, np.nan, -9.0, 9.0], [4.0, np.nan, -0.5, 0.5], [-0.0, 0.0],
             [np.nan] * 5, [3.0] + [np.nan] * 700 + [2.0]]
    for c in cases:
        x = np.array(c, dt)
        X = pkg.Vector(ctx, x, dt)
        assert X.argmax() == oracle.argmax(x), c
        assert X.argmin() == oracle.argmin(x), c
    with pytest.raises(pkg.HfhError):
        pkg.Vector(ctx, 0, dt).argmax()


@pytest.mark.parametrize("dt", DTYPES)
def test_add_values_in_order(pkg, ctx, oracle, dt):
    rng = np.random.default_rng(3)
    n = 5000
    y = rng.normal(size=n).astype(dt)
    idx = rng.integers(0, n, 20000).astype(np.int32)   # many repeats
    v = (rng.normal(size=20000) * np.where(rng.random(20000) < 0.1, 1e7, 1.0)).astype(dt)
    Y = pkg.Vector(ctx, y, dt)
    Y.add_values(idx, v)
    assert np.array_equal(Y.numpy(), oracle.add_values(y, idx, v))
    # add_value is a single entry
    Y.add_values([4], [1.5])
    assert Y.numpy()[4] == oracle.add_values(oracle.add_values(y, idx, v), [4], [1.5])[4]
    with pytest.raises(pkg.HfhError):
        Y.add_values([n], [1.0])


@pytest.mark.parametrize("dt", DTYPES)
def test_set_values_last_write_wins(pkg, ctx, dt):
    X = pkg.Vector(ctx, np.zeros(64, dt), dt)
    X.set_values([3, 5, 3, 3, 9], [1.0, 2.0, 3.0, 4.0, 5.0])
    ref = np.zeros(64, dt)
    for i, v in zip([3, 5, 3, 3, 9], [1.0, 2.0, 3.0, 4.0, 5.0]):
        ref[i] = v
    assert np.array_equal(X.numpy(), ref)


@pytest.mark.parametrize("dt", DTYPES)
def test_indexset(pkg, ctx, dt):
    # lvector_cpu.cc:217-297: set / get the index set, gather and scatter through it
    rng = np.random.default_rng(5)
    x = rng.normal(size=300).astype(dt)
    X = pkg.Vector(ctx, x, dt)
    assert len(X.get_indexset()) == 0
    assert len(X.get_indexed_values()) == 0
    idx = np.array([7, 0, 299, 7, 150], np.int32)
    X.set_indexset(idx)
    assert np.array_equal(X.get_indexset(), idx)
    assert np.array_equal(X.get_indexed_values(), x[idx])
    vals = np.array([1.0, 2.0, 3.0, 4.0, 5.0], dt)
    X.set_indexed_values(vals)
    ref = x.copy()
    for i, v in zip(idx, vals):
        ref[i] = v
    assert np.array_equal(X.numpy(), ref)
    with pytest.raises(pkg.HfhError):
        X.set_indexset([300])
    X.set_indexset([])
    assert len(X.get_indexset()) == 0


@pytest.mark.parametrize("dt", DTYPES)
def test_partial_add_subvector(pkg, ctx, oracle, dt):
    rng = np.random.default_rng(6)
    y, x = rng.normal(size=1000).astype(dt), rng.normal(size=700).astype(dt)
    Y, Xv = pkg.Vector(ctx, y, dt), pkg.Vector(ctx, x, dt)
    Y.add_range(100, Xv, 33, 600, 0.37)
    assert np.array_equal(Y.numpy(), oracle.add_range(y, 100, x, 33, 600, dt(0.37)))
    with pytest.raises(pkg.HfhError):
        Y.add_range(500, Xv, 0, 600, 1.0)
    # one vector with disjoint ranges
    Y2 = pkg.Vector(ctx, y, dt)
    Y2.add_range(0, Y2, 500, 400, -2.0)
    assert np.array_equal(Y2.numpy(), oracle.add_range(y, 0, y.copy(), 500, 400, dt(-2.0)))
    with pytest.raises(pkg.HfhError):
        Y2.add_range(0, Y2, 10, 400, 1.0)


@pytest.mark.parametrize("dt", DTYPES)
def test_rot_and_rotm(pkg, ctx, oracle, dt):
    rng = np.random.default_rng(8)
    a, b = rng.normal(size=4099).astype(dt), rng.normal(size=4099).astype(dt)
    A, B = pkg.Vector(ctx, a, dt), pkg.Vector(ctx, b, dt)
    c, s = np.cos(0.7), np.sin(0.7)
    A.rot(B, c, s)
    ra, rb = oracle.rot(a, b, dt(c), dt(s))
    assert np.array_equal(A.numpy(), ra) and np.array_equal(B.numpy(), rb)
    for p in ([-1.0, 0.3, -1.7, 2.2, 0.9], [0.0, 9.0, -0.4, 1.3, 9.0], [1.0, 0.6, 9.0, 9.0, -1.1],
              [-2.0, 1.0, 1.0, 1.0, 1.0], [0.5, 1.0, 1.0, 1.0, 1.0]):
        U, V = pkg.Vector(ctx, a, dt), pkg.Vector(ctx, b, dt)
        U.rotm(V, p)
        ru, rv = oracle.rotm(a, b, np.array(p, dt))
        assert np.array_equal(U.numpy(), ru) and np.array_equal(V.numpy(), rv), p


@pytest.mark.parametrize("dt", DTYPES)
def test_matrix_value_ops(pkg, ctx, oracle, dt):
    row, col, val, rhs = fv3(pkg, 9, seed=4)
    val = val.astype(dt)
    OA = oracle.Csr(row, col, val)
    rows = np.array([0, 17, 17, 400, 728], np.int32)
    A = pkg.CsrMatrix(ctx, row, col, val)
    A.zero_rows(rows, 2.5)
    assert np.array_equal(A.download()[2], oracle.zero_rows(OA, rows, dt(2.5)))
    A = pkg.CsrMatrix(ctx, row, col, val)
    A.zero_cols([3, 81, 3], -1.0)
    assert np.array_equal(A.download()[2], oracle.zero_cols(OA, [3, 81, 3], dt(-1.0)))
    for off in (False, True):
        A = pkg.CsrMatrix(ctx, row, col, val)
        A.scale(0.37, off)
        assert np.array_equal(A.download()[2], oracle.mat_scale(OA, dt(0.37), off))
    # the SpMV follows the new values (its copies are refilled)
    x = np.random.default_rng(1).normal(size=len(rhs)).astype(dt)
    Y = pkg.Vector(ctx, len(rhs), dt)
    A.spmv(pkg.Vector(ctx, x, dt), Y)
    assert np.array_equal(Y.numpy(), oracle.spmv(oracle.Csr(row, col, oracle.mat_scale(OA, dt(0.37), True)), x))
    A.zeros()
    assert not A.download()[2].any()
    A = pkg.CsrMatrix(ctx, row, col, val)
    Z = pkg.Vector(ctx, np.full(len(rhs), 7.0, dt), dt)
    A.pjacobi(pkg.Vector(ctx, x, dt), Z)
    assert np.array_equal(Z.numpy(), oracle.pjacobi(OA, x))
    d = np.diag(OA.to_dense()).astype(dt)
    assert np.array_equal(A.diag(10, 500).numpy(), d[10:500])
    inv = oracle.invdiag(OA)
    assert np.array_equal(A.diag(3, 700, invert=True).numpy(), inv[3:700])
    with pytest.raises(pkg.HfhError):
        A.zero_rows([len(rhs)], 1.0)
    with pytest.raises(pkg.HfhError):
        A.zero_cols([], 1.0)
