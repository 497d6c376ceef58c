"""Preconditioner setup on the device: level schedules, first-fit
multicolouring and the symmetric permutation against the host passes (which
the oracle pins), bit for bit, and an MC-SGS solve set up entirely on the
device against one set up on the host."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def ragged(rng, n=700):
    """Non-symmetric random pattern: ragged rows, some without a diagonal,
    a run of empty rows."""
    row, cols = [0], []
    for i in range(n):
        c = set()
        if not 200 <= i < 230:
            c = set(rng.choice(n, size=int(rng.integers(1, 10)), replace=False).tolist())
            if i % 7:
                c.add(i)
        c = sorted(c)
        cols.extend(c)
        row.append(row[-1] + len(c))
    row = np.array(row, np.int32)
    col = np.array(cols, np.int32)
    return row, col, rng.uniform(-1, 1, len(col))


def cases(pkg, name):
    if name == "fv_3d":
        return pkg.problem(pkg.FV_DIFFUSION, 10, p0=1.0, seed=3)[:3]
    if name == "q1_2d":
        return pkg.problem(pkg.Q1_2D, 40)[:3]
    if name == "q1_3d":
        return pkg.problem(pkg.Q1_3D, 12)[:3]
    if name == "upwind":
        return pkg.problem(pkg.UPWIND_CONVDIFF, 12, p0=1e-1, p1=2.0, seed=4)[:3]
    if name == "dense100":        # 100 colours: the search above 64
        n = 100
        row = (np.arange(n + 1) * n).astype(np.int32)
        col = np.tile(np.arange(n), n).astype(np.int32)
        return row, col, np.random.default_rng(9).uniform(-1, 1, n * n)
    return ragged(np.random.default_rng(13))


NAMES = ["fv_3d", "q1_2d", "q1_3d", "upwind", "dense100", "ragged"]


@pytest.mark.parametrize("name", NAMES)
def test_levels_match_host(pkg, ctx, name):
    row, col, val = cases(pkg, name)
    A = pkg.CsrMatrix(ctx, row, col, val)
    for upper in (False, True):
        p_dev, r_dev = A.levels(upper)
        p_host, r_host = pkg.levels(row, col, upper)
        assert np.array_equal(p_dev, p_host) and np.array_equal(r_dev, r_host)


@pytest.mark.parametrize("name", NAMES)
def test_multicoloring_and_permutation_match_host(pkg, ctx, oracle, name):
    row, col, val = cases(pkg, name)
    A = pkg.CsrMatrix(ctx, row, col, val)
    nc, color, sizes, perm = A.multicoloring()
    nc_h, color_h, sizes_h, perm_h = pkg.multicoloring(row, col)
    assert nc == nc_h and np.array_equal(color, color_h)
    assert np.array_equal(sizes, sizes_h) and np.array_equal(perm, perm_h)
    if name == "dense100":
        assert nc == 100
    P = A.permuted(perm)
    prow, pcol, pval = P.download()
    hrow, hcol, hval = pkg.csr_permute(row, col, val, perm)
    assert np.array_equal(prow, hrow) and np.array_equal(pcol, hcol) and np.array_equal(pval, hval)
    x = np.random.default_rng(2).uniform(-1, 1, len(row) - 1)
    Y = pkg.Vector(ctx, len(x))
    P.spmv(pkg.Vector(ctx, x), Y)
    assert np.array_equal(Y.numpy(), oracle.spmv(oracle.Csr(hrow, hcol, hval), x))


def test_multicoloring_without_off_diagonal_is_refused(pkg, ctx):
    n = 50
    row = np.arange(n + 1, dtype=np.int32)
    col = np.arange(n, dtype=np.int32)
    A = pkg.CsrMatrix(ctx, row, col, np.full(n, 2.0))
    with pytest.raises(pkg.HfhError):
        A.multicoloring()
    with pytest.raises(ValueError):
        pkg.multicoloring(row, col)
    with pytest.raises(pkg.HfhError):
        A.permuted(np.array([0] * n))          # not a permutation


def test_mcsgs_pcg_with_device_setup(pkg, ctx):
    """Colouring, permutation and sweep copies from the device: the same PCG
    iterates as the host setup."""
    row, col, val, rhs = pkg.problem(pkg.FV_DIFFUSION, 16, p0=1.0, seed=5)
    nc, color, sizes, perm = pkg.multicoloring(row, col)
    prow, pcol, pval = pkg.csr_permute(row, col, val, perm)
    Ah = pkg.CsrMatrix(ctx, prow, pcol, pval)
    Ph = pkg.Preconditioner.mcsgs(Ah, sizes)
    A = pkg.CsrMatrix(ctx, row, col, val)
    nc_d, _, sizes_d, perm_d = A.multicoloring()
    Ad = A.permuted(perm_d)
    Pd = pkg.Preconditioner.mcsgs(Ad, sizes_d)
    b = rhs[perm]
    Xh, Xd = pkg.Vector(ctx, len(b)), pkg.Vector(ctx, len(b))
    ih = pkg.pcg(Ah, pkg.Vector(ctx, b), Xh, Ph, rtol=1e-10)
    idv = pkg.pcg(Ad, pkg.Vector(ctx, b), Xd, Pd, rtol=1e-10)
    assert ih["iters"] == idv["iters"]
    assert np.array_equal(Xh.numpy(), Xd.numpy())
