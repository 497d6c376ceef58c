"""The seeded benchmark systems assembled on the device against the host
generator (whose matrices the closed-form stencil tests and the reference's
set_C system pin): row pointer, columns, values and right-hand side bit for
bit, FP64 and FP32."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    ("q1_2d_40", dict(kind=1, nx=40)),
    ("q1_3d_12", dict(kind=2, nx=12)),
    ("q1_3d_rect", dict(kind=2, nx=9, ny=5, nz=7)),
    ("fv_3d_10", dict(kind=3, nx=10, p0=1.0, seed=3)),
    ("fv_2d_24", dict(kind=3, nx=24, nz=1, p0=0.5, seed=7)),
    ("upwind_3d_12", dict(kind=4, nx=12, p0=1e-2, p1=5.0, seed=4)),
    ("upwind_2d_30", dict(kind=4, nx=30, nz=1, p0=1e-1, p1=2.0, seed=5)),
    ("fv_3d_48", dict(kind=3, nx=48, p0=1.0, seed=11)),
]


@pytest.mark.parametrize("name,spec", CASES, ids=[c[0] for c in CASES])
@pytest.mark.parametrize("dtype", [np.float64, np.float32])
def test_assembly_matches_host_generator(pkg, ctx, oracle, name, spec, dtype):
    A, b = pkg.assemble(ctx, dtype=dtype, **spec)
    row, col, val, rhs = pkg.problem(**spec)
    r, c, v = A.download()
    assert np.array_equal(r, row) and np.array_equal(c, col)
    assert np.array_equal(v, val.astype(dtype))
    assert np.array_equal(b.numpy(), rhs.astype(dtype))
    # the assembled matrix is a working operator: SpMV bit-identical to the oracle
    x = np.random.default_rng(1).uniform(-1, 1, len(rhs)).astype(dtype)
    Y = pkg.Vector(ctx, len(x), dtype)
    A.spmv(pkg.Vector(ctx, x, dtype), Y)
    assert np.array_equal(Y.numpy(), oracle.spmv(oracle.Csr(row, col, val.astype(dtype)), x))


def test_assembly_refuses_mismatched_matrix(pkg, ctx):
    lib = pkg.hip()
    h = C.c_void_p()
    assert lib.hfh_csr_create(ctx.h, 1000, 1000, 123, pkg.F64, C.byref(h)) == 0
    try:
        # 10^3 FV cells have 6400 entries, not 123; the FV problem needs coefficients
        kc = np.ones(1000)
        assert lib.hfh_problem_assemble(h, 3, 10, 10, 10, 1.0, 0.0, 0, kc.ctypes.data, None) == pkg.ERR_INVALID
        assert lib.hfh_problem_assemble(h, 3, 10, 10, 10, 1.0, 0.0, 0, None, None) == pkg.ERR_INVALID
        assert lib.hfh_problem_assemble(h, 9, 10, 10, 10, 1.0, 0.0, 0, None, None) == pkg.ERR_INVALID
        assert lib.hfh_problem_assemble(h, 2, 11, 10, 10, 1.0, 0.0, 0, None, None) == pkg.ERR_INVALID
    finally:
        lib.hfh_csr_destroy(h)
