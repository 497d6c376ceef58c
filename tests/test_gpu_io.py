"""External systems through the MatrixMarket path run on the HIP platform:
the reference's set_C system (test/PETSc/data/set_C) is written as a
MatrixMarket file from its fixture, read back, solved on the device and
compared with the reference's sol.txt (7 significant digits)."""
import os

import numpy as np
import pytest
import scipy.sparse as sp

from conftest import relnorm

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_set_C_from_matrix_market(pkg, ctx, tmp_path):
    d = np.load(os.path.join(GOLD, "set_C.npz"))
    A = sp.coo_matrix((d["vals"], (d["rows"], d["cols"]))).tocsr()
    A.sort_indices()
    path = tmp_path / "set_C.mtx"
    pkg.mm_write(str(path), A.indptr, A.indices, A.data)
    row, col, val, ncol = pkg.mm_read(str(path))
    M = pkg.CsrMatrix(ctx, row, col, val)
    X = pkg.Vector(ctx, d["x0"])
    info = pkg.pcg(M, pkg.Vector(ctx, d["rhs"]), X, pkg.Preconditioner.jacobi(M), maxits=5000, atol=1e-15,
                   rtol=1e-12)
    assert info["state"] in (pkg.SUCCESS_ABS, pkg.SUCCESS_REL)
    x = X.numpy()
    assert relnorm(x, d["sol"]) < 2e-6
    assert np.allclose(x, d["sol"], rtol=1e-5, atol=1e-6)
    # and GMRES(30) with ILU(0) factored on the device
    LU = pkg.CsrMatrix(ctx, row, col, val)
    assert LU.ilu0(row, col) == 0
    X.upload(d["x0"])
    info = pkg.gmres(M, pkg.Vector(ctx, d["rhs"]), X, pkg.Preconditioner.ilu0(LU, row, col), m=30, maxits=2000,
                     atol=1e-15, rtol=1e-12)
    assert relnorm(X.numpy(), d["sol"]) < 2e-6
