"""Pins the CPU oracle against the reference's own known answers, golden
fixtures and what the mathematics fixes (closed forms, invariants, dense
brute force).  CPU only."""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import relnorm

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def tridiag_coo(O, n, diag=2.0, off=-1.0, full=False):
    rows, cols, vals = [], [], []
    for i in range(n):
        for j in range(n) if full else (i - 1, i, i + 1):
            if 0 <= j < n:
                rows.append(i)
                cols.append(j)
                vals.append(diag if i == j else (off if abs(i - j) == 1 else 0.0))
    return O.csr_from_coo(n, rows, cols, vals)


def fv_matrix(pkg, O, nx, ny=None, nz=None, sigma=1.0, seed=7):
    row, col, val, rhs = pkg.problem(pkg.FV_DIFFUSION, nx, ny, nz, p0=sigma, seed=seed)
    return O.Csr(row, col, val), rhs


def to_scipy(A):
    return sp.csr_matrix((A.val.astype(np.float64), A.col, A.row), shape=(A.n, A.ncol))


# --------------------------------------------------------------- reference known answers

def test_spmv_known_answers(oracle):
    # test-legacy/lmp_test.cc:53-82: TransformFromCOO on tridiag(-1,2,-1)
    rows = [0, 0, 1, 1, 1, 2, 2]
    cols = [0, 1, 0, 1, 2, 1, 2]
    vals = [2.0, -1.0, -1.0, 2.0, -1.0, -1.0, 2.0]
    A = oracle.csr_from_coo(3, rows, cols, vals)
    assert np.array_equal(oracle.spmv(A, np.array([1.0, -2.0, 0.0])), [4.0, -5.0, 2.0])
    # test-legacy/cuda_test.cc:75-77
    y = oracle.spmv(A, np.array([1.2, -2.3, 0.7]))
    assert np.allclose(y, [4.7, -6.5, 3.7], atol=1e-12)
    assert np.array_equal(oracle.spmv(A, np.array([1.0, -2.0, 0.0]), omp=True), [4.0, -5.0, 2.0])


def test_csr_unordered_structure(oracle):
    # test/LinearAlgebra/CPUsimple_CSR_lMatrix.cc:159-200
    A = oracle.csr_from_coo(3, [1, 2, 0, 0, 1], [2, 1, 0, 1, 1])
    assert list(A.row) == [0, 2, 4, 5]
    assert list(A.col) == [0, 1, 1, 2, 1]
    B = oracle.csr_from_coo(3, [2, 0, 0, 1], [1, 0, 1, 1])
    assert list(B.row) == [0, 2, 3, 4]
    assert list(B.col) == [0, 1, 1, 1]


def test_csr_add_values(oracle):
    # test/LinearAlgebra/CPUsimple_CSR_lMatrix.cc:79-101
    A = oracle.csr_from_coo(3, [0, 0, 1, 1, 1, 2, 2], [0, 1, 0, 1, 2, 1, 2])
    A.val[:] = 0
    oracle.csr_add_values(A, [0, 1, 2], [0, 1, 2], np.arange(1.0, 10.0))
    assert list(A.val) == [1.0, 2.0, 4.0, 5.0, 6.0, 8.0, 9.0]


@pytest.mark.parametrize("prec", ["none", "jacobi"])
def test_cg_3x3(oracle, prec):
    # test/LinearAlgebra/cg.cc:89-91 (and test-legacy/cg_test.cc:87-89)
    A = tridiag_coo(oracle, 3, full=True)
    P = oracle.Prec.jacobi(A) if prec == "jacobi" else None
    x, info = oracle.cg(A, np.array([1.0, -2.0, 0.0]), P=P, ctrl=(100, 1e-15, 1e-12, 1e6))
    assert info["ret"] == 0 and info["iters"] <= 3
    assert np.allclose(x, [-0.25, -1.5, -0.75], atol=1e-12)


@pytest.mark.parametrize("flex", [False, True])
def test_gmres_3x3(oracle, flex):
    # test/LinearAlgebra/gmres.cc (same system, basis 3)
    A = tridiag_coo(oracle, 3, full=True)
    fn = oracle.fgmres if flex else oracle.gmres
    x, info = fn(A, np.array([1.0, -2.0, 0.0]), P=oracle.Prec.jacobi(A), m=3, ctrl=(100, 1e-15, 1e-12, 1e6))
    assert info["ret"] == 0
    assert np.allclose(x, [-0.25, -1.5, -0.75], atol=1e-12)


def test_equation1_fixture(oracle):
    # test/LinearAlgebra/solver.cc:106-156 with data/Equation1.*
    d = np.load(os.path.join(GOLD, "equation1.npz"))
    A = oracle.csr_from_coo(int(d["n"]), d["rows"], d["cols"], d["vals"])
    x, info = oracle.cg(A, d["rhs"], x0=d["x0"], ctrl=(100, 1e-15, 1e-12, 1e6))
    assert np.allclose(x, d["sol"], atol=1e-12)


def test_set_C_fixture(oracle):
    # test/PETSc/PETScLinearSolverReadData.cc:71-215 on data/set_C (tolerance 1e-2 there)
    d = np.load(os.path.join(GOLD, "set_C.npz"))
    A = oracle.csr_from_coo(int(d["n"]), d["rows"], d["cols"], d["vals"])
    for solve in (lambda: oracle.gmres(A, d["rhs"], x0=d["x0"], P=oracle.Prec.jacobi(A), m=30,
                                       ctrl=(5000, 1e-15, 1e-12, 1e6)),
                  lambda: oracle.cg(A, d["rhs"], x0=d["x0"], P=oracle.Prec.jacobi(A),
                                    ctrl=(5000, 1e-15, 1e-12, 1e6))):
        x, info = solve()
        assert info["ret"] == 0
        # sol.txt holds 7 significant digits
        assert relnorm(x, d["sol"]) < 2e-6
        assert np.allclose(x, d["sol"], rtol=1e-5, atol=1e-6)


# --------------------------------------------------------------- BLAS-1 orders and closed forms

def test_dot_unroll5_order(oracle):
    # CPUsimple_lvector.cc:160-209: the N mod 5 head first, then groups of five
    # summed left to right.  With x = [1, 1e16, 1, -1e16, 1, 1], y = 1:
    # head = 1; group = ((((1e16 + 1) - 1e16) + 1) + 1) = 2 -> 3.  A plain
    # left-to-right sum gives 2, a pairwise sum gives 2 as well.
    x = np.array([1.0, 1e16, 1.0, -1e16, 1.0, 1.0])
    assert oracle.dot(x, np.ones(6)) == 3.0
    # N = 10, no head: ((1e16+1)-1e16+1+1) + (1+...) = 2 + 5
    x = np.array([1e16, 1.0, -1e16, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0])
    assert oracle.dot(x, np.ones(10)) == 7.0
    # float variant keeps the same order in float arithmetic
    xf = np.array([1.0, 2.0**25, 1.0, -(2.0**25), 1.0, 1.0], np.float32)
    assert oracle.dot(xf, np.ones(6, np.float32)) == np.float32(3.0)


def test_blas1_closed_forms(oracle):
    assert oracle.nrm2(np.array([3.0, 4.0])) == 5.0
    # ScaleAdd is y = a*y + x (not y + a*x)
    assert list(oracle.scaleadd(np.array([1.0, 2.0]), 3.0, np.array([10.0, 20.0]))) == [13.0, 26.0]
    assert list(oracle.axpy(np.array([1.0, 2.0]), np.array([10.0, 20.0]), 3.0)) == [31.0, 62.0]
    assert list(oracle.scale(np.array([1.0, -2.0]), -0.5)) == [-0.5, 1.0]
    assert list(oracle.ewmult(np.array([2.0, 3.0]), np.array([4.0, -1.0]))) == [8.0, -3.0]


def test_invdiag_and_casts(oracle, pkg):
    A, _ = fv_matrix(pkg, oracle, 7, 6, 5)
    assert np.array_equal(oracle.invdiag(A), 1.0 / np.diag(A.to_dense()))
    # rows without a stored diagonal: the reference leaves its buffer entry
    # unset (lmatrix_csr_cpu.cc:2417-2424); the oracle defines it as 0
    B = oracle.csr_from_coo(3, [0, 1, 2], [1, 1, 0], [1.0, 4.0, 2.0])
    assert list(oracle.invdiag(B)) == [0.0, 0.25, 0.0]
    # CastFrom: static_cast, round to nearest even (numpy's astype is IEEE RN)
    x = np.random.default_rng(9).normal(size=1000) * np.logspace(-30, 30, 1000)
    x[:4] = [1.0 + 2.0**-24, 1.0 + 3 * 2.0**-24, -(1.0 + 2.0**-24), 0.0]
    f = oracle.cast_d2f(x)
    assert np.array_equal(f, x.astype(np.float32))
    assert f[0] == np.float32(1.0) and f[1] == np.float32(1.0 + 2.0**-22)
    assert np.array_equal(oracle.cast_f2d(f), f.astype(np.float64))


def test_spmv_matches_dense(oracle, pkg):
    rng = np.random.default_rng(1)
    A, _ = fv_matrix(pkg, oracle, 9, 7, 5)
    x = rng.uniform(-1, 1, A.n)
    assert relnorm(oracle.spmv(A, x), A.to_dense() @ x) < 1e-15
    y0 = rng.uniform(-1, 1, A.n)
    assert relnorm(oracle.spmv_add(A, x, y0), y0 + A.to_dense() @ x) < 1e-15
    # linearity
    z = rng.uniform(-1, 1, A.n)
    assert relnorm(oracle.spmv(A, 2.0 * x + z), 2.0 * oracle.spmv(A, x) + oracle.spmv(A, z)) < 1e-14


# --------------------------------------------------------------- preconditioner applies

def test_sgs_is_inverse_of_sgs_operator(oracle, pkg):
    A, _ = fv_matrix(pkg, oracle, 6, 5, 4)
    D = np.diag(np.diag(A.to_dense()))
    L = np.tril(A.to_dense(), -1)
    U = np.triu(A.to_dense(), 1)
    M = (D + L) @ np.linalg.inv(D) @ (D + U)
    r = np.random.default_rng(2).uniform(-1, 1, A.n)
    z = oracle.sgs(A, r)
    assert relnorm(M @ z, r) < 1e-13


@pytest.mark.parametrize("omega", [1.0, 1.3])
def test_mcsgs_operator(oracle, pkg, omega):
    A, _ = fv_matrix(pkg, oracle, 6, 5, 4)
    nc, color, sizes, perm = oracle.multicoloring(A)
    Ap = oracle.csr_reorder(A, perm)
    Dn = Ap.to_dense()
    D = np.diag(np.diag(Dn))
    L = np.tril(Dn, -1)
    U = np.triu(Dn, 1)
    r = np.random.default_rng(3).uniform(-1, 1, A.n)
    z = oracle.mcsgs(Ap, sizes, r, omega=omega)
    # forward (D + wL) y = r, t = D y / (w(2-w)), backward (D + wU) z = t
    M = omega * (2 - omega) * (D + omega * L) @ np.linalg.inv(D) @ (D + omega * U)
    assert relnorm(M @ z, r) < 1e-13
    if omega == 1.0:
        # a proper colouring makes MC-SGS natural-order SGS on the permuted matrix
        assert relnorm(z, oracle.sgs(Ap, r)) < 1e-13


def test_ilu0_tridiagonal_is_exact_lu(oracle):
    n = 40
    A = tridiag_coo(oracle, n, diag=3.0, off=-1.0)
    A.val[A.col > np.repeat(np.arange(n), np.diff(A.row))] = -0.5  # nonsymmetric
    LU, miss = oracle.ilu0(A)
    assert miss == 0
    r = np.random.default_rng(4).uniform(-1, 1, n)
    assert relnorm(oracle.ilu_solve(LU, r), np.linalg.solve(A.to_dense(), r)) < 1e-14


def test_ilu0_matches_pattern(oracle, pkg):
    # defining property of ILU(0): (L U)_ij = a_ij for every (i, j) in the pattern
    row, col, val, _ = pkg.problem(pkg.UPWIND_CONVDIFF, 5, p0=1e-2, p1=5.0, seed=1)
    A = oracle.Csr(row, col, val)
    LU, miss = oracle.ilu0(A)
    assert miss == 0
    F = LU.to_dense()
    L = np.tril(F, -1) + np.eye(A.n)
    U = np.triu(F)
    P = L @ U
    Ad = A.to_dense()
    mask = np.zeros_like(Ad, dtype=bool)
    for i in range(A.n):
        mask[i, A.col[A.row[i]:A.row[i + 1]]] = True
    assert np.max(np.abs(P[mask] - Ad[mask])) < 1e-14 * np.max(np.abs(Ad))
    r = np.random.default_rng(5).uniform(-1, 1, A.n)
    assert relnorm(P @ oracle.ilu_solve(LU, r), r) < 1e-13


# --------------------------------------------------------------- colouring and level schedules

def first_fit(A):
    colors = np.zeros(A.n, np.int64)
    for j in range(A.n):
        taken = {colors[c] for c in A.col[A.row[j]:A.row[j + 1]] if c != j}
        c = 1
        while c in taken:
            c += 1
        colors[j] = c
    return colors


@pytest.mark.parametrize("kind,dims,ncol", [("fv3", (6, 5, 4), 2), ("q1_2d", (9, 7), 4), ("q1_3d", (5, 4, 6), 8),
                                            ("upwind", (5,), 2)])
def test_multicoloring(oracle, pkg, kind, dims, ncol):
    if kind == "fv3":
        row, col, val, _ = pkg.problem(pkg.FV_DIFFUSION, *dims)
    elif kind == "q1_2d":
        row, col, val, _ = pkg.problem(pkg.Q1_2D, dims[0], dims[1])
    elif kind == "q1_3d":
        row, col, val, _ = pkg.problem(pkg.Q1_3D, *dims)
    else:
        row, col, val, _ = pkg.problem(pkg.UPWIND_CONVDIFF, dims[0], p0=1e-2, p1=5.0)
    A = oracle.Csr(row, col, val)
    nc, color, sizes, perm = oracle.multicoloring(A)
    assert nc == ncol
    # the reference's rescans reach the textbook first-fit colouring
    assert np.array_equal(color, first_fit(A))
    # proper colouring
    for i in range(A.n):
        nb = A.col[A.row[i]:A.row[i + 1]]
        assert np.all(color[nb[nb != i]] != color[i])
    # permutation: colour by colour, increasing rows within a colour
    assert np.array_equal(perm, np.lexsort((np.arange(A.n), color)))
    assert list(sizes) == [int(np.sum(color == c)) for c in range(1, nc + 1)]
    if kind == "fv3":  # red-black: colour = parity of i+j+k
        nx, ny, nz = dims
        k, j, i = np.meshgrid(range(nz), range(ny), range(nx), indexing="ij")
        assert np.array_equal(color, ((i + j + k) % 2 + 1).ravel())


def test_levelscheduling_grid(oracle, pkg):
    nx, ny, nz = 6, 5, 4
    row, col, val, _ = pkg.problem(pkg.FV_DIFFUSION, nx, ny, nz)
    A = oracle.Csr(row, col, val)
    nl, rep, sizes, perm = oracle.levelscheduling(A)
    # level of cell (i, j, k) is i + j + k; rows ascending within a level
    k, j, i = np.meshgrid(range(nz), range(ny), range(nx), indexing="ij")
    lev = (i + j + k).ravel()
    assert nl == rep == (nx - 1) + (ny - 1) + (nz - 1) + 1
    assert np.array_equal(perm, np.lexsort((np.arange(A.n), lev)))
    assert list(sizes) == [int(np.sum(lev == l)) for l in range(nl)]


def test_levelscheduling_chain_quirk(oracle):
    # a chain needs n levels; the reference then never records the count
    # (lmatrix_csr_cpu.cc:2324-2329) and reports 0 levels
    A = tridiag_coo(oracle, 6)
    nl, rep, sizes, perm = oracle.levelscheduling(A)
    assert nl == 6 and rep == 0 and list(perm) == list(range(6))


def test_reorder_is_symmetric_permutation(oracle, pkg):
    A, _ = fv_matrix(pkg, oracle, 5, 4, 3)
    perm = np.random.default_rng(6).permutation(A.n).astype(np.int32)
    B = oracle.csr_reorder(A, perm)
    Ad = A.to_dense()
    assert np.array_equal(B.to_dense(), Ad[np.ix_(perm, perm)])
    for i in range(B.n):
        assert np.all(np.diff(B.col[B.row[i]:B.row[i + 1]]) > 0)
    x = np.arange(A.n, dtype=np.float64)
    assert np.array_equal(oracle.vec_reorder(x, perm), x[perm])


# --------------------------------------------------------------- Krylov drivers

def test_cg_matches_scipy_and_bound(oracle, pkg):
    A, b = fv_matrix(pkg, oracle, 24, 24, 1, sigma=0.5, seed=3)
    S = to_scipy(A)
    invd = oracle.invdiag(A)
    x, info = oracle.cg(A, b, P=oracle.Prec.jacobi(A), ctrl=(2000, 0.0, 1e-8, 1e6))
    count = [0]
    xs, _ = spla.cg(S, b, rtol=1e-8, atol=0.0, maxiter=2000, M=sp.diags(invd),
                    callback=lambda xk: count.__setitem__(0, count[0] + 1))
    assert abs(info["iters"] - count[0]) <= 1
    xd = np.linalg.solve(S.toarray(), b)
    assert relnorm(x, xd) < 1e-6
    # residual history: recursive residual of the last step is the reported one
    assert info["hist"][-1] == info["res"] and info["res"] < 1e-8 * info["res0"]
    # classical bound with the Jacobi-scaled condition number
    Ds = np.sqrt(invd)
    ev = np.linalg.eigvalsh((Ds[:, None] * S.toarray()) * Ds[None, :])
    kappa = ev[-1] / ev[0]
    assert info["iters"] <= np.ceil(0.5 * np.sqrt(kappa) * np.log(2 / 1e-8)) + 2


def test_cg_diagonal_one_iteration(oracle):
    d = np.array([1.0, 2.0, 5.0, 7.0])
    A = oracle.csr_from_coo(4, [0, 1, 2, 3], [0, 1, 2, 3], d)
    x, info = oracle.cg(A, np.ones(4), P=oracle.Prec.jacobi(A), ctrl=(10, 1e-15, 1e-12, 1e6))
    assert info["iters"] == 1 and np.allclose(x, 1 / d)


def _krylov_min_residuals(Ad, Minv, r0, kmax):
    AM = Ad @ Minv
    K = [r0 / np.linalg.norm(r0)]
    out = []
    for j in range(1, kmax + 1):
        v = AM @ K[-1]
        for q in K:  # keep the explicit basis well conditioned
            v = v - (q @ v) * q
        K.append(v / np.linalg.norm(v))
        Q = np.array(K[:j]).T
        y, *_ = np.linalg.lstsq(AM @ Q, r0, rcond=None)
        out.append(np.linalg.norm(r0 - AM @ Q @ y))
    return np.array(out)


def test_gmres_minimises_residual(oracle, pkg):
    row, col, val, rhs = pkg.problem(pkg.UPWIND_CONVDIFF, 5, p0=1e-2, p1=5.0, seed=2)
    A = oracle.Csr(row, col, val)
    LU, _ = oracle.ilu0(A)
    P = oracle.Prec(oracle.PREC_ILU0, LU)
    x, info = oracle.gmres(A, rhs, P=P, m=30, ctrl=(30, 0.0, 1e-12, 1e6))
    F = LU.to_dense()
    Minv = np.linalg.inv((np.tril(F, -1) + np.eye(A.n)) @ np.triu(F))
    k = min(12, len(info["hist"]) - 1)
    ref = _krylov_min_residuals(A.to_dense(), Minv, rhs, k)
    got = info["hist"][1:k + 1]
    # rounding limits the agreement once the residual is tiny
    assert np.all(np.abs(got - ref) <= 1e-6 * ref + 1e-12 * np.linalg.norm(rhs))
    # FGMRES with a fixed preconditioner is GMRES
    xf, inf = oracle.fgmres(A, rhs, P=P, m=30, ctrl=(30, 0.0, 1e-12, 1e6))
    assert np.max(np.abs(inf["hist"] - info["hist"]) / info["hist"]) < 1e-8
    assert relnorm(xf, x) < 1e-8


def test_gmres_no_preconditioning(oracle, pkg):
    # GMRES::SolveNoPrecond (gmres.cc:103-264): the minimal residual over the
    # Krylov space of A itself; restarts reach the tolerance; the 3x3 answer
    row, col, val, rhs = pkg.problem(pkg.UPWIND_CONVDIFF, 5, p0=1e-2, p1=5.0, seed=3)
    A = oracle.Csr(row, col, val)
    x, info = oracle.gmres(A, rhs, m=30, ctrl=(30, 0.0, 1e-12, 1e6))
    k = min(12, len(info["hist"]) - 1)
    ref = _krylov_min_residuals(A.to_dense(), np.eye(A.n), rhs, k)
    got = info["hist"][1:k + 1]
    assert np.all(np.abs(got - ref) <= 1e-6 * ref + 1e-12 * np.linalg.norm(rhs))
    for flex in (False, True):
        fn = oracle.fgmres if flex else oracle.gmres
        x, info = fn(A, rhs, m=7, ctrl=(5000, 0.0, 1e-9, 1e6))
        assert info["ret"] == 0 and info["iters"] > 7
        assert np.linalg.norm(rhs - A.to_dense() @ x) < 2e-9 * np.linalg.norm(rhs)
    T = tridiag_coo(oracle, 3, full=True)
    x, info = oracle.gmres(T, np.array([1.0, -2.0, 0.0]), m=3, ctrl=(100, 1e-15, 1e-12, 1e6))
    assert np.allclose(x, [-0.25, -1.5, -0.75], atol=1e-12)


def test_gmres_restart(oracle, pkg):
    row, col, val, rhs = pkg.problem(pkg.UPWIND_CONVDIFF, 6, p0=1e-2, p1=5.0, seed=2)
    A = oracle.Csr(row, col, val)
    x, info = oracle.gmres(A, rhs, P=oracle.Prec.jacobi(A), m=5, ctrl=(2000, 0.0, 1e-8, 1e6))
    assert info["ret"] == 0 and info["iters"] > 5
    true = np.linalg.norm(rhs - A.to_dense() @ x)
    assert true < 2e-8 * np.linalg.norm(rhs)


def test_control_failures(oracle, pkg):
    A, b = fv_matrix(pkg, oracle, 8, 8, 1)
    x, info = oracle.cg(A, b, ctrl=(1, 0.0, 1e-12, 1e6))
    assert info["status"] == oracle.FAIL_MAXITS and info["ret"] == 1 and info["iters"] == 1
    # residual growth beyond dtol * res0 stops with divergence
    rng = np.random.default_rng(0)
    Q, _ = np.linalg.qr(rng.normal(size=(30, 30)))
    M = Q @ np.diag(np.logspace(0, 6, 30)) @ Q.T
    r_, c_ = np.nonzero(np.ones((30, 30)))
    D = oracle.csr_from_coo(30, r_, c_, M.ravel())
    x, info = oracle.cg(D, rng.uniform(-1, 1, 30), ctrl=(200, 0.0, 1e-14, 1.0 + 1e-9))
    assert info["status"] == oracle.FAIL_DIV and info["ret"] == 1
    # NaN in b: no test passes, the loop runs to maxits
    bn = b.copy()
    bn[3] = np.nan
    x, info = oracle.cg(A, bn, ctrl=(7, 0.0, 1e-8, 1e6))
    assert info["status"] == oracle.FAIL_MAXITS and info["iters"] == 7


def test_mpir(oracle, pkg):
    A, b = fv_matrix(pkg, oracle, 10, 10, 10, sigma=1.0, seed=5)
    nrm = np.linalg.norm(b)
    x, info = oracle.mpir(A, b, Pl=oracle.Prec.jacobi(A), max_ecl=30, abs_tol=1e-10 * nrm, rel_tol=1e-12,
                          ecl=(1000, 0.0, 1e-3, 1e6))
    assert info["ret"] == 0 and info["n_ech"] == 0
    true = np.linalg.norm(b - to_scipy(A) @ x)
    assert true < 1e-10 * nrm and abs(true - info["res"]) < 1e-6 * true
    # every loop reduces the true residual by about the inner tolerance
    h = info["hist"]
    assert np.all(h[1:] < h[:-1]) and np.all(h[1:] / h[:-1] < 1e-2)
    assert relnorm(x, spla.spsolve(to_scipy(A).tocsc(), b)) < 1e-8


def _float_gate(res0):
    """A double d whose float rounding lies on the other side of res0: the
    gate 'res0 > d' has one answer in double and the other in float."""
    lo = np.float32(res0)
    if float(lo) > res0:
        lo = np.nextafter(lo, np.float32(0))
    hi = np.nextafter(lo, np.float32(np.inf))
    mid = (float(lo) + float(hi)) / 2
    d = (res0 + mid) / 2      # between res0 and the rounding midpoint
    assert (res0 > d) != (res0 > float(np.float32(d)))
    return d, res0 > float(np.float32(d))


def test_mpir_res0_reset(oracle, pkg):
    """mpir.cc:275 resets res0_ to the residual just computed, after every
    correction, and ContinueIteration (mpir.cc:308-327) tests res_ > res0_*rel.
    With rel_tol < 1 that test compares the residual with itself, so it never
    ends the loop: exactly max_ecl corrections run and Solve returns
    kSolverExceeded (mpir.cc:300-303).  A rule relative to the initial residual
    would stop after the first correction, which already gains 10x here."""
    A, b = fv_matrix(pkg, oracle, 8, 8, 8, sigma=0.5, seed=3)
    for k in (1, 3, 5):
        x, info = oracle.mpir(A, b, Pl=oracle.Prec.jacobi(A), max_ecl=k, abs_tol=0.0, rel_tol=0.5,
                              ecl=(1000, 0.0, 0.125, 1e6))
        assert info["n_ecl"] == k and info["ret"] == 1
        h = info["hist"]
        assert len(h) == k + 1 and h[1] < 0.5 * h[0]
    # the error-correction gates use res0_ too: a relative gate of 0.5 never
    # closes, an absolute gate above the residual closes before any correction
    x, info = oracle.mpir(A, b, Pl=oracle.Prec.jacobi(A), max_ecl=4, abs_tol=0.0, rel_tol=0.0,
                          ecl=(1000, 0.0, 0.5, 1e6))
    assert info["n_ecl"] == 4
    nb = np.linalg.norm(b)
    x, info = oracle.mpir(A, b, Pl=oracle.Prec.jacobi(A), max_ecl=4, abs_tol=0.0, rel_tol=0.0,
                          ecl=(1000, 2.0 * nb, 0.125, 1e6))
    assert info["n_ecl"] == 0 and info["ret"] == 1 and np.all(x == 0.0)
    # abs_tol_ecl_ is a DataType_low member (mpir.h:225): the gate compares the
    # FP64 residual with the tolerance rounded to float
    res0 = float(np.linalg.norm(b))
    d, runs = _float_gate(res0)
    x, info = oracle.mpir(A, b, Pl=oracle.Prec.jacobi(A), max_ecl=1, abs_tol=0.0, rel_tol=0.0,
                          ecl=(1000, d, 0.125, 1e6))
    assert info["n_ecl"] == (1 if runs else 0)


# --------------------------------------------------------------- stationary Jacobi (jacobi.cc)

def test_swap_diag_front_and_nodiag(oracle):
    # SwapDiagElementsToRowFront trades the diagonal with the row's first entry
    A = oracle.csr_from_coo(3, [0, 0, 0, 1, 1, 2, 2], [0, 1, 2, 0, 1, 1, 2], [4.0, 1.0, 2.0, 3.0, 5.0, 6.0, 7.0])
    B = oracle.swap_diag_front(A)
    assert list(B.col) == [0, 1, 2, 1, 0, 2, 1]
    assert list(B.val) == [4.0, 1.0, 2.0, 5.0, 3.0, 7.0, 6.0]
    x = np.array([1.0, -2.0, 0.5])
    D = np.diag(np.diag(A.to_dense()))
    assert np.array_equal(oracle.spmv_nodiag(B, x), (A.to_dense() - D) @ x)
    # the stored order is the summation order: with the diagonal at the row
    # end, the swap moves the row's first entry last
    C = oracle.csr_from_coo(4, [0, 1, 2, 3, 3, 3, 3], [0, 1, 2, 0, 1, 2, 3], [1.0, 1.0, 1.0, 1.0, 1e16, -1e16, 2.0])
    Cs = oracle.swap_diag_front(C)
    assert list(Cs.col[3:]) == [3, 1, 2, 0]
    z = np.ones(4)
    assert oracle.spmv_nodiag(C, z)[3] == 0.0      # (1 + 1e16) - 1e16: the 1 is lost
    assert oracle.spmv_nodiag(Cs, z)[3] == 1.0     # (1e16 - 1e16) + 1


def test_jacobi_iteration(oracle, pkg):
    A, b = fv_matrix(pkg, oracle, 6, 5, 4, sigma=0.5, seed=4)
    S = oracle.swap_diag_front(A)
    Ad = A.to_dense()
    d = np.diag(Ad)
    x0 = np.random.default_rng(1).uniform(-1, 1, A.n)
    # one normal step: x = D^-1 (b - (A - D) x0); one damped step: (1-w) x0 + w (that)
    x1, _ = oracle.jacobi(S, b, x0=x0, mode=2, inner_iter=1, smooth_iter=1)
    ref = (b - (Ad - np.diag(d)) @ x0) / d
    assert relnorm(x1, ref) < 1e-15
    x1d, _ = oracle.jacobi(S, b, x0=x0, mode=3, w=0.7, inner_iter=1, smooth_iter=1)
    assert relnorm(x1d, 0.3 * x0 + 0.7 * ref) < 1e-15
    # solve mode: converges to A x = b at the tolerance, counting steps in inner_iter blocks
    x, info = oracle.jacobi(S, b, mode=0, inner_iter=2, ctrl=(20000, 0.0, 1e-8, 1e6))
    assert info["status"] == oracle.SUCCESS_REL and info["iters"] % 2 == 0
    assert np.linalg.norm(b - Ad @ x) < 1e-8 * np.linalg.norm(b)
    # the residual contracts at the spectral radius of I - D^-1 A
    rho = np.max(np.abs(np.linalg.eigvals(np.eye(A.n) - Ad / d[:, None])))
    h = info["hist"]
    assert h[-1] / h[0] < 50 * rho ** info["iters"]
    # a diagonal matrix is solved by one step
    Dm = oracle.csr_from_coo(4, [0, 1, 2, 3], [0, 1, 2, 3], [2.0, 4.0, 8.0, 0.5])
    xd, inf = oracle.jacobi(oracle.swap_diag_front(Dm), np.ones(4), mode=0, ctrl=(10, 0.0, 1e-12, 1e6))
    assert inf["iters"] == 1 and list(xd) == [0.5, 0.25, 0.125, 2.0]


# --------------------------------------------------------------- the rest of lVector / lMatrix

def test_norm1_order_and_closed_form(oracle):
    # CPUsimple_lvector.cc:95-140: the N mod 5 head first, left to right:
    # ((2^53 + 1) + 1) loses both ones; any order adding the ones first keeps them
    assert oracle.norm1(np.array([2.0**53, 1.0, 1.0])) == 2.0**53
    assert oracle.norm1(np.array([1.0, 1.0, 2.0**53])) == 2.0**53 + 2
    # groups of five: the group is summed first, then added to the total
    x = np.array([2.0**53] + [1.0] * 4 + [1.0] * 5 + [1.0])
    assert oracle.norm1(x) == 2.0**53 + 4 + 4
    x = np.random.default_rng(3).uniform(-8, 8, 1001).round()
    assert oracle.norm1(x) == np.abs(x).sum()
    assert oracle.norm1(x.astype(np.float32)) == np.float32(np.abs(x).sum())


def test_argmax_argmin(oracle):
    rng = np.random.default_rng(4)
    for n in (1, 2, 7, 1000):
        x = rng.integers(-5, 6, n).astype(np.float64)
        # ties keep the first index (numpy's argmax / argmin do as well)
        assert oracle.argmax(x) == int(np.argmax(np.abs(x)))
        assert oracle.argmin(x) == int(np.argmin(np.abs(x)))
    # the loop starts at |x_0|: a NaN there is never replaced; elsewhere it never wins
    assert oracle.argmax(np.array([np.nan, 5.0, 9.0])) == 0
    assert oracle.argmax(np.array([1.0, np.nan, -9.0, 9.0])) == 2
    assert oracle.argmin(np.array([4.0, np.nan, -0.5, 0.5])) == 2
    assert oracle.argmax(np.array([-0.0, 0.0])) == 0


def test_add_values_and_ranges(oracle):
    y = np.array([1.0, 2.0, 3.0])
    # a repeated index adds its values in order: (1 + 1e16) - 1e16 = 0, not 1
    out = oracle.add_values(y, [0, 2, 0, 0], [1e16, 5.0, -1e16, 0.5])
    assert list(out) == [0.5, 2.0, 8.0]
    y = np.arange(6.0)
    out = oracle.add_range(y, 2, np.array([10.0, 20.0, 30.0, 40.0]), 1, 3, 0.5)
    assert list(out) == [0.0, 1.0, 12.0, 18.0, 24.0, 5.0]


def test_rot_rotm_rotg(oracle):
    rng = np.random.default_rng(5)
    a, b = rng.normal(size=50), rng.normal(size=50)
    # a quarter turn: a' = b, b' = -a exactly
    ra, rb = oracle.rot(a, b, 0.0, 1.0)
    assert np.array_equal(ra, b) and np.array_equal(rb, -a)
    th = 0.3
    ra, rb = oracle.rot(a, b, np.cos(th), np.sin(th))
    assert np.allclose(np.hypot(ra, rb), np.hypot(a, b), rtol=1e-14)
    # Rotm: the argument vector v is the first operand (CPUsimple_lvector.cc:540-600)
    u, v = rng.integers(-9, 9, 20).astype(float), rng.integers(-9, 9, 20).astype(float)
    H = np.array([[2.0, 3.0], [-1.0, 4.0]])    # [[h11, h12], [h21, h22]]
    ru, rv = oracle.rotm(u, v, [-1.0, H[0, 0], H[1, 0], H[0, 1], H[1, 1]])
    assert np.array_equal(rv, H[0, 0] * v + H[0, 1] * u) and np.array_equal(ru, H[1, 0] * v + H[1, 1] * u)
    ru, rv = oracle.rotm(u, v, [0.0, 99.0, -1.0, 3.0, 99.0])
    assert np.array_equal(rv, v + 3.0 * u) and np.array_equal(ru, -v + u)
    ru, rv = oracle.rotm(u, v, [1.0, 2.0, 99.0, 99.0, 5.0])
    assert np.array_equal(rv, 2.0 * v + u) and np.array_equal(ru, 5.0 * u - v)
    ru, rv = oracle.rotm(u, v, [-2.0, 2.0, 2.0, 2.0, 2.0])
    assert np.array_equal(ru, u) and np.array_equal(rv, v)
    # Rotg: r = +-sqrt(a^2 + b^2) with the sign of the larger entry, c = a/r,
    # s = b/r; z = s when |a| > |b|, else 1/c (1 when c = 0)
    r, z, c, s = oracle.rotg(3.0, 4.0)
    assert abs(r - 5.0) < 1e-15 and abs(c - 0.6) < 1e-15 and abs(s - 0.8) < 1e-15 and abs(z - 1 / c) < 1e-14
    r, z, c, s = oracle.rotg(-4.0, 3.0)
    assert abs(r + 5.0) < 1e-15 and abs(c - 0.8) < 1e-15 and abs(s + 0.6) < 1e-15 and z == s
    assert oracle.rotg(0.0, 0.0) == (0.0, 0.0, 1.0, 0.0)
    r, z, c, s = oracle.rotg(0.0, -2.0)
    assert r == -2.0 and c == 0.0 and s == 1.0 and z == 1.0


def test_matrix_value_ops(oracle, pkg):
    A, _ = fv_matrix(pkg, oracle, 5, 4, 3, sigma=0.7, seed=2)
    D = A.to_dense()
    rows = [0, 7, 7, 59]
    V = oracle.zero_rows(A, rows, 2.5)
    Z = oracle.Csr(A.row, A.col, V).to_dense()
    ref = D.copy()
    ref[rows, :] = 0
    ref[rows, rows] = 2.5
    assert np.array_equal(Z, ref)
    V = oracle.zero_cols(A, [3, 11], -1.0)
    ref = D.copy()
    ref[:, [3, 11]] = np.where(D[:, [3, 11]] != 0, 0.0, 0.0)
    ref[3, 3] = ref[11, 11] = -1.0
    assert np.array_equal(oracle.Csr(A.row, A.col, V).to_dense(), ref)
    assert np.array_equal(oracle.Csr(A.row, A.col, oracle.mat_scale(A, 0.5)).to_dense(), D * 0.5)
    off = D * 3.0
    np.fill_diagonal(off, np.diag(D))
    assert np.array_equal(oracle.Csr(A.row, A.col, oracle.mat_scale(A, 3.0, True)).to_dense(), off)
    x = np.random.default_rng(2).normal(size=A.n)
    assert np.array_equal(oracle.pjacobi(A, x), x / np.diag(D))


def test_threaded_rows_same_bits(oracle, pkg):
    """The oracle's row-parallel loops (SpMV, element-wise updates, the rows
    of one colour in the multicolour sweeps) give the single-thread bits: the
    CG + MC-SGS and GMRES + ILU(0) solves that the benchmark times on all host
    cores are the reference-order solves."""
    row, col, val, rhs = pkg.problem(pkg.FV_DIFFUSION, 20, p0=1.0, seed=5)
    nc, color, sizes, perm = pkg.multicoloring(row, col)
    prow, pcol, pval = pkg.csr_permute(row, col, val, perm)
    A = oracle.Csr(prow, pcol, pval)
    b = rhs[perm]
    urow, ucol, uval, ub = pkg.problem(pkg.UPWIND_CONVDIFF, 14, p0=1e-2, p1=5.0, seed=2)
    U = oracle.Csr(urow, ucol, uval)
    LU, miss = oracle.ilu0(U)
    assert miss == 0
    runs = []
    try:
        for t in (1, 4):
            oracle.set_threads(t)
            assert oracle.get_threads() == t
            x1, i1 = oracle.cg(A, b, P=oracle.Prec(oracle.PREC_MCSGS, A, sizes=sizes),
                               ctrl=(2000, 0.0, 1e-10, 1e6))
            x2, i2 = oracle.gmres(U, ub, P=oracle.Prec(oracle.PREC_ILU0, LU), m=30, ctrl=(2000, 0.0, 1e-8, 1e6))
            y = oracle.spmv(U, ub)
            runs.append((x1, i1, x2, i2, y))
    finally:
        oracle.set_threads(1)
    (a1, ai1, a2, ai2, ay), (b1, bi1, b2, bi2, by) = runs
    assert ai1["iters"] == bi1["iters"] and np.array_equal(ai1["hist"], bi1["hist"]) and np.array_equal(a1, b1)
    assert ai2["iters"] == bi2["iters"] and np.array_equal(ai2["hist"], bi2["hist"]) and np.array_equal(a2, b2)
    assert np.array_equal(ay, by)
    assert ai1["iters"] > 20 and ai2["iters"] > 5
