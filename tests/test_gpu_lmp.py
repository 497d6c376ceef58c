"""The C++ LinearSolver templates (CG / GMRES / FGMRES / MPIR over single-rank
CoupledVector / CoupledMatrix of HIP lVectors) against the oracle: the same
iteration, so the same iteration count (+-2) and residual history."""
import numpy as np
import pytest

from conftest import relnorm

pytestmark = pytest.mark.gpu


def true_res(row, col, val, x, b):
    import scipy.sparse as sp
    return np.linalg.norm(b - sp.csr_matrix((val, col, row)) @ x) / np.linalg.norm(b)


@pytest.mark.parametrize("prec", ["none", "jacobi", "sgs", "ilu0", "mcsgs"])
def test_lmp_cg(pkg, oracle, prec):
    row, col, val, rhs = pkg.problem(pkg.FV_DIFFUSION, 12, p0=1.0, seed=2)
    code = {"none": pkg.PREC_NONE, "jacobi": pkg.PREC_JACOBI, "sgs": pkg.PREC_SGS, "ilu0": pkg.PREC_ILU0,
            "mcsgs": pkg.PREC_MCSGS}[prec]
    x, info = pkg.lmp_solve(row, col, val, rhs, solver=pkg.SOLVER_CG, precond=code, maxits=2000, atol=0.0,
                            rtol=1e-10)
    assert info["ret"] == 0
    OA = oracle.Csr(row, col, val)
    if prec == "mcsgs":
        nc, color, sizes, perm = oracle.multicoloring(OA)
        Ap = oracle.csr_reorder(OA, perm)
        xo_p, oi = oracle.cg(Ap, rhs[perm], P=oracle.Prec(oracle.PREC_MCSGS, Ap, sizes=sizes),
                             ctrl=(2000, 0.0, 1e-10, 1e6))
        xo = np.empty_like(xo_p)
        xo[perm] = xo_p
    else:
        P = {"none": None, "jacobi": oracle.Prec.jacobi(OA), "sgs": oracle.Prec(oracle.PREC_SGS, OA),
             "ilu0": oracle.Prec(oracle.PREC_ILU0, oracle.ilu0(OA)[0])}[prec]
        xo, oi = oracle.cg(OA, rhs, P=P, ctrl=(2000, 0.0, 1e-10, 1e6))
    assert abs(info["iters"] - oi["iters"]) <= 2
    k = min(15, len(info["hist"]), len(oi["hist"]))
    assert np.allclose(info["hist"][:k], oi["hist"][:k], rtol=1e-9)
    assert true_res(row, col, val, x, rhs) < 3e-10
    assert relnorm(x, xo) < 1e-8


@pytest.mark.parametrize("solver", ["gmres", "fgmres"])
def test_lmp_gmres_ilu0(pkg, oracle, solver):
    row, col, val, rhs = pkg.problem(pkg.UPWIND_CONVDIFF, 10, p0=1e-2, p1=5.0, seed=3)
    code = pkg.SOLVER_GMRES if solver == "gmres" else pkg.SOLVER_FGMRES
    x, info = pkg.lmp_solve(row, col, val, rhs, solver=code, precond=pkg.PREC_ILU0, m=10, maxits=500, atol=0.0,
                            rtol=1e-8)
    OA = oracle.Csr(row, col, val)
    fn = oracle.gmres if solver == "gmres" else oracle.fgmres
    xo, oi = fn(OA, rhs, P=oracle.Prec(oracle.PREC_ILU0, oracle.ilu0(OA)[0]), m=10, ctrl=(500, 0.0, 1e-8, 1e6))
    assert info["ret"] == 0 and abs(info["iters"] - oi["iters"]) <= 2
    k = min(10, len(info["hist"]), len(oi["hist"]))
    assert np.allclose(info["hist"][:k], oi["hist"][:k], rtol=1e-8)
    assert true_res(row, col, val, x, rhs) < 2e-8


def test_lmp_cg_float_descriptor(pkg, oracle):
    row, col, val, rhs = pkg.problem(pkg.FV_DIFFUSION, 10, p0=0.5, seed=4)
    x, info = pkg.lmp_solve(row, col, val, rhs, solver=pkg.SOLVER_CG, precond=pkg.PREC_JACOBI, single=True,
                            maxits=1000, atol=0.0, rtol=1e-5)
    OA = oracle.Csr(row, col, val.astype(np.float32))
    xo, oi = oracle.cg(OA, rhs.astype(np.float32), P=oracle.Prec.jacobi(OA), ctrl=(1000, 0.0, 1e-5, 1e6))
    assert abs(info["iters"] - oi["iters"]) <= 2
    assert relnorm(x, xo) < 1e-4


@pytest.mark.parametrize("solver", ["gmres", "fgmres"])
def test_lmp_gmres_no_preconditioning(pkg, oracle, solver):
    # GMRES::SolveNoPrecond (gmres.cc:103-264), which FGMRES also runs
    # without a preconditioner; restarts of basis 10
    row, col, val, rhs = pkg.problem(pkg.UPWIND_CONVDIFF, 8, p0=1e-2, p1=5.0, seed=4)
    code = pkg.SOLVER_GMRES if solver == "gmres" else pkg.SOLVER_FGMRES
    x, info = pkg.lmp_solve(row, col, val, rhs, solver=code, precond=pkg.PREC_NONE, m=10, maxits=3000, atol=0.0,
                            rtol=1e-8)
    OA = oracle.Csr(row, col, val)
    fn = oracle.gmres if solver == "gmres" else oracle.fgmres
    xo, oi = fn(OA, rhs, m=10, ctrl=(3000, 0.0, 1e-8, 1e6))
    assert info["ret"] == 0 and oi["ret"] == 0 and info["iters"] > 10
    assert abs(info["iters"] - oi["iters"]) <= 2
    k = min(10, len(info["hist"]), len(oi["hist"]))
    assert np.allclose(info["hist"][:k], oi["hist"][:k], rtol=1e-8)
    assert true_res(row, col, val, x, rhs) < 2e-8
    assert relnorm(x, xo) < 1e-6


def test_lmp_mpir(pkg, oracle):
    row, col, val, rhs = pkg.problem(pkg.FV_DIFFUSION, 12, p0=1.0, seed=5)
    nb = np.linalg.norm(rhs)
    x, info = pkg.lmp_mpir(row, col, val, rhs, precond_low=pkg.PREC_JACOBI, max_ecl=30, abs_tol=1e-10 * nb,
                           rel_tol=1e-12, ecl=(1000, 0.0, 1e-3, 1e6))
    OA = oracle.Csr(row, col, val)
    xo, oi = oracle.mpir(OA, rhs, Pl=oracle.Prec.jacobi(OA), max_ecl=30, abs_tol=1e-10 * nb, rel_tol=1e-12,
                         ecl=(1000, 0.0, 1e-3, 1e6))
    assert info["ret"] == 0 and abs(info["n_ecl"] - oi["n_ecl"]) <= 1
    assert true_res(row, col, val, x, rhs) < 1e-10
    assert relnorm(x, xo) < 1e-8


def test_lmp_mpir_res0_reset(pkg, oracle):
    """The template keeps the reference's res0_ reset (mpir.cc:275): with
    rel_tol 0.5 and no reachable abs_tol, exactly max_ecl corrections run and
    Solve returns kSolverExceeded, as the oracle's MPIR."""
    row, col, val, rhs = pkg.problem(pkg.FV_DIFFUSION, 10, p0=0.5, seed=3)
    OA = oracle.Csr(row, col, val)
    for k in (1, 3):
        x, info = pkg.lmp_mpir(row, col, val, rhs, precond_low=pkg.PREC_JACOBI, max_ecl=k, abs_tol=0.0,
                               rel_tol=0.5, ecl=(1000, 0.0, 0.125, 1e6))
        xo, oi = oracle.mpir(OA, rhs, Pl=oracle.Prec.jacobi(OA), max_ecl=k, abs_tol=0.0, rel_tol=0.5,
                             ecl=(1000, 0.0, 0.125, 1e6))
        assert info["n_ecl"] == oi["n_ecl"] == k and info["ret"] == oi["ret"] == 1
        assert info["hist"][1] < 0.5 * info["hist"][0]
        assert relnorm(info["hist"], oi["hist"]) < 1e-5


def test_lmp_errors(pkg):
    row, col, val, rhs = pkg.problem(pkg.FV_DIFFUSION, 4)
    with pytest.raises(pkg.HfhError):
        pkg.lmp_solve(row, col, val, rhs, solver=7)
    val2 = val.copy()
    val2[row[3]:row[4]][col[row[3]:row[4]] == 3] = 0.0  # zero diagonal in row 3
    with pytest.raises(pkg.HfhError) as e:
        pkg.lmp_solve(row, col, val2, rhs, precond=pkg.PREC_JACOBI)
    assert e.value.code == pkg.ERR_ZERO_PIVOT
