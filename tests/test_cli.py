"""The local solver benchmark's command line (examples/benchmark_local_solver/
benchmark_local_solver.cc:226-311): argument handling on the CPU; solves of
the reference's Equation1 files and the set_C system on the GPU."""
import importlib.util
import io
import os

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
MM = os.path.join(ROOT, "tests", "golden", "mm")
GOLD = os.path.join(ROOT, "tests", "golden")


def cli():
    spec = importlib.util.spec_from_file_location(
        "benchmark_local_solver", os.path.join(ROOT, "hiflow3_2.0_amd", "benchmark_local_solver.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def eq1(solver, interface):
    return ["benchmark_local_solver.py", os.path.join(MM, "Equation1.A.txt"), os.path.join(MM, "Equation1.rhs.txt"),
            os.path.join(MM, "Equation1.sol.txt"), os.path.join(MM, "Equation1.x0.txt"), solver, interface, "1", "2"]


def test_usage_and_unsupported_choices():
    m = cli()
    out = io.StringIO()
    assert m.run(["benchmark_local_solver.py"], out) == 1 and "Usage:" in out.getvalue()
    out = io.StringIO()
    assert m.run(eq1("bicgstab", "hip"), out) == 1 and "Solver bicgstab not supported." in out.getvalue()
    out = io.StringIO()
    assert m.run(eq1("cg", "cublas2"), out) == 1 and "Algo cublas2 not supported." in out.getvalue()


@pytest.mark.gpu
@pytest.mark.parametrize("solver", ["cg", "gmres"])
@pytest.mark.parametrize("interface", ["hip", "lmp"])
def test_equation1(solver, interface):
    out = io.StringIO()
    assert cli().run(eq1(solver, interface), out) == 0, out.getvalue()
    text = out.getvalue()
    dot = float(text.split("dot(x - x_result) = ")[1].split()[0])
    assert dot < 1e-20 and "WARNING" not in text and "Done successfully." in text


@pytest.mark.gpu
def test_set_C_through_the_command_line(pkg, tmp_path):
    import scipy.sparse as sp
    d = np.load(os.path.join(GOLD, "set_C.npz"))
    A = sp.coo_matrix((d["vals"], (d["rows"], d["cols"]))).tocsr()
    A.sort_indices()
    b, sol = d["rhs"], d["sol"]
    pkg.mm_write(str(tmp_path / "A.mtx"), A.indptr, A.indices, A.data)
    pkg.vec_write(str(tmp_path / "b.txt"), b)
    pkg.vec_write(str(tmp_path / "x0.txt"), np.zeros(len(b)))
    pkg.vec_write(str(tmp_path / "xr.txt"), sol)
    argv = ["benchmark_local_solver.py", str(tmp_path / "A.mtx"), str(tmp_path / "b.txt"), str(tmp_path / "xr.txt"),
            str(tmp_path / "x0.txt"), "cg", "hip", "1", "1"]
    out = io.StringIO()
    assert cli().run(argv, out) == 0, out.getvalue()
    assert "WARNING" not in out.getvalue()
