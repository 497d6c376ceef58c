"""HFH_GRAPHS=0: the drivers and the serving engine launch their iteration
chunks directly instead of replaying captured graphs.  The kernels and their
order are the same, so the results are bit-identical; a child process with the
variable set (it is read once per process) solves the same systems as this
one."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

SCRIPT = r"""
import json, sys
import numpy as np
sys.path.insert(0, "tests")
from conftest import load_pkg
P = load_pkg()
ctx = P.Context(0)
out = {}
# CG + Jacobi (graph of 16 iterations per chunk)
row, col, val, rhs = P.problem(P.Q1_3D, 21)
A = P.CsrMatrix(ctx, row, col, val)
X = P.Vector(ctx, len(rhs))
inf = P.pcg(A, P.Vector(ctx, rhs), X, P.Preconditioner.jacobi(A), maxits=2000, atol=0.0, rtol=1e-8, history=True)
out["cg_iters"] = inf["iters"]
out["cg_hist"] = inf["hist"].tolist()
out["cg_x"] = X.numpy().tolist()
# serving engine
E = P.BatchEngine(ctx, max_systems=8, hbm_budget=64 << 20, chunk=8, rtol=1e-10, atol=0.0, outer_maxits=40,
                  inner_rtol=1e-4, inner_maxits=4000)
k, b = P.fv_cells(17, 13, 5, 0.5, 3)
rid = E.submit_fv(17, 13, 5, k, b)
while E.step()[0]:
    pass
x, info = E.collect(rid, len(b))
out["batch_x"] = x.tolist()
out["batch_info"] = list(info) if not isinstance(info, dict) else info
E.close()
print("JSON" + json.dumps(out))
"""


def _run(env_graphs):
    env = dict(os.environ)
    if env_graphs is None:
        env.pop("HFH_GRAPHS", None)
    else:
        env["HFH_GRAPHS"] = env_graphs
    r = subprocess.run([sys.executable, "-c", SCRIPT], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=110)
    assert r.returncode == 0, r.stderr[-2000:]
    line = [l for l in r.stdout.splitlines() if l.startswith("JSON")][-1]
    return json.loads(line[4:])


def test_direct_launches_equal_graph_replays():
    print("nograph case: graphs on, then HFH_GRAPHS=0", flush=True)
    g = _run(None)
    d = _run("0")
    assert g["cg_iters"] == d["cg_iters"]
    assert g["cg_hist"] == d["cg_hist"]
    assert np.array_equal(np.array(g["cg_x"]), np.array(d["cg_x"]))
    assert np.array_equal(np.array(g["batch_x"]), np.array(d["batch_x"]))
    assert g["batch_info"] == d["batch_info"]
