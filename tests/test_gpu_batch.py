"""Serving engine (hfh_batch_*): FP32-inner / FP64-outer CG for many systems
in shared batched launches.

Each result is checked against the oracle (its MPIR<double, float> restatement
with an FP32 Jacobi-CG inner solver, the same algorithm, and an FP64 CG solve
to 1e-14), and each system's result must be bit-identical whether it was
solved alone or admitted into a running mix of other systems."""
import numpy as np
import pytest

from conftest import relnorm

pytestmark = pytest.mark.gpu

RTOL = 1e-10
INNER_RTOL = 1e-4


def engine(pkg, ctx, **kw):
    args = dict(max_systems=64, hbm_budget=256 << 20, chunk=8, rtol=RTOL, atol=0.0, outer_maxits=40,
                inner_rtol=INNER_RTOL, inner_maxits=4000)
    args.update(kw)
    return pkg.BatchEngine(ctx, **args)


def fv(pkg, nx, ny, nz, seed, sigma=0.5):
    k, b = pkg.fv_cells(nx, ny, nz, sigma, seed)
    row, col, val, rhs = pkg.problem(pkg.FV_DIFFUSION, nx, ny, nz, p0=sigma, seed=seed)
    assert np.array_equal(rhs, b)
    return k, b, row, col, val


def run_to_end(E, max_steps=100000):
    steps = 0
    while True:
        active, _ = E.step()
        steps += 1
        if active == 0:
            return steps
        assert steps < max_steps


CASES = [(40, 40, 1, 11), (12, 12, 12, 12), (100, 60, 1, 13), (20, 24, 18, 14)]


@pytest.mark.parametrize("nx,ny,nz,seed", CASES)
def test_batch_single_vs_oracle(pkg, oracle, ctx, nx, ny, nz, seed):
    k, b, row, col, val = fv(pkg, nx, ny, nz, seed)
    n = len(b)
    E = engine(pkg, ctx)
    rid = E.submit_fv(nx, ny, nz, k, b)
    assert rid is not None
    run_to_end(E)
    assert E.finished() == [rid]
    x, info = E.collect(rid, n)
    E.close()
    assert info["rc"] == pkg.OK and info["state"] == pkg.SUCCESS_REL
    A = oracle.Csr(row, col, val)
    # true residual, recomputed by the oracle SpMV, meets the stated tolerance
    res = np.linalg.norm(b - oracle.spmv(A, x))
    assert res <= RTOL * np.linalg.norm(b)
    assert abs(info["res"] - res) <= 1e-12 * res + 1e-300
    assert info["res0"] == pytest.approx(np.linalg.norm(b), rel=1e-13)
    # the oracle's MPIR with an FP32 Jacobi-CG inner solver: same outer loop
    xm, om = oracle.mpir(A, b, Pl=oracle.Prec.jacobi(A), max_ecl=40, abs_tol=RTOL * np.linalg.norm(b),
                         rel_tol=0.0, ecl=(4000, 0.0, INNER_RTOL, 1e6))
    assert om["ret"] == 0
    assert abs(info["outer_iters"] - om["n_ecl"]) <= 1
    assert abs(info["inner_iters"] - om["iters"]) <= 0.1 * om["iters"] + 5
    # against an FP64 solve to 1e-14: error bounded by cond * tolerance
    xr, oi = oracle.cg(A, b, P=oracle.Prec.jacobi(A), ctrl=(20000, 0.0, 1e-14, 1e6))
    assert relnorm(x, xr) < 1e-6
    assert relnorm(xm, xr) < 1e-6


def test_batch_fv_assembly_matches_host_csr(pkg, ctx):
    """Device-assembled FV systems give bit-identical results to the host CSR."""
    for nx, ny, nz, seed in CASES[:2]:
        k, b, row, col, val = fv(pkg, nx, ny, nz, seed)
        E = engine(pkg, ctx)
        r1 = E.submit_fv(nx, ny, nz, k, b)
        r2 = E.submit_csr(row, col, val, b)
        run_to_end(E)
        x1, i1 = E.collect(r1, len(b))
        x2, i2 = E.collect(r2, len(b))
        E.close()
        assert np.array_equal(x1, x2)
        assert (i1["outer_iters"], i1["inner_iters"]) == (i2["outer_iters"], i2["inner_iters"])


def test_batch_alone_equals_mixed(pkg, ctx):
    """Continuous batching: systems admitted while others run; every result is
    bit-identical to the same system solved alone."""
    specs = [(30, 30, 1, 21), (10, 10, 10, 22), (70, 50, 1, 23), (16, 16, 16, 24), (9, 7, 1, 25),
             (120, 100, 1, 26), (13, 11, 9, 27)]
    data = {s: fv(pkg, *s) for s in specs}
    alone = {}
    for s in specs:
        k, b = data[s][0], data[s][1]
        E = engine(pkg, ctx)
        rid = E.submit_fv(s[0], s[1], s[2], k, b)
        run_to_end(E)
        alone[s] = E.collect(rid, len(b))
        E.close()
    E = engine(pkg, ctx, max_systems=4)
    pending = list(specs)
    ids = {}
    results = {}
    steps = 0
    while pending or ids:
        # admit as slots free up, one per step, so arrivals overlap running work
        if pending:
            s = pending[0]
            k, b = data[s][0], data[s][1]
            rid = E.submit_fv(s[0], s[1], s[2], k, b)
            if rid is not None:
                ids[rid] = s
                pending.pop(0)
        E.step()
        steps += 1
        for rid in E.finished():
            s = ids.pop(rid)
            results[s] = E.collect(rid, len(data[s][1]))
        assert steps < 100000
    replays, busy_ms, row_iters = E.stats()
    E.close()
    for s in specs:
        xa, ia = alone[s]
        xm, im = results[s]
        assert np.array_equal(xa, xm), s
        assert ia == im, s
        assert ia["rc"] == pkg.OK
    # engine statistics: every replay read back is counted (the last one, run
    # after the final system finished, is not read back), and the
    # row-iterations are the sum of n x inner iterations of the systems
    assert steps - 1 <= replays <= steps and busy_ms > 0
    assert row_iters == sum(s[0] * s[1] * s[2] * results[s][1]["inner_iters"] for s in specs)


def test_batch_budget_admission(pkg, ctx):
    """Admission by the HBM budget: a request that does not fit now is refused
    (None) and admitted after a finished system is collected; a request larger
    than the whole budget is an error."""
    nx, ny, nz = 60, 60, 1
    need = pkg.BatchEngine.bytes_fv(nx, ny, nz)
    E = engine(pkg, ctx, hbm_budget=int(need * 2.5))
    ks = [fv(pkg, nx, ny, nz, 40 + i) for i in range(3)]
    r0 = E.submit_fv(nx, ny, nz, ks[0][0], ks[0][1])
    r1 = E.submit_fv(nx, ny, nz, ks[1][0], ks[1][1])
    assert r0 is not None and r1 is not None
    assert E.submit_fv(nx, ny, nz, ks[2][0], ks[2][1]) is None
    active, used, cap = E.usage()
    assert active == 2 and used == 2 * need and cap == int(need * 2.5)
    run_to_end(E)
    E.collect(r0, nx * ny)
    r2 = E.submit_fv(nx, ny, nz, ks[2][0], ks[2][1])
    assert r2 is not None
    run_to_end(E)
    for rid in (r1, r2):
        x, info = E.collect(rid, nx * ny)
        assert info["rc"] == pkg.OK
    big = pkg.fv_cells(200, 200, 1, 0.5, 1)
    with pytest.raises(pkg.HfhError):
        E.submit_fv(200, 200, 1, big[0], big[1])
    assert E.usage()[:2] == (0, 0)
    E.close()


def test_batch_maxits_failure_is_reported(pkg, ctx):
    """A system stopped by its correction limit is reported as not converged;
    every kernel still exits normally (the limit is a host parameter)."""
    k, b, row, col, val = fv(pkg, 50, 50, 1, 31)
    E = engine(pkg, ctx, outer_maxits=1, inner_maxits=5)
    rid = E.submit_fv(50, 50, 1, k, b)
    run_to_end(E)
    x, info = E.collect(rid, len(b))
    E.close()
    assert info["rc"] == pkg.ERR_NOT_CONVERGED and info["state"] == pkg.FAIL_MAXITS
    assert info["outer_iters"] == 1 and info["inner_iters"] == 5
    assert info["res"] < info["res0"]


def test_batch_rejects_bad_requests(pkg, ctx):
    E = engine(pkg, ctx)
    row = np.array([0, 1, 2], np.int32)
    col = np.array([1, 0], np.int32)   # no diagonal
    with pytest.raises(pkg.HfhError):
        E.submit_csr(row, col, np.array([1.0, 1.0]), np.ones(2))
    with pytest.raises(pkg.HfhError):
        E.submit_fv(4, 4, 1, -np.ones(16), np.ones(16))
    E.close()


def test_small_systems_persistent_equals_tiles(pkg, ctx):
    """With small_tiles = 4, systems of at most 4 tiles (4,096 rows) run a
    step's iterations in one persistent workgroup each (kb_small) beside the
    tile launches of the larger ones.  Group g of the workgroup does what the
    tile path does for tile g, in the same order: the results equal the tile
    path's bit for bit, in a mix with larger systems."""
    specs = [(9, 7, 1, 31), (40, 40, 1, 32), (16, 16, 12, 33), (64, 32, 2, 35), (20, 20, 20, 34), (130, 120, 1, 36)]
    data = {s: fv(pkg, *s) for s in specs}
    tiles = [(a * b * c + 1023) // 1024 for a, b, c, _ in specs]
    assert tiles[:4] == [1, 2, 3, 4] and min(tiles[4:]) > 4
    res = {}
    for small_tiles in (4, 0):
        E = engine(pkg, ctx, small_tiles=small_tiles)
        ids = {}
        for s in specs:
            k, b = data[s][0], data[s][1]
            ids[E.submit_fv(s[0], s[1], s[2], k, b)] = s
        run_to_end(E)
        res[small_tiles] = {ids[r]: E.collect(r, len(data[ids[r]][1])) for r in list(ids)}
        E.close()
    for s in specs:
        xa, ia = res[4][s]
        xb, ib = res[0][s]
        print(f"small path case {s}: outer {ia['outer_iters']} inner {ia['inner_iters']}", flush=True)
        assert ia["state"] in (pkg.SUCCESS_ABS, pkg.SUCCESS_REL)
        assert np.array_equal(xa, xb), s
        assert ia == ib, s
