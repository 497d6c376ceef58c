"""Host layer of the product (problem generators, colouring, level schedules,
ILU(0) factorisation, permutation) against the oracle, closed forms and the
reference's set_C fixture.  CPU only."""
import os

import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from conftest import relnorm

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def csr(pkg, kind, *dims, **kw):
    row, col, val, rhs = pkg.problem(kind, *dims, **kw)
    return row, col, val, rhs


def test_problem_sizes_closed_form(pkg):
    # nnz formulas of the benchmark configurations (evaluated by counting rows)
    import ctypes as C
    L = pkg.hfh.lmp()
    n, nnz = C.c_int(), C.c_longlong()

    def size(kind, nx, ny, nz):
        assert L.hfl_problem_size(kind, nx, ny, nz, 1.0, 0.0, 0, C.byref(n), C.byref(nnz)) == 0
        return n.value, nnz.value
    assert size(pkg.Q1_2D, 255, 255, 1) == (65025, (3 * 255 - 2) ** 2)            # 582,169
    assert size(pkg.Q1_3D, 127, 127, 127) == (2048383, (3 * 127 - 2) ** 3)        # 54,439,939
    assert size(pkg.FV_DIFFUSION, 192, 192, 192) == (192 ** 3, 7 * 192 ** 3 - 6 * 192 ** 2)
    assert size(pkg.UPWIND_CONVDIFF, 128, 128, 128) == (128 ** 3, 7 * 128 ** 3 - 6 * 128 ** 2)
    assert size(pkg.FV_DIFFUSION, 100, 100, 1) == (10000, 5 * 10000 - 4 * 100)


def test_q1_stencils(pkg):
    row, col, val, _ = csr(pkg, pkg.Q1_2D, 7, 6)
    S = sp.csr_matrix((val, col, row))
    d = S.diagonal()
    assert np.allclose(d, 8.0 / 3.0) and np.allclose(val[val != d.repeat(np.diff(row))], -1.0 / 3.0)
    assert abs(S - S.T).max() == 0
    # a row with all 8 neighbours sums to zero
    full = np.diff(row) == 9
    assert np.allclose(np.asarray(S.sum(axis=1)).ravel()[full], 0.0, atol=1e-15)
    nx = 6
    row, col, val, _ = csr(pkg, pkg.Q1_3D, nx)
    h = 1.0 / (nx + 1)
    S = sp.csr_matrix((val, col, row))
    assert np.allclose(S.diagonal(), 8 * h / 3)
    assert set(np.round(val / h, 12)) == {round(8 / 3, 12), 0.0, round(-1 / 6, 12), round(-1 / 12, 12)}
    full = np.diff(row) == 27
    assert np.allclose(np.asarray(S.sum(axis=1)).ravel()[full], 0.0, atol=1e-15)
    assert abs(S - S.T).max() == 0
    for i in range(len(row) - 1):
        assert np.all(np.diff(col[row[i]:row[i + 1]]) > 0)


def test_fv_harmonic_coefficients(pkg):
    nx, ny, nz, sigma, seed = 5, 4, 3, 1.0, 11
    row, col, val, _ = csr(pkg, pkg.FV_DIFFUSION, nx, ny, nz, p0=sigma, seed=seed)
    k = np.exp(sigma * np.array([pkg.hfh.lmp().hfl_normal01(seed, 2, c) for c in range(nx * ny * nz)]))
    S = sp.csr_matrix((val, col, row)).toarray()
    idx = lambda i, j, kk: i + nx * (j + ny * kk)  # noqa: E731
    for (i, j, kk) in [(0, 0, 0), (2, 1, 1), (4, 3, 2)]:
        c = idx(i, j, kk)
        diag = 0.0
        for di, dj, dk in [(-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)]:
            ii, jj, k2 = i + di, j + dj, kk + dk
            if 0 <= ii < nx and 0 <= jj < ny and 0 <= k2 < nz:
                t = 2 * k[c] * k[idx(ii, jj, k2)] / (k[c] + k[idx(ii, jj, k2)])
                assert S[c, idx(ii, jj, k2)] == pytest.approx(-t, rel=1e-15)
                diag += t
            else:
                diag += 2 * k[c]
        assert S[c, c] == pytest.approx(diag, rel=1e-14)
    assert np.abs(S - S.T).max() == 0
    # strictly diagonally dominant rows on the boundary, weakly inside
    off = np.abs(S).sum(axis=1) - np.abs(np.diag(S))
    assert np.all(np.diag(S) >= off - 1e-12)


@pytest.mark.parametrize("dims", [(7, 5, 1), (6, 5, 4)])
def test_fv_cells_give_the_generator_matrix(pkg, dims):
    """hfl_fv_cells returns the cell coefficients and rhs behind the FV
    generator: its harmonic-mean stencil, in the generator's order, equals the
    generated CSR values bit for bit (the serving engine assembles from it)."""
    nx, ny, nz = dims
    k, b = pkg.fv_cells(nx, ny, nz, 0.5, 21)
    row, col, val, rhs = csr(pkg, pkg.FV_DIFFUSION, nx, ny, nz, p0=0.5, seed=21)
    assert np.array_equal(b, rhs)
    for c in range(nx * ny * nz):
        i, j, kk = c % nx, (c // nx) % ny, c // (nx * ny)
        nb = [(c - nx * ny, nz > 1, nz > 1 and kk > 0), (c - nx, True, j > 0), (c - 1, True, i > 0),
              (c + 1, True, i < nx - 1), (c + nx, True, j < ny - 1), (c + nx * ny, nz > 1, nz > 1 and kk < nz - 1)]
        diag, off = 0.0, {}
        for q, (cn, axis, has) in enumerate(nb):
            if not axis:
                continue
            if has:
                t = 2.0 * k[c] * k[cn] / (k[c] + k[cn])
                off[cn] = -t
                diag += t
            else:
                diag += 2.0 * k[c]
        off[c] = diag
        cols = col[row[c]:row[c + 1]]
        assert list(cols) == sorted(off)
        assert [val[row[c] + e] for e in range(len(cols))] == [off[cc] for cc in cols]


def test_upwind_is_m_matrix(pkg):
    row, col, val, _ = csr(pkg, pkg.UPWIND_CONVDIFF, 8, p0=1e-2, p1=5.0)
    S = sp.csr_matrix((val, col, row)).toarray()
    D = np.diag(S)
    assert np.all(D > 0)
    off = S - np.diag(D)
    assert np.all(off <= 0)
    assert np.all(D >= np.abs(off).sum(axis=1) - 1e-12)
    assert np.abs(S - S.T).max() > 0.01  # convection makes it nonsymmetric
    # cell Peclet 5 at |b0| = sqrt(0.59): the largest coupling is eps + |b_x| h
    # with |b_x| largest at the cell centre nearest y = 0 or 1 (|y - 1/2| = 1/2 - h/2)
    h = 1.0 / 8
    assert (-off).max() / 1e-2 == pytest.approx(1 + 5.0 * (0.5 - h / 2) / np.sqrt(0.59), rel=1e-12)


def test_generator_matches_set_C(pkg):
    # set_C (test/PETSc/data/set_C) is -0.01 lap u = 1 on a 64x64-cell Q1 mesh
    # with the Dirichlet rows kept (diag 1) and FE cell-wise numbering.  The
    # interior rows carry the 8/3, -1/3 stencil times 0.01, and the interior
    # solution is a permutation of the solution of the generated system.
    d = np.load(os.path.join(GOLD, "set_C.npz"))
    rows, cols, vals = d["rows"], d["cols"], d["vals"]
    diag = vals[rows == cols]
    interior = np.where(np.abs(diag - 0.01 * 8 / 3) < 1e-6)[0]
    assert len(interior) == 63 * 63
    m = np.isin(rows, interior) & (rows != cols) & (vals != 0)
    assert np.allclose(vals[m], -0.01 / 3, rtol=1e-4)
    row, col, val, _ = csr(pkg, pkg.Q1_2D, 63, 63, p0=0.01)
    S = sp.csr_matrix((val, col, row))
    h2 = 1.0 / 64 ** 2
    u = spla.spsolve(S.tocsc(), np.full(S.shape[0], h2))
    assert np.allclose(np.sort(u), np.sort(d["sol"][interior]), rtol=2e-6, atol=1e-7)


def test_multicoloring_matches_oracle(pkg, oracle):
    for kind, dims in [(pkg.FV_DIFFUSION, (7, 6, 5)), (pkg.Q1_3D, (6, 5, 4)), (pkg.Q1_2D, (9, 8)),
                       (pkg.UPWIND_CONVDIFF, (6,))]:
        row, col, val, _ = csr(pkg, kind, *dims)
        nc, color, sizes, perm = pkg.multicoloring(row, col)
        onc, ocolor, osizes, operm = oracle.multicoloring(oracle.Csr(row, col, val))
        assert nc == onc and np.array_equal(color, ocolor) and np.array_equal(sizes, osizes)
        assert np.array_equal(perm, operm)


def test_levels_match_oracle(pkg, oracle):
    for kind, dims in [(pkg.FV_DIFFUSION, (7, 6, 5)), (pkg.UPWIND_CONVDIFF, (6,)), (pkg.Q1_2D, (9, 8))]:
        row, col, val, _ = csr(pkg, kind, *dims)
        ptr, rows = pkg.levels(row, col)
        nl, rep, sizes, perm = oracle.levelscheduling(oracle.Csr(row, col, val))
        assert len(ptr) - 1 == nl
        assert np.array_equal(np.diff(ptr), sizes) and np.array_equal(rows, perm)
        # backward levels: every upper neighbour sits in an earlier level
        uptr, urows = pkg.levels(row, col, upper=True)
        lev = np.empty(len(row) - 1, np.int64)
        for l in range(len(uptr) - 1):
            lev[urows[uptr[l]:uptr[l + 1]]] = l
        for i in range(len(row) - 1):
            c = col[row[i]:row[i + 1]]
            assert np.all(lev[c[c > i]] < lev[i])


def test_ilu0_matches_oracle_bitwise(pkg, oracle):
    row, col, val, _ = csr(pkg, pkg.UPWIND_CONVDIFF, 7, p0=1e-2, p1=5.0, seed=3)
    lu = pkg.ilu0_factor(row, col, val)
    LU, miss = oracle.ilu0(oracle.Csr(row, col, val))
    assert miss == 0 and np.array_equal(lu, LU.val)


def test_permute_matches_oracle(pkg, oracle):
    row, col, val, _ = csr(pkg, pkg.FV_DIFFUSION, 6, 5, 4, seed=2)
    perm = np.random.default_rng(0).permutation(len(row) - 1).astype(np.int32)
    prow, pcol, pval = pkg.csr_permute(row, col, val, perm)
    B = oracle.csr_reorder(oracle.Csr(row, col, val), perm)
    assert np.array_equal(prow, B.row) and np.array_equal(pcol, B.col) and np.array_equal(pval, B.val)


def test_rhs_is_seeded_uniform(pkg):
    _, _, _, b1 = csr(pkg, pkg.FV_DIFFUSION, 20, 20, 1, seed=5)
    _, _, _, b2 = csr(pkg, pkg.FV_DIFFUSION, 20, 20, 1, seed=5)
    _, _, _, b3 = csr(pkg, pkg.FV_DIFFUSION, 20, 20, 1, seed=6)
    assert np.array_equal(b1, b2) and not np.array_equal(b1, b3)
    assert b1.min() >= -1 and b1.max() < 1 and abs(b1.mean()) < 0.1
    assert relnorm(b1.std(), 1 / np.sqrt(3)) < 0.1


# --------------------------------------------------------------- MatrixMarket and vector files

MM_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mm")


def _scipy_csr(path):
    import scipy.io
    S = scipy.io.mmread(path).tocsr()
    S.sort_indices()
    return S


@pytest.mark.parametrize("name", ["Equation1.A.txt", "A3.mtx", "A4.mtx"])
def test_mm_read_reference_files(pkg, name):
    """The reference's own MatrixMarket test files read into sorted CSR; scipy's
    reader gives the same matrix."""
    path = os.path.join(MM_DIR, name)
    row, col, val, ncol = pkg.mm_read(path)
    S = _scipy_csr(path)
    assert ncol == S.shape[1] and len(row) - 1 == S.shape[0]
    assert np.array_equal(row, S.indptr) and np.array_equal(col, S.indices) and np.array_equal(val, S.data)
    if name == "Equation1.A.txt":
        d = np.load(os.path.join(os.path.dirname(MM_DIR), "equation1.npz"))
        assert np.array_equal(sp.csr_matrix((val, col, row)).toarray(),
                              sp.coo_matrix((d["vals"], (d["rows"], d["cols"]))).toarray())


def test_mm_symmetric_and_roundtrip(pkg, tmp_path):
    p = tmp_path / "sym.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real symmetric\n% comment\n4 4 5\n1 1 4.0\n2 1 -1.5\n"
                 "3 2 -2.0\n3 3 5.0\n4 1 0.25\n")
    row, col, val, ncol = pkg.mm_read(str(p))
    S = _scipy_csr(str(p))
    assert np.array_equal(row, S.indptr) and np.array_equal(col, S.indices) and np.array_equal(val, S.data)
    assert len(val) == 8                                      # three mirrored entries
    # write -> read of the set_C system keeps every entry (its values have 15 digits)
    d = np.load(os.path.join(os.path.dirname(MM_DIR), "set_C.npz"))
    A = sp.coo_matrix((d["vals"], (d["rows"], d["cols"]))).tocsr()
    A.sort_indices()
    q = tmp_path / "setC.mtx"
    pkg.mm_write(str(q), A.indptr, A.indices, A.data)
    r2, c2, v2, n2 = pkg.mm_read(str(q))
    assert np.array_equal(r2, A.indptr) and np.array_equal(c2, A.indices) and np.array_equal(v2, A.data)
    assert q.read_text().splitlines()[2].split() == ["1", "1", "1.00000000000000e+00"]


def test_mm_rejects_unsupported(pkg, tmp_path):
    for banner in ("%%MatrixMarket matrix array real general", "%%MatrixMarket matrix coordinate complex general",
                   "%%MatrixMarket matrix coordinate pattern general", "%%MatrixMarket vector coordinate real general"):
        p = tmp_path / "bad.mtx"
        p.write_text(banner + "\n2 2 1\n1 1 1.0\n")
        with pytest.raises(pkg.HfhError):
            pkg.mm_read(str(p))
    p = tmp_path / "short.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real general\n2 2 3\n1 1 1.0\n")
    with pytest.raises(pkg.HfhError):
        pkg.mm_read(str(p))


def test_vector_files(pkg, tmp_path):
    """CPU_lVector::WriteFile writes std::scientific with 6 digits; ReadFile reads
    whitespace-separated values."""
    p = tmp_path / "v.txt"
    pkg.vec_write(str(p), [1.0, -2.5e-3, 123456.789, 0.0])
    assert p.read_text() == "1.000000e+00\n-2.500000e-03\n1.234568e+05\n0.000000e+00\n"
    assert list(pkg.vec_read(str(p))) == [1.0, -2.5e-3, 123456.8, 0.0]
