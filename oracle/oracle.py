"""ctypes front end of the CPU oracle (oracle/libhfo.so).

TEST INFRASTRUCTURE ONLY: used by tests/, the smoke check and bench.py's
cpu_baseline leg as the checker.  The product path never imports it.

The C library restates the reference CPU path loop for loop; see hfo.h for
the reference file:line of each routine.
"""
import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "libhfo.so")

PREC_NONE, PREC_JACOBI, PREC_SGS, PREC_MCSGS, PREC_ILU0 = 0, 1, 2, 3, 4
ITERATE, SUCCESS_ABS, SUCCESS_REL, FAIL_DIV, FAIL_MAXITS = 0, 1, 2, 3, 4


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        _lib = C.CDLL(_LIB)
        _declare(_lib)
    return _lib


class Ctrl(C.Structure):
    _fields_ = [("maxits", C.c_int), ("atol", C.c_double), ("rtol", C.c_double),
                ("dtol", C.c_double)]


class Result(C.Structure):
    _fields_ = [("iters", C.c_int), ("status", C.c_int), ("ret", C.c_int),
                ("res", C.c_double), ("res0", C.c_double)]


class PrecD(C.Structure):
    _fields_ = [("kind", C.c_int), ("invd", C.c_void_p), ("row", C.c_void_p),
                ("col", C.c_void_p), ("val", C.c_void_p), ("ncolors", C.c_int),
                ("color_sizes", C.c_void_p), ("omega", C.c_double)]


class PrecF(C.Structure):
    _fields_ = [("kind", C.c_int), ("invd", C.c_void_p), ("row", C.c_void_p),
                ("col", C.c_void_p), ("val", C.c_void_p), ("ncolors", C.c_int),
                ("color_sizes", C.c_void_p), ("omega", C.c_float)]


class MpirParams(C.Structure):
    _fields_ = [("max_ech", C.c_int), ("max_ecl", C.c_int), ("abs_tol", C.c_double),
                ("rel_tol", C.c_double), ("div_tol", C.c_double), ("ecl", Ctrl),
                ("ech", Ctrl)]


class MpirResult(C.Structure):
    _fields_ = [("ret", C.c_int), ("iters", C.c_int), ("n_ecl", C.c_int),
                ("n_ech", C.c_int), ("res", C.c_double)]


def _declare(L):
    vp, i, d, f = C.c_void_p, C.c_int, C.c_double, C.c_float
    sig = {
        "hfo_csr_from_coo_d": (None, [i, i, vp, vp, vp, vp, vp, vp]),
        "hfo_csr_add_values_d": (None, [i, vp, vp, vp, vp, i, vp, i, vp]),
        "hfo_csr_reorder_d": (None, [i, vp, vp, vp, vp, vp, vp, vp]),
        "hfo_vec_reorder_d": (None, [i, vp, vp, vp]),
        "hfo_multicoloring": (i, [i, vp, vp, vp, vp, vp, vp]),
        "hfo_levelscheduling": (i, [i, vp, vp, vp, vp, vp]),
        "hfo_mc_dropoffs_d": (i, [i, vp, vp, vp, i, vp]),
        "hfo_cast_d2f": (None, [i, vp, vp]),
        "hfo_cast_f2d": (None, [i, vp, vp]),
        "hfo_mpir": (i, [i, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
        "hfo_set_threads": (None, [i]),
        "hfo_get_threads": (i, []),
    }
    for s, t in (("d", d), ("f", f)):
        sig.update({
            f"hfo_spmv_{s}": (None, [i, vp, vp, vp, vp, vp]),
            f"hfo_spmv_omp_{s}": (None, [i, vp, vp, vp, vp, vp]),
            f"hfo_spmv_add_{s}": (None, [i, vp, vp, vp, vp, vp]),
            f"hfo_dot_{s}": (t, [i, vp, vp]),
            f"hfo_nrm2_{s}": (t, [i, vp]),
            f"hfo_axpy_{s}": (None, [i, vp, vp, t]),
            f"hfo_scaleadd_{s}": (None, [i, vp, t, vp]),
            f"hfo_scale_{s}": (None, [i, vp, t]),
            f"hfo_ewmult_{s}": (None, [i, vp, vp]),
            f"hfo_invdiag_{s}": (None, [i, vp, vp, vp, vp]),
            f"hfo_sgs_{s}": (None, [i, vp, vp, vp, vp, vp]),
            f"hfo_ilu0_{s}": (i, [i, vp, vp, vp]),
            f"hfo_ilu_solve_{s}": (None, [i, vp, vp, vp, vp, vp]),
            f"hfo_mcsgs_{s}": (None, [i, vp, vp, vp, i, vp, t, vp, vp]),
            f"hfo_prec_apply_{s}": (None, [i, vp, vp, vp]),
            f"hfo_cg_{s}": (i, [i, vp, vp, vp, vp, vp, vp, vp, vp, vp]),
            f"hfo_gmres_{s}": (i, [i, vp, vp, vp, vp, vp, vp, i, vp, vp, vp]),
            f"hfo_fgmres_{s}": (i, [i, vp, vp, vp, vp, vp, vp, i, vp, vp, vp]),
            f"hfo_swap_diag_front_{s}": (None, [i, vp, vp, vp]),
            f"hfo_spmv_nodiag_{s}": (None, [i, vp, vp, vp, vp, vp]),
            f"hfo_jacobi_{s}": (i, [i, vp, vp, vp, vp, vp, i, t, i, i, vp, vp, vp]),
            f"hfo_norm1_{s}": (t, [i, vp]),
            f"hfo_argmax_{s}": (i, [i, vp]),
            f"hfo_argmin_{s}": (i, [i, vp]),
            f"hfo_add_values_{s}": (None, [vp, vp, i, vp]),
            f"hfo_add_range_{s}": (None, [vp, i, vp, i, i, t]),
            f"hfo_rot_{s}": (None, [i, vp, vp, t, t]),
            f"hfo_rotm_{s}": (None, [i, vp, vp, vp]),
            f"hfo_rotg_{s}": (None, [vp, vp, vp, vp]),
            f"hfo_zero_rows_{s}": (None, [vp, vp, vp, vp, i, t]),
            f"hfo_zero_cols_{s}": (None, [i, vp, vp, vp, vp, i, t]),
            f"hfo_mat_scale_{s}": (None, [i, vp, vp, vp, t, i]),
            f"hfo_pjacobi_{s}": (None, [i, vp, vp, vp, vp, vp]),
        })
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args


def set_threads(t):
    """Threads of the oracle's row-parallel loops (results do not depend on it)."""
    lib().hfo_set_threads(int(t))


def get_threads():
    return lib().hfo_get_threads()


def _p(a):
    # data_as keeps a reference to the array, so temporaries live through the call
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _sfx(dtype):
    return "f" if np.dtype(dtype) == np.float32 else "d"


class Csr:
    """Host CSR triple (int32 row/col, float64 or float32 values)."""

    def __init__(self, row, col, val, ncol=None):
        self.row = np.ascontiguousarray(row, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val)
        self.n = len(self.row) - 1
        self.ncol = self.n if ncol is None else ncol

    @property
    def nnz(self):
        return int(self.row[-1])

    def astype(self, dtype):
        return Csr(self.row, self.col, self.val.astype(dtype), self.ncol)

    def to_dense(self):
        A = np.zeros((self.n, self.ncol), dtype=np.float64)
        for i in range(self.n):
            for j in range(self.row[i], self.row[i + 1]):
                A[i, self.col[j]] += self.val[j]
        return A


def csr_from_coo(nrow, rows, cols, data=None):
    rows = np.ascontiguousarray(rows, dtype=np.int32)
    cols = np.ascontiguousarray(cols, dtype=np.int32)
    nnz = len(rows)
    rp = np.zeros(nrow + 1, np.int32)
    ci = np.zeros(nnz, np.int32)
    vv = np.zeros(nnz, np.float64)
    dd = None if data is None else np.ascontiguousarray(data, dtype=np.float64)
    lib().hfo_csr_from_coo_d(nrow, nnz, _p(rows), _p(cols), _p(dd), _p(rp), _p(ci), _p(vv))
    return Csr(rp, ci, vv)


def csr_add_values(A, rows, cols, values):
    rows = np.ascontiguousarray(rows, np.int32)
    cols = np.ascontiguousarray(cols, np.int32)
    values = np.ascontiguousarray(values, np.float64)
    lib().hfo_csr_add_values_d(A.n, _p(A.row), _p(A.col), _p(A.val), _p(rows), len(rows),
                               _p(cols), len(cols), _p(values))


def spmv(A, x, omp=False):
    s = _sfx(A.val.dtype)
    x = np.ascontiguousarray(x, dtype=A.val.dtype)
    y = np.zeros(A.n, dtype=A.val.dtype)
    fn = getattr(lib(), f"hfo_spmv_omp_{s}" if omp else f"hfo_spmv_{s}")
    fn(A.n, _p(A.row), _p(A.col), _p(A.val), _p(x), _p(y))
    return y


def spmv_add(A, x, y):
    s = _sfx(A.val.dtype)
    y = np.array(y, dtype=A.val.dtype)
    getattr(lib(), f"hfo_spmv_add_{s}")(A.n, _p(A.row), _p(A.col), _p(A.val),
                                        _p(np.ascontiguousarray(x, A.val.dtype)), _p(y))
    return y


def dot(x, y):
    s = _sfx(x.dtype)
    return getattr(lib(), f"hfo_dot_{s}")(len(x), _p(x), _p(np.ascontiguousarray(y, x.dtype)))


def nrm2(x):
    return getattr(lib(), f"hfo_nrm2_{_sfx(x.dtype)}")(len(x), _p(x))


def axpy(y, x, a):
    y = np.array(y)
    getattr(lib(), f"hfo_axpy_{_sfx(y.dtype)}")(len(y), _p(y), _p(np.ascontiguousarray(x, y.dtype)), a)
    return y


def scaleadd(y, a, x):
    y = np.array(y)
    getattr(lib(), f"hfo_scaleadd_{_sfx(y.dtype)}")(len(y), _p(y), a, _p(np.ascontiguousarray(x, y.dtype)))
    return y


def scale(y, a):
    y = np.array(y)
    getattr(lib(), f"hfo_scale_{_sfx(y.dtype)}")(len(y), _p(y), a)
    return y


def ewmult(y, x):
    y = np.array(y)
    getattr(lib(), f"hfo_ewmult_{_sfx(y.dtype)}")(len(y), _p(y), _p(np.ascontiguousarray(x, y.dtype)))
    return y


def norm1(x):
    return getattr(lib(), f"hfo_norm1_{_sfx(x.dtype)}")(len(x), _p(x))


def argmax(x):
    return getattr(lib(), f"hfo_argmax_{_sfx(x.dtype)}")(len(x), _p(x))


def argmin(x):
    return getattr(lib(), f"hfo_argmin_{_sfx(x.dtype)}")(len(x), _p(x))


def add_values(y, idx, v):
    y = np.array(y)
    idx = np.ascontiguousarray(idx, np.int32)
    getattr(lib(), f"hfo_add_values_{_sfx(y.dtype)}")(_p(y), _p(idx), len(idx), _p(np.ascontiguousarray(v, y.dtype)))
    return y


def add_range(y, s, x, ss, m, w):
    y = np.array(y)
    getattr(lib(), f"hfo_add_range_{_sfx(y.dtype)}")(_p(y), s, _p(np.ascontiguousarray(x, y.dtype)), ss, m, w)
    return y


def rot(a, b, c, s):
    a, b = np.array(a), np.array(b, dtype=np.asarray(a).dtype)
    getattr(lib(), f"hfo_rot_{_sfx(a.dtype)}")(len(a), _p(a), _p(b), c, s)
    return a, b


def rotm(u, v, p):
    u, v = np.array(u), np.array(v, dtype=np.asarray(u).dtype)
    getattr(lib(), f"hfo_rotm_{_sfx(u.dtype)}")(len(u), _p(u), _p(v), _p(np.ascontiguousarray(p, u.dtype)))
    return u, v


def rotg(a, b, dtype=np.float64):
    v = np.array([a, b, 0, 0], dtype)
    getattr(lib(), f"hfo_rotg_{_sfx(dtype)}")(_p(v[0:]), _p(v[1:]), _p(v[2:]), _p(v[3:]))
    return tuple(v)


def zero_rows(A, idx, alpha):
    val = A.val.copy()
    idx = np.ascontiguousarray(idx, np.int32)
    getattr(lib(), f"hfo_zero_rows_{_sfx(val.dtype)}")(_p(A.row), _p(A.col), _p(val), _p(idx), len(idx), alpha)
    return val


def zero_cols(A, idx, alpha):
    val = A.val.copy()
    idx = np.ascontiguousarray(idx, np.int32)
    getattr(lib(), f"hfo_zero_cols_{_sfx(val.dtype)}")(A.n, _p(A.row), _p(A.col), _p(val), _p(idx), len(idx), alpha)
    return val


def mat_scale(A, a, offdiag_only=False):
    val = A.val.copy()
    getattr(lib(), f"hfo_mat_scale_{_sfx(val.dtype)}")(A.n, _p(A.row), _p(A.col), _p(val), a, int(offdiag_only))
    return val


def pjacobi(A, x, out=None):
    out = np.zeros(A.n, A.val.dtype) if out is None else np.array(out, A.val.dtype)
    getattr(lib(), f"hfo_pjacobi_{_sfx(A.val.dtype)}")(A.n, _p(A.row), _p(A.col), _p(A.val),
                                                       _p(np.ascontiguousarray(x, A.val.dtype)), _p(out))
    return out


def invdiag(A):
    d = np.zeros(A.n, A.val.dtype)
    getattr(lib(), f"hfo_invdiag_{_sfx(A.val.dtype)}")(A.n, _p(A.row), _p(A.col), _p(A.val), _p(d))
    return d


def cast_d2f(x):
    x = np.ascontiguousarray(x, np.float64)
    out = np.zeros(len(x), np.float32)
    lib().hfo_cast_d2f(len(x), _p(x), _p(out))
    return out


def cast_f2d(x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(len(x), np.float64)
    lib().hfo_cast_f2d(len(x), _p(x), _p(out))
    return out


def sgs(A, r):
    out = np.zeros(A.n, A.val.dtype)
    getattr(lib(), f"hfo_sgs_{_sfx(A.val.dtype)}")(A.n, _p(A.row), _p(A.col), _p(A.val),
                                                   _p(np.ascontiguousarray(r, A.val.dtype)), _p(out))
    return out


def ilu0(A):
    """Returns (LU factors as Csr, number of missing diagonals)."""
    LU = Csr(A.row, A.col, A.val.copy(), A.ncol)
    miss = getattr(lib(), f"hfo_ilu0_{_sfx(A.val.dtype)}")(A.n, _p(LU.row), _p(LU.col), _p(LU.val))
    return LU, miss


def ilu_solve(LU, r):
    out = np.zeros(LU.n, LU.val.dtype)
    getattr(lib(), f"hfo_ilu_solve_{_sfx(LU.val.dtype)}")(LU.n, _p(LU.row), _p(LU.col), _p(LU.val),
                                                          _p(np.ascontiguousarray(r, LU.val.dtype)), _p(out))
    return out


def multicoloring(A):
    """Returns (ncolors, colour of each node (1-based), colour sizes, perm[new]=old)."""
    nc = C.c_int(0)
    color = np.zeros(A.n, np.int32)
    sizes = np.zeros(max(A.n, 1), np.int32)
    perm = np.zeros(A.n, np.int32)
    bad = lib().hfo_multicoloring(A.n, _p(A.row), _p(A.col), C.byref(nc), _p(color), _p(sizes), _p(perm))
    if bad:
        raise RuntimeError(f"multicoloring: {bad} rows without a colour")
    return nc.value, color, sizes[: nc.value].copy(), perm


def levelscheduling(A):
    """Returns (nlevels, nlevels as the reference reports it, level sizes, perm)."""
    rep = C.c_int(0)
    sizes = np.zeros(max(A.n, 1), np.int32)
    perm = np.zeros(A.n, np.int32)
    nl = lib().hfo_levelscheduling(A.n, _p(A.row), _p(A.col), C.byref(rep), _p(sizes), _p(perm))
    return nl, rep.value, sizes[:nl].copy(), perm


def csr_reorder(A, perm):
    perm = np.ascontiguousarray(perm, np.int32)
    rp = np.zeros(A.n + 1, np.int32)
    ci = np.zeros(A.nnz, np.int32)
    vv = np.zeros(A.nnz, np.float64)
    lib().hfo_csr_reorder_d(A.n, _p(A.row), _p(A.col), _p(A.val.astype(np.float64)), _p(perm),
                            _p(rp), _p(ci), _p(vv))
    return Csr(rp, ci, vv)


def vec_reorder(v, perm):
    out = np.zeros(len(v), np.float64)
    lib().hfo_vec_reorder_d(len(v), _p(np.ascontiguousarray(v, np.float64)),
                            _p(np.ascontiguousarray(perm, np.int32)), _p(out))
    return out


def mc_dropoffs(A, sizes):
    sizes = np.ascontiguousarray(sizes, np.int32)
    return lib().hfo_mc_dropoffs_d(A.n, _p(A.row), _p(A.col), _p(A.val), len(sizes), _p(sizes))


def mcsgs(A, sizes, r, omega=1.0):
    sizes = np.ascontiguousarray(sizes, np.int32)
    out = np.zeros(A.n, A.val.dtype)
    getattr(lib(), f"hfo_mcsgs_{_sfx(A.val.dtype)}")(A.n, _p(A.row), _p(A.col), _p(A.val), len(sizes),
                                                     _p(sizes), omega,
                                                     _p(np.ascontiguousarray(r, A.val.dtype)), _p(out))
    return out


class Prec:
    """Preconditioner description for the solvers (keeps its arrays alive)."""

    def __init__(self, kind=PREC_NONE, A=None, invd=None, sizes=None, omega=1.0):
        self.kind, self.A, self.invd, self.omega = kind, A, invd, omega
        self.sizes = None if sizes is None else np.ascontiguousarray(sizes, np.int32)

    @staticmethod
    def jacobi(A):
        return Prec(PREC_JACOBI, A, invd=invdiag(A))

    def struct(self, dtype):
        S = PrecF if np.dtype(dtype) == np.float32 else PrecD
        keep = []

        def conv(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype)
            keep.append(a)
            return a.ctypes.data
        s = S()
        s.kind = self.kind
        s.invd = conv(self.invd)
        if self.A is not None:
            s.row, s.col, s.val = _p(self.A.row), _p(self.A.col), conv(self.A.val)
        s.ncolors = 0 if self.sizes is None else len(self.sizes)
        s.color_sizes = _p(self.sizes)
        s.omega = self.omega
        s._keep = keep
        return s


def _solve(fn_name, A, b, x0, P, ctrl, extra=()):
    s = _sfx(A.val.dtype)
    x = np.array(x0 if x0 is not None else np.zeros(A.n), dtype=A.val.dtype)
    b = np.ascontiguousarray(b, A.val.dtype)
    c = Ctrl(*ctrl)
    r = Result()
    hist = np.zeros(ctrl[0] + 2, np.float64)
    ps = (P or Prec()).struct(A.val.dtype)
    getattr(lib(), f"{fn_name}_{s}")(A.n, _p(A.row), _p(A.col), _p(A.val), _p(b), _p(x),
                                     C.byref(ps), *extra, C.byref(c), C.byref(r), _p(hist))
    return x, {"iters": r.iters, "status": r.status, "ret": r.ret, "res": r.res, "res0": r.res0,
               "hist": hist[: r.iters + 1].copy()}


def cg(A, b, x0=None, P=None, ctrl=(1000, 1e-15, 1e-8, 1e6)):
    return _solve("hfo_cg", A, b, x0, P, ctrl)


def gmres(A, b, x0=None, P=None, m=30, ctrl=(1000, 1e-15, 1e-8, 1e6)):
    return _solve("hfo_gmres", A, b, x0, P, ctrl, (m,))


def fgmres(A, b, x0=None, P=None, m=30, ctrl=(1000, 1e-15, 1e-8, 1e6)):
    return _solve("hfo_fgmres", A, b, x0, P, ctrl, (m,))


def swap_diag_front(A):
    """CPU_CSR_lMatrix::SwapDiagElementsToRowFront on a copy of A."""
    B = Csr(A.row.copy(), A.col.copy(), A.val.copy(), A.ncol)
    getattr(lib(), f"hfo_swap_diag_front_{_sfx(B.val.dtype)}")(B.n, _p(B.row), _p(B.col), _p(B.val))
    return B


def spmv_nodiag(A, x):
    y = np.zeros(A.n, A.val.dtype)
    getattr(lib(), f"hfo_spmv_nodiag_{_sfx(A.val.dtype)}")(A.n, _p(A.row), _p(A.col), _p(A.val),
                                                            _p(np.ascontiguousarray(x, A.val.dtype)), _p(y))
    return y


def jacobi(A, b, x0=None, mode=0, w=1.0, inner_iter=1, smooth_iter=1, ctrl=(1000, 1e-15, 1e-8, 1e6)):
    """Jacobi<LAD> on the operator A after Prepare (diagonal swapped to the row front)."""
    x = np.array(x0 if x0 is not None else np.zeros(A.n), dtype=A.val.dtype)
    b = np.ascontiguousarray(b, A.val.dtype)
    c = Ctrl(*ctrl)
    r = Result()
    hist = np.zeros(ctrl[0] + 2, np.float64)
    getattr(lib(), f"hfo_jacobi_{_sfx(A.val.dtype)}")(A.n, _p(A.row), _p(A.col), _p(A.val), _p(b), _p(x), mode, w,
                                                       inner_iter, smooth_iter, C.byref(c), C.byref(r), _p(hist))
    nh = 0 if mode >= 2 else (r.iters // max(inner_iter, 1)) + 1
    return x, {"iters": r.iters, "status": r.status, "ret": r.ret, "res": r.res, "res0": r.res0,
               "hist": hist[:nh].copy()}


def mpir(A, b, x0=None, Pl=None, Ph=None, max_ech=0, max_ecl=10, abs_tol=1e-12, rel_tol=1e-12,
         div_tol=1e6, ecl=(1000, 1e-6, 1e-1, 1e6), ech=(1000, 1e-12, 1e-12, 1e6)):
    """MPIR<double, float> with CG error-correction solvers (no smoothers)."""
    x = np.array(x0 if x0 is not None else np.zeros(A.n), dtype=np.float64)
    b = np.ascontiguousarray(b, np.float64)
    vf = A.val.astype(np.float32)
    p = MpirParams(max_ech, max_ecl, abs_tol, rel_tol, div_tol, Ctrl(*ecl), Ctrl(*ech))
    r = MpirResult()
    pl = (Pl or Prec()).struct(np.float32)
    ph = (Ph or Prec()).struct(np.float64)
    hist = np.zeros(max_ech + max_ecl + 2, np.float64)
    lib().hfo_mpir(A.n, _p(A.row), _p(A.col), _p(A.val), _p(vf), _p(b), _p(x), C.byref(pl),
                   C.byref(ph), C.byref(p), C.byref(r), _p(hist))
    nloop = r.n_ecl + r.n_ech
    return x, {"ret": r.ret, "iters": r.iters, "n_ecl": r.n_ecl, "n_ech": r.n_ech, "res": r.res,
               "hist": hist[: nloop + 1].copy()}
