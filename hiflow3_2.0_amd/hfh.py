"""Python binding of the gfx950 HIP platform (libhfhip.so) and of the host
layer (libhflmp.so), over their C ABIs (include/hfh.h, include/hflmp.h).

There is no fallback: if the HIP library is missing or no GPU is present the
calls raise.  Arrays cross the boundary as numpy buffers (host) or opaque
device handles; the device never sees a torch tensor.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
HIP_LIB = os.path.join(_HERE, "libhfhip.so")
LMP_LIB = os.path.join(_HERE, "libhflmp.so")

F64, F32 = 0, 1
PREC_NONE, PREC_JACOBI, PREC_SGS, PREC_MCSGS, PREC_ILU0 = 0, 1, 2, 3, 4
ITERATE, SUCCESS_ABS, SUCCESS_REL, FAIL_DIV, FAIL_MAXITS = 0, 1, 2, 3, 4
OK, ERR_INVALID, ERR_NOMEM, ERR_HIP, ERR_ZERO_PIVOT, ERR_NOT_CONVERGED, ERR_BUDGET, ERR_UNSUPPORTED = \
    0, -1, -2, -3, -4, -5, -6, -7

Q1_2D, Q1_3D, FV_DIFFUSION, UPWIND_CONVDIFF = 1, 2, 3, 4
SWEEP_SYNCFREE, SWEEP_LEVELS, SWEEP_SLICES = 0, 1, 2
SPMV_AUTO, SPMV_STREAM, SPMV_VECTOR, SPMV_WAVE, SPMV_SELL, SPMV_SELL_DICT, SPMV_SELL_PAIRS, SPMV_SELL_LDS = \
    0, 1, 2, 3, 4, 5, 6, 7


class HfhError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"hfh error {code}: {msg}")
        self.code = code


class Control(C.Structure):
    _fields_ = [("maxits", C.c_int), ("atol", C.c_double), ("rtol", C.c_double), ("dtol", C.c_double)]


class SolveInfo(C.Structure):
    _fields_ = [("iters", C.c_int), ("state", C.c_int), ("res", C.c_double), ("res0", C.c_double),
                ("time_ms", C.c_double)]


class DcInfo(C.Structure):
    _fields_ = [("outer_iters", C.c_int), ("inner_iters", C.c_int), ("state", C.c_int), ("res", C.c_double),
                ("res0", C.c_double), ("time_ms", C.c_double)]


class BatchParams(C.Structure):
    _fields_ = [("max_systems", C.c_int), ("hbm_budget", C.c_longlong), ("chunk", C.c_int), ("rtol", C.c_double),
                ("atol", C.c_double), ("outer_maxits", C.c_int), ("inner_rtol", C.c_double),
                ("inner_maxits", C.c_int), ("small_tiles", C.c_int)]


class BatchResult(C.Structure):
    _fields_ = [("outer_iters", C.c_int), ("inner_iters", C.c_int), ("state", C.c_int), ("res", C.c_double),
                ("res0", C.c_double)]


_hip = None
_lmp = None

# C ABI signatures: name -> (restype, argtypes)
_vp, _i, _d, _ll = C.c_void_p, C.c_int, C.c_double, C.c_longlong
_P = C.POINTER
HIP_API = {
    "hfh_last_error": (C.c_char_p, []),
    "hfh_version": (C.c_char_p, []),
    "hfh_device_count": (_i, [_P(_i)]),
    "hfh_ctx_create": (_i, [_i, _P(_vp)]),
    "hfh_ctx_destroy": (_i, [_vp]),
    "hfh_ctx_sync": (_i, [_vp]),
    "hfh_ctx_info": (_i, [_vp, _P(_i), _P(_ll), C.c_char_p, _i]),
    "hfh_ctx_set_timing": (_i, [_vp, _i]),
    "hfh_ctx_timing": (_i, [_vp, _i, _P(_ll), _P(_d)]),
    "hfh_ctx_timing_reset": (_i, [_vp]),
    "hfh_vec_create": (_i, [_vp, _i, _i, _P(_vp)]),
    "hfh_vec_destroy": (_i, [_vp]),
    "hfh_vec_size": (_i, [_vp, _P(_i)]),
    "hfh_vec_dtype": (_i, [_vp, _P(_i)]),
    "hfh_vec_upload": (_i, [_vp, _vp, _i, _i]),
    "hfh_vec_download": (_i, [_vp, _vp, _i, _i]),
    "hfh_vec_set_values": (_i, [_vp, _vp, _i, _vp]),
    "hfh_vec_get_values": (_i, [_vp, _vp, _i, _vp]),
    "hfh_vec_zeros": (_i, [_vp]),
    "hfh_vec_copy": (_i, [_vp, _vp]),
    "hfh_vec_copy_range": (_i, [_vp, _i, _vp, _i, _i]),
    "hfh_vec_cast": (_i, [_vp, _vp]),
    "hfh_vec_reorder": (_i, [_vp, _vp, _vp]),
    "hfh_dot": (_i, [_vp, _vp, _P(_d)]),
    "hfh_dot_dev": (_i, [_vp, _vp, _vp]),
    "hfh_vec_device_ptr": (_i, [_vp, _P(_vp)]),
    "hfh_nrm2": (_i, [_vp, _P(_d)]),
    "hfh_axpy": (_i, [_vp, _vp, _d]),
    "hfh_scaleadd": (_i, [_vp, _d, _vp]),
    "hfh_scale": (_i, [_vp, _d]),
    "hfh_ewmult": (_i, [_vp, _vp]),
    "hfh_norm1": (_i, [_vp, _P(_d)]),
    "hfh_iamax": (_i, [_vp, _P(_i)]),
    "hfh_iamin": (_i, [_vp, _P(_i)]),
    "hfh_normmax": (_i, [_vp, _P(_d)]),
    "hfh_vec_add_values": (_i, [_vp, _vp, _i, _vp]),
    "hfh_vec_add_range": (_i, [_vp, _i, _vp, _i, _i, _d]),
    "hfh_vec_set_indexset": (_i, [_vp, _vp, _i]),
    "hfh_vec_get_indexset": (_i, [_vp, _vp, _P(_i)]),
    "hfh_vec_get_indexed_values": (_i, [_vp, _vp]),
    "hfh_vec_set_indexed_values": (_i, [_vp, _vp]),
    "hfh_rot": (_i, [_vp, _vp, _d, _d]),
    "hfh_rotm": (_i, [_vp, _vp, _vp]),
    "hfh_csr_diag": (_i, [_vp, _i, _i, _i, _vp]),
    "hfh_csr_zeros": (_i, [_vp]),
    "hfh_csr_scale": (_i, [_vp, _d, _i]),
    "hfh_csr_zero_rows": (_i, [_vp, _vp, _i, _d]),
    "hfh_csr_zero_cols": (_i, [_vp, _vp, _i, _d]),
    "hfh_csr_pjacobi": (_i, [_vp, _vp, _vp]),
    "hfh_csr_create": (_i, [_vp, _i, _i, _i, _i, _P(_vp)]),
    "hfh_csr_destroy": (_i, [_vp]),
    "hfh_csr_dims": (_i, [_vp, _P(_i), _P(_i), _P(_i), _P(_i)]),
    "hfh_csr_upload": (_i, [_vp, _vp, _vp, _vp]),
    "hfh_csr_upload_values": (_i, [_vp, _vp]),
    "hfh_csr_download": (_i, [_vp, _vp, _vp, _vp]),
    "hfh_csr_cast": (_i, [_vp, _vp]),
    "hfh_spmv": (_i, [_vp, _vp, _vp]),
    "hfh_spmv_add": (_i, [_vp, _vp, _vp]),
    "hfh_csr_invdiag": (_i, [_vp, _vp]),
    "hfh_csr_set_spmv": (_i, [_vp, _i, _i, _i]),
    "hfh_csr_get_spmv": (_i, [_vp, _P(_i), _P(_i)]),
    "hfh_jacobi_create": (_i, [_vp, _P(_vp)]),
    "hfh_sgs_create": (_i, [_vp, _i, _vp, _vp, _i, _vp, _vp, _P(_vp)]),
    "hfh_mcsgs_create": (_i, [_vp, _i, _vp, _d, _P(_vp)]),
    "hfh_ilu0_create": (_i, [_vp, _i, _vp, _vp, _i, _vp, _vp, _P(_vp)]),
    "hfh_csr_ilu0": (_i, [_vp, _i, _vp, _vp, _P(_i)]),
    "hfh_prec_apply": (_i, [_vp, _vp, _vp]),
    "hfh_prec_kind": (_i, [_vp, _P(_i)]),
    "hfh_prec_set_sweep": (_i, [_vp, _i, _i]),
    "hfh_prec_sweep_status": (_i, [_vp]),
    "hfh_prec_sweep_blocks": (_i, [_vp, _P(C.c_int), _P(C.c_int), _P(C.c_int), _P(C.c_int)]),
    "hfh_prec_destroy": (_i, [_vp]),
    "hfh_pcg_solve": (_i, [_vp, _vp, _vp, _vp, _P(Control), _P(SolveInfo), _vp]),
    "hfh_pcg_solve_noalloc": (_i, [_vp, _vp, _vp, _vp, _P(Control), _P(SolveInfo), _vp]),
    "hfh_gmres_solve": (_i, [_vp, _vp, _i, _i, _vp, _vp, _P(Control), _P(SolveInfo), _vp]),
    "hfh_csr_levels": (_i, [_vp, _i, _P(_i), _vp, _vp]),
    "hfh_csr_multicoloring": (_i, [_vp, _P(_i), _vp, _vp, _vp]),
    "hfh_csr_permute": (_i, [_vp, _vp, _vp]),
    "hfh_problem_assemble": (_i, [_vp, _i, _i, _i, _i, _d, _d, C.c_ulonglong, _vp, _vp]),
    "hfh_csr_swap_diag_front": (_i, [_vp]),
    "hfh_csr_is_sorted": (_i, [_vp, _P(_i)]),
    "hfh_spmv_nodiag": (_i, [_vp, _vp, _vp]),
    "hfh_jacobi_solve": (_i, [_vp, _vp, _vp, _i, _d, _i, _i, _P(Control), _P(SolveInfo), _vp]),
    "hfh_dc_solve": (_i, [_vp, _vp, _vp, _vp, _vp, _P(Control), _P(Control), _P(DcInfo), _vp]),
    "hfh_dc_solve_ex": (_i, [_vp, _vp, _vp, _vp, _vp, _P(Control), _P(Control), _P(C.c_int), _P(DcInfo), _vp]),
    "hfh_batch_create": (_i, [_vp, _P(BatchParams), _P(_vp)]),
    "hfh_batch_destroy": (_i, [_vp]),
    "hfh_batch_bytes_csr": (_i, [_i, _vp, _P(_ll)]),
    "hfh_batch_bytes_fv": (_i, [_i, _i, _i, _P(_ll)]),
    "hfh_batch_submit_csr": (_i, [_vp, _i, _vp, _vp, _vp, _vp, _vp, _P(_i)]),
    "hfh_batch_submit_fv": (_i, [_vp, _i, _i, _i, _vp, _vp, _vp, _P(_i)]),
    "hfh_batch_step": (_i, [_vp, _P(_i), _P(_i)]),
    "hfh_batch_finished": (_i, [_vp, _vp, _i, _P(_i)]),
    "hfh_batch_collect": (_i, [_vp, _i, _vp, _P(BatchResult)]),
    "hfh_batch_usage": (_i, [_vp, _P(_i), _P(_ll), _P(_ll)]),
    "hfh_batch_stats": (_i, [_vp, _P(_ll), _P(C.c_double), _P(_ll)]),
}

class LmpParams(C.Structure):
    _fields_ = [("maxits", C.c_int), ("atol", C.c_double), ("rtol", C.c_double), ("dtol", C.c_double)]


class LmpResult(C.Structure):
    _fields_ = [("iters", C.c_int), ("ret", C.c_int), ("state", C.c_int), ("res", C.c_double),
                ("hist", C.POINTER(C.c_double)), ("hist_cap", C.c_int), ("nhist", C.c_int)]


LMP_API = {
    "hfl_problem_size": (_i, [_i, _i, _i, _i, _d, _d, C.c_ulonglong, _P(_i), _P(_ll)]),
    "hfl_problem_fill": (_i, [_i, _i, _i, _i, _d, _d, C.c_ulonglong, _vp, _vp, _vp, _vp]),
    "hfl_fv_cells": (_i, [_i, _i, _i, _d, C.c_ulonglong, _vp, _vp]),
    "hfl_mm_read": (_i, [C.c_char_p, _P(_i), _P(_i), _P(_ll), _vp, _vp, _vp]),
    "hfl_mm_write": (_i, [C.c_char_p, _i, _i, _vp, _vp, _vp]),
    "hfl_vec_read": (_i, [C.c_char_p, _P(_i), _vp]),
    "hfl_vec_write": (_i, [C.c_char_p, _i, _vp]),
    "hfl_uniform_pm1": (_d, [C.c_ulonglong, C.c_ulonglong, C.c_ulonglong]),
    "hfl_normal01": (_d, [C.c_ulonglong, C.c_ulonglong, C.c_ulonglong]),
    "hfl_multicoloring": (_i, [_i, _vp, _vp, _P(_i), _vp, _vp, _vp]),
    "hfl_levels_lower": (_i, [_i, _vp, _vp, _P(_i), _vp, _vp]),
    "hfl_levels_upper": (_i, [_i, _vp, _vp, _P(_i), _vp, _vp]),
    "hfl_ilu0": (_i, [_i, _vp, _vp, _vp]),
    "hfl_csr_permute": (_i, [_i, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "hfl_last_error": (C.c_char_p, []),
    "hfl_lmp_solve": (_i, [_i, _i, _i, _d, _i, _i, _vp, _vp, _vp, _vp, _vp, _P(LmpParams), _P(LmpResult)]),
    "hfl_lmp_mpir": (_i, [_i, _i, _vp, _vp, _vp, _vp, _vp, _i, _d, _d, _d, _P(LmpParams), _P(LmpResult)]),
}

SOLVER_CG, SOLVER_GMRES, SOLVER_FGMRES = 0, 1, 2


def _declare(lib, api):
    for name, (res, args) in api.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args


def hip():
    """The HIP platform library; raises if it has not been built."""
    global _hip
    if _hip is None:
        if not os.path.exists(HIP_LIB):
            raise ImportError(f"{HIP_LIB} is missing: run hiflow3_2.0_amd/build.py (or __graft_entry__.build())")
        lib = C.CDLL(HIP_LIB, mode=C.RTLD_GLOBAL)
        _declare(lib, HIP_API)
        _hip = lib
    return _hip


def lmp():
    global _lmp
    if _lmp is None:
        hip()
        if not os.path.exists(LMP_LIB):
            raise ImportError(f"{LMP_LIB} is missing: run hiflow3_2.0_amd/build.py")
        lib = C.CDLL(LMP_LIB)
        _declare(lib, LMP_API)
        _lmp = lib
    return _lmp


def _ck(rc):
    if rc != OK:
        raise HfhError(rc, hip().hfh_last_error().decode())
    return rc


def _p(a):
    # data_as keeps a reference to the array, so temporaries live through the call
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _dt(dtype):
    return F32 if np.dtype(dtype) == np.float32 else F64


def _np(dt):
    return np.float32 if dt == F32 else np.float64


def device_count():
    c = C.c_int(0)
    _ck(hip().hfh_device_count(C.byref(c)))
    return c.value


class Context:
    """One HIP stream plus reduction scratch on one device."""

    def __init__(self, device=0):
        h = C.c_void_p()
        _ck(hip().hfh_ctx_create(device, C.byref(h)))
        self.h = h

    def close(self):
        """Destroys the context once no vector, matrix or preconditioner uses it
        (the library refuses earlier; the context then lives to process exit)."""
        if self.h:
            hip().hfh_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        _ck(hip().hfh_ctx_sync(self.h))

    def info(self):
        cu, hbm = C.c_int(), C.c_longlong()
        name = C.create_string_buffer(128)
        _ck(hip().hfh_ctx_info(self.h, C.byref(cu), C.byref(hbm), name, 128))
        return {"num_cu": cu.value, "hbm_bytes": hbm.value, "arch": name.value.decode()}

    def set_timing(self, on):
        _ck(hip().hfh_ctx_set_timing(self.h, 1 if on else 0))

    def timing(self, kclass=0):
        n, ms = C.c_longlong(), C.c_double()
        _ck(hip().hfh_ctx_timing(self.h, kclass, C.byref(n), C.byref(ms)))
        return n.value, ms.value

    def timing_reset(self):
        _ck(hip().hfh_ctx_timing_reset(self.h))


class Vector:
    def __init__(self, ctx, n_or_array, dtype=np.float64):
        self.ctx = ctx
        if isinstance(n_or_array, (int, np.integer)):
            arr, n = None, int(n_or_array)
        else:
            arr = np.ascontiguousarray(n_or_array, dtype=dtype)
            n = len(arr)
        self.n, self.dtype = n, np.dtype(dtype)
        h = C.c_void_p()
        _ck(hip().hfh_vec_create(ctx.h, n, _dt(dtype), C.byref(h)))
        self.h = h
        if arr is not None:
            self.upload(arr)

    def __del__(self):
        try:
            if self.h:
                hip().hfh_vec_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def upload(self, arr, offset=0):
        arr = np.ascontiguousarray(arr, dtype=self.dtype)
        _ck(hip().hfh_vec_upload(self.h, _p(arr), offset, len(arr)))

    def numpy(self, offset=0, count=None):
        count = self.n - offset if count is None else count
        out = np.empty(count, dtype=self.dtype)
        _ck(hip().hfh_vec_download(self.h, _p(out), offset, count))
        return out

    def set_values(self, index, values):
        index = np.ascontiguousarray(index, np.int32)
        values = np.ascontiguousarray(values, self.dtype)
        _ck(hip().hfh_vec_set_values(self.h, _p(index), len(index), _p(values)))

    def get_values(self, index):
        index = np.ascontiguousarray(index, np.int32)
        out = np.empty(len(index), self.dtype)
        _ck(hip().hfh_vec_get_values(self.h, _p(index), len(index), _p(out)))
        return out

    def zeros(self):
        _ck(hip().hfh_vec_zeros(self.h))

    def copy_from(self, other):
        _ck(hip().hfh_vec_copy(self.h, other.h))

    def copy_range(self, dst_off, src, src_off, count):
        _ck(hip().hfh_vec_copy_range(self.h, dst_off, src.h, src_off, count))

    def cast_from(self, other):
        _ck(hip().hfh_vec_cast(self.h, other.h))

    def reorder_from(self, other, perm):
        perm = np.ascontiguousarray(perm, np.int32)
        _ck(hip().hfh_vec_reorder(self.h, other.h, _p(perm)))

    def dot(self, other):
        d = C.c_double()
        _ck(hip().hfh_dot(self.h, other.h, C.byref(d)))
        return d.value

    def device_ptr(self):
        p = C.c_void_p()
        _ck(hip().hfh_vec_device_ptr(self.h, C.byref(p)))
        return p.value

    def dot_dev(self, other, dst, index=0):
        """x.y stored into the FP64 vector dst at index, without a host wait."""
        _ck(hip().hfh_dot_dev(self.h, other.h, C.c_void_p(dst.device_ptr() + 8 * index)))

    def nrm2(self):
        d = C.c_double()
        _ck(hip().hfh_nrm2(self.h, C.byref(d)))
        return d.value

    def axpy(self, x, a):
        _ck(hip().hfh_axpy(self.h, x.h, float(a)))

    def scaleadd(self, a, x):
        _ck(hip().hfh_scaleadd(self.h, float(a), x.h))

    def scale(self, a):
        _ck(hip().hfh_scale(self.h, float(a)))

    def ewmult(self, x):
        _ck(hip().hfh_ewmult(self.h, x.h))

    def norm1(self):
        d = C.c_double()
        _ck(hip().hfh_norm1(self.h, C.byref(d)))
        return d.value

    def normmax(self):
        d = C.c_double()
        _ck(hip().hfh_normmax(self.h, C.byref(d)))
        return d.value

    def argmax(self):
        i = C.c_int()
        _ck(hip().hfh_iamax(self.h, C.byref(i)))
        return i.value

    def argmin(self):
        i = C.c_int()
        _ck(hip().hfh_iamin(self.h, C.byref(i)))
        return i.value

    def add_values(self, index, values):
        idx = np.ascontiguousarray(index, np.int32)
        v = np.ascontiguousarray(values, self.dtype)
        _ck(hip().hfh_vec_add_values(self.h, _p(idx), len(idx), _p(v)))

    def add_range(self, dst_off, src, src_off, count, weight):
        _ck(hip().hfh_vec_add_range(self.h, dst_off, src.h, src_off, count, float(weight)))

    def set_indexset(self, index):
        idx = np.ascontiguousarray(index, np.int32)
        _ck(hip().hfh_vec_set_indexset(self.h, _p(idx), len(idx)))

    def get_indexset(self):
        n = C.c_int()
        _ck(hip().hfh_vec_get_indexset(self.h, None, C.byref(n)))
        idx = np.zeros(n.value, np.int32)
        _ck(hip().hfh_vec_get_indexset(self.h, _p(idx), C.byref(n)))
        return idx

    def get_indexed_values(self):
        out = np.zeros(len(self.get_indexset()), self.dtype)
        _ck(hip().hfh_vec_get_indexed_values(self.h, _p(out)))
        return out

    def set_indexed_values(self, values):
        v = np.ascontiguousarray(values, self.dtype)
        _ck(hip().hfh_vec_set_indexed_values(self.h, _p(v)))

    def rot(self, other, c, s):
        _ck(hip().hfh_rot(self.h, other.h, float(c), float(s)))

    def rotm(self, other, param):
        p = np.ascontiguousarray(param, np.float64)
        _ck(hip().hfh_rotm(self.h, other.h, _p(p)))


class CsrMatrix:
    def __init__(self, ctx, row, col, val, ncol=None, dtype=None):
        self.ctx = ctx
        row = np.ascontiguousarray(row, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        dtype = np.dtype(dtype or np.asarray(val).dtype)
        val = np.ascontiguousarray(val, dtype)
        self.nrow = len(row) - 1
        self.ncol = self.nrow if ncol is None else ncol
        self.nnz = len(col)
        self.dtype = dtype
        h = C.c_void_p()
        _ck(hip().hfh_csr_create(ctx.h, self.nrow, self.ncol, self.nnz, _dt(dtype), C.byref(h)))
        self.h = h
        _ck(hip().hfh_csr_upload(self.h, _p(row), _p(col), _p(val)))

    @classmethod
    def cast_of(cls, other, dtype):
        self = cls.__new__(cls)
        self.ctx, self.nrow, self.ncol, self.nnz = other.ctx, other.nrow, other.ncol, other.nnz
        self.dtype = np.dtype(dtype)
        h = C.c_void_p()
        _ck(hip().hfh_csr_create(other.ctx.h, self.nrow, self.ncol, self.nnz, _dt(dtype), C.byref(h)))
        self.h = h
        _ck(hip().hfh_csr_cast(self.h, other.h))
        return self

    def __del__(self):
        try:
            if self.h:
                hip().hfh_csr_destroy(self.h)
                self.h = None
        except Exception:
            pass

    def spmv(self, x, y):
        _ck(hip().hfh_spmv(self.h, x.h, y.h))

    def spmv_add(self, x, y):
        _ck(hip().hfh_spmv_add(self.h, x.h, y.h))

    def upload_values(self, val):
        """New values for the same structure (CopyFrom with an unchanged pattern)."""
        v = np.ascontiguousarray(val, self.dtype)
        _ck(hip().hfh_csr_upload_values(self.h, _p(v)))

    def download(self):
        """(row, col, val) of the device matrix."""
        nr, nc, nz, dt = C.c_int(), C.c_int(), C.c_int(), C.c_int()
        _ck(hip().hfh_csr_dims(self.h, C.byref(nr), C.byref(nc), C.byref(nz), C.byref(dt)))
        row = np.empty(nr.value + 1, np.int32)
        col = np.empty(nz.value, np.int32)
        val = np.empty(nz.value, np.float64 if dt.value == F64 else np.float32)
        _ck(hip().hfh_csr_download(self.h, _p(row), _p(col), _p(val)))
        return row, col, val

    def get_spmv(self):
        """(algorithm, param1) of the SpMV layout in use."""
        a, p1 = C.c_int(), C.c_int()
        _ck(hip().hfh_csr_get_spmv(self.h, C.byref(a), C.byref(p1)))
        return a.value, p1.value

    def set_spmv(self, algo, param1=0, param2=0):
        """Select the SpMV kernel (SPMV_AUTO / SPMV_STREAM / SPMV_VECTOR)."""
        _ck(hip().hfh_csr_set_spmv(self.h, algo, param1, param2))

    def ilu0(self, row, col):
        """Factor this matrix in place into ILU(0) factors on the device
        (row/col: host copy of the pattern, for the level schedule).  Returns
        the number of lower entries skipped for lack of a diagonal."""
        ptr, rows = levels(row, col)
        miss = C.c_int()
        _ck(hip().hfh_csr_ilu0(self.h, len(ptr) - 1, _p(ptr), _p(rows), C.byref(miss)))
        return miss.value

    def invdiag(self):
        v = Vector(self.ctx, self.nrow, self.dtype)
        _ck(hip().hfh_csr_invdiag(self.h, v.h))
        return v

    def diag(self, start=0, end=None, invert=False):
        end = self.nrow if end is None else end
        v = Vector(self.ctx, end - start, self.dtype)
        _ck(hip().hfh_csr_diag(self.h, start, end, 1 if invert else 0, v.h))
        return v

    def zeros(self):
        _ck(hip().hfh_csr_zeros(self.h))

    def scale(self, alpha, offdiag_only=False):
        _ck(hip().hfh_csr_scale(self.h, float(alpha), 1 if offdiag_only else 0))

    def zero_rows(self, rows, alpha):
        r = np.ascontiguousarray(rows, np.int32)
        _ck(hip().hfh_csr_zero_rows(self.h, _p(r), len(r), float(alpha)))

    def zero_cols(self, cols, alpha):
        c = np.ascontiguousarray(cols, np.int32)
        _ck(hip().hfh_csr_zero_cols(self.h, _p(c), len(c), float(alpha)))

    def pjacobi(self, x, y):
        _ck(hip().hfh_csr_pjacobi(self.h, x.h, y.h))

    def levels(self, upper=False):
        """Level schedule computed on the device: (level_ptr, level_rows), the
        same arrays as levels(row, col, upper)."""
        nl = C.c_int()
        ptr = np.zeros(self.nrow + 1, np.int32)
        rows = np.zeros(self.nrow, np.int32)
        _ck(hip().hfh_csr_levels(self.h, 1 if upper else 0, C.byref(nl), _p(ptr), _p(rows)))
        return ptr[: nl.value + 1].copy(), rows

    def multicoloring(self):
        """First-fit multicolouring on the device: (ncolors, color_of,
        color_sizes, perm), the same as multicoloring(row, col)."""
        nc = C.c_int()
        color = np.zeros(self.nrow, np.int32)
        sizes = np.zeros(self.nrow, np.int32)
        perm = np.zeros(self.nrow, np.int32)
        _ck(hip().hfh_csr_multicoloring(self.h, C.byref(nc), _p(color), _p(sizes), _p(perm)))
        return nc.value, color, sizes[: nc.value].copy(), perm

    def permuted(self, perm):
        """A new device matrix P with P[i, j] = A[perm[i], perm[j]] (columns
        ascending), built on the device."""
        P = CsrMatrix.__new__(CsrMatrix)
        P.ctx, P.nrow, P.ncol, P.nnz, P.dtype = self.ctx, self.nrow, self.ncol, self.nnz, self.dtype
        h = C.c_void_p()
        _ck(hip().hfh_csr_create(self.ctx.h, self.nrow, self.ncol, self.nnz, _dt(self.dtype), C.byref(h)))
        P.h = h
        _ck(hip().hfh_csr_permute(self.h, _p(np.ascontiguousarray(perm, np.int32)), P.h))
        return P

    def swap_diag_front(self):
        """SwapDiagElementsToRowFront in place (what Jacobi's Prepare does to
        the operator); the columns are no longer ascending afterwards."""
        _ck(hip().hfh_csr_swap_diag_front(self.h))

    def is_sorted(self):
        v = C.c_int()
        _ck(hip().hfh_csr_is_sorted(self.h, C.byref(v)))
        return bool(v.value)

    def spmv_nodiag(self, x, y):
        """VectorMultNoDiag: y = (A - D) x, stored order."""
        _ck(hip().hfh_spmv_nodiag(self.h, x.h, y.h))


class Preconditioner:
    def __init__(self, h, keep=()):
        self.h = h
        self._keep = keep

    def __del__(self):
        try:
            if self.h:
                hip().hfh_prec_destroy(self.h)
                self.h = None
        except Exception:
            pass

    @staticmethod
    def jacobi(A):
        h = C.c_void_p()
        _ck(hip().hfh_jacobi_create(A.h, C.byref(h)))
        return Preconditioner(h, (A,))

    @staticmethod
    def mcsgs(A, color_sizes, omega=1.0):
        sizes = np.ascontiguousarray(color_sizes, np.int32)
        h = C.c_void_p()
        _ck(hip().hfh_mcsgs_create(A.h, len(sizes), _p(sizes), float(omega), C.byref(h)))
        return Preconditioner(h, (A,))

    @staticmethod
    def _trsv(fn, A, row, col):
        lo_ptr, lo_rows = levels(row, col, upper=False)
        up_ptr, up_rows = levels(row, col, upper=True)
        h = C.c_void_p()
        _ck(fn(A.h, len(lo_ptr) - 1, _p(lo_ptr), _p(lo_rows), len(up_ptr) - 1, _p(up_ptr), _p(up_rows),
               C.byref(h)))
        return Preconditioner(h, (A,))

    @staticmethod
    def sgs(A, row, col):
        """Natural-order symmetric Gauss-Seidel; row/col: host copy of A's pattern."""
        return Preconditioner._trsv(hip().hfh_sgs_create, A, row, col)

    @staticmethod
    def ilu0(LU, row, col):
        """ILU(0) apply; LU holds the factors (see ilu0_factor)."""
        return Preconditioner._trsv(hip().hfh_ilu0_create, LU, row, col)

    def apply(self, x, y):
        _ck(hip().hfh_prec_apply(self.h, x.h, y.h))

    def set_sweep(self, mode, waves_per_cu=0):
        """SGS / ILU(0): SWEEP_SYNCFREE (default: geometric blocks on a
        structured grid, else level slices), SWEEP_SLICES or SWEEP_LEVELS
        schedule; waves_per_cu of a sync-free sweep (0: default)."""
        _ck(hip().hfh_prec_set_sweep(self.h, mode, waves_per_cu))

    def sweep_blocks(self):
        """(forward blocks, forward block levels, backward blocks, backward
        block levels) of the blocked sync-free sweeps; zeros when the
        pattern is not a structured grid's."""
        v = [C.c_int() for _ in range(4)]
        _ck(hip().hfh_prec_sweep_blocks(self.h, *[C.byref(x) for x in v]))
        return tuple(x.value for x in v)

    def sweep_status(self):
        """Raises if a sync-free sweep gave up waiting since the last check."""
        _ck(hip().hfh_prec_sweep_status(self.h))


def _info(inf):
    return {"iters": inf.iters, "state": inf.state, "res": inf.res, "res0": inf.res0, "time_ms": inf.time_ms}


def pcg(A, b, x, P=None, maxits=1000, atol=1e-15, rtol=1e-8, dtol=1e6, history=False, check=True,
        noalloc=False):
    """CG<LAD>::Solve on the device; returns the info dict (and history).
    noalloc: the allocation-free entry point (work space in the library's
    static device array; Jacobi or no preconditioner)."""
    ctl = Control(maxits, atol, rtol, dtol)
    inf = SolveInfo()
    hist = np.zeros(maxits + 2) if history else None
    fn = hip().hfh_pcg_solve_noalloc if noalloc else hip().hfh_pcg_solve
    rc = fn(A.h, P.h if P else None, b.h, x.h, C.byref(ctl), C.byref(inf), _p(hist))
    if rc != OK and (check or rc != ERR_NOT_CONVERGED):
        _ck(rc)
    out = _info(inf)
    out["rc"] = rc
    if history:
        out["hist"] = hist[: inf.iters + 1].copy()
    return out


def gmres(A, b, x, P=None, m=30, flexible=False, maxits=1000, atol=1e-15, rtol=1e-8, dtol=1e6,
          history=False, check=True):
    ctl = Control(maxits, atol, rtol, dtol)
    inf = SolveInfo()
    hist = np.zeros(maxits + 2) if history else None
    rc = hip().hfh_gmres_solve(A.h, P.h if P else None, m, 1 if flexible else 0, b.h, x.h, C.byref(ctl),
                               C.byref(inf), _p(hist))
    if rc != OK and (check or rc != ERR_NOT_CONVERGED):
        _ck(rc)
    out = _info(inf)
    out["rc"] = rc
    if history:
        out["hist"] = hist[: inf.iters + 1].copy()
    return out


JACOBI_SOLVE_NORMAL, JACOBI_SOLVE_DAMPED, JACOBI_SMOOTH_NORMAL, JACOBI_SMOOTH_DAMPED = 0, 1, 2, 3


def jacobi(A, b, x, mode=JACOBI_SOLVE_NORMAL, w=1.0, inner_iter=1, smooth_iter=1, maxits=1000, atol=1e-15,
           rtol=1e-8, dtol=1e6, history=False, check=True):
    """Jacobi<LAD>::Solve on the device (A after swap_diag_front, as Prepare
    leaves it); returns the info dict (and the checked residuals)."""
    ctl = Control(maxits, atol, rtol, dtol)
    inf = SolveInfo()
    hist = np.zeros(maxits // max(inner_iter, 1) + 2) if history else None
    rc = hip().hfh_jacobi_solve(A.h, b.h, x.h, mode, w, inner_iter, smooth_iter, C.byref(ctl), C.byref(inf),
                                _p(hist))
    if rc != OK and (check or rc != ERR_NOT_CONVERGED):
        _ck(rc)
    out = _info(inf)
    out["rc"] = rc
    if history and mode < 2:
        out["hist"] = hist[: inf.iters // max(inner_iter, 1) + 1].copy()
    return out


def dc(Ah, Al, b, x, Pl=None, outer=(20, 1e-10, 0.0, 1e6), inner=(1000, 0.0, 1e-4, 1e6), check=True,
       keep_direction=False):
    """FP64 defect correction with FP32 inner PCG (MPIR<D,S>::Solve); with
    keep_direction the inner CG continues its search direction across
    corrections (reliable updates)."""
    oc, ic = Control(*outer), Control(*inner)
    inf = DcInfo()
    hist = np.zeros(outer[0] + 2)
    opt = C.c_int(1 if keep_direction else 0)
    rc = hip().hfh_dc_solve_ex(Ah.h, Al.h, Pl.h if Pl else None, b.h, x.h, C.byref(oc), C.byref(ic), C.byref(opt),
                               C.byref(inf), _p(hist))
    if rc != OK and (check or rc != ERR_NOT_CONVERGED):
        _ck(rc)
    return {"outer_iters": inf.outer_iters, "inner_iters": inf.inner_iters, "state": inf.state, "res": inf.res,
            "res0": inf.res0, "time_ms": inf.time_ms, "rc": rc, "hist": hist[: inf.outer_iters + 1].copy()}


class BatchEngine:
    """Serving engine (hfh_batch_*): continuous batching of SPD solves by
    FP32-inner / FP64-outer CG.  submit_* return a request id, or None when
    the request does not fit the free slots / HBM budget now."""

    def __init__(self, ctx, max_systems=256, hbm_budget=1 << 30, chunk=16, rtol=1e-10, atol=0.0,
                 outer_maxits=30, inner_rtol=1e-4, inner_maxits=5000, small_tiles=0):
        self.ctx = ctx
        self.params = BatchParams(max_systems, int(hbm_budget), chunk, rtol, atol, outer_maxits, inner_rtol,
                                  inner_maxits, small_tiles)
        h = C.c_void_p()
        _ck(hip().hfh_batch_create(ctx.h, C.byref(self.params), C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            hip().hfh_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def bytes_csr(row):
        b = C.c_longlong()
        row = np.ascontiguousarray(row, np.int32)
        _ck(hip().hfh_batch_bytes_csr(len(row) - 1, _p(row), C.byref(b)))
        return b.value

    @staticmethod
    def bytes_fv(nx, ny, nz):
        b = C.c_longlong()
        _ck(hip().hfh_batch_bytes_fv(nx, ny, nz, C.byref(b)))
        return b.value

    def _admit(self, rc, rid):
        if rc == ERR_BUDGET:
            return None
        _ck(rc)
        return rid.value

    def submit_csr(self, row, col, val, b, x0=None):
        row = np.ascontiguousarray(row, np.int32)
        col = np.ascontiguousarray(col, np.int32)
        val = np.ascontiguousarray(val, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        x0 = None if x0 is None else np.ascontiguousarray(x0, np.float64)
        rid = C.c_int()
        rc = hip().hfh_batch_submit_csr(self.h, len(row) - 1, _p(row), _p(col), _p(val), _p(b), _p(x0),
                                        C.byref(rid))
        return self._admit(rc, rid)

    def submit_fv(self, nx, ny, nz, kcell, b, x0=None):
        kcell = np.ascontiguousarray(kcell, np.float64)
        b = np.ascontiguousarray(b, np.float64)
        x0 = None if x0 is None else np.ascontiguousarray(x0, np.float64)
        rid = C.c_int()
        rc = hip().hfh_batch_submit_fv(self.h, nx, ny, nz, _p(kcell), _p(b), _p(x0), C.byref(rid))
        return self._admit(rc, rid)

    def step(self):
        """One chunk of iterations; returns (active, finished_this_step)."""
        a, f = C.c_int(), C.c_int()
        _ck(hip().hfh_batch_step(self.h, C.byref(a), C.byref(f)))
        return a.value, f.value

    def finished(self):
        cnt = C.c_int()
        _ck(hip().hfh_batch_finished(self.h, None, 0, C.byref(cnt)))
        ids = np.zeros(max(cnt.value, 1), np.int32)
        _ck(hip().hfh_batch_finished(self.h, _p(ids), cnt.value, C.byref(cnt)))
        return [int(i) for i in ids[: cnt.value]]

    def collect(self, rid, n):
        """Returns (x, info dict); info['rc'] is the collect status."""
        x = np.empty(n, np.float64)
        r = BatchResult()
        rc = hip().hfh_batch_collect(self.h, rid, _p(x), C.byref(r))
        if rc not in (OK, ERR_NOT_CONVERGED):
            _ck(rc)
        return x, {"outer_iters": r.outer_iters, "inner_iters": r.inner_iters, "state": r.state, "res": r.res,
                   "res0": r.res0, "rc": rc}

    def stats(self):
        """(replays, GPU ms of the replays, inner row-iterations) since creation."""
        r, ms, ri = _ll(), C.c_double(), _ll()
        _ck(hip().hfh_batch_stats(self.h, C.byref(r), C.byref(ms), C.byref(ri)))
        return r.value, ms.value, ri.value

    def usage(self):
        a, u, c = C.c_int(), C.c_longlong(), C.c_longlong()
        _ck(hip().hfh_batch_usage(self.h, C.byref(a), C.byref(u), C.byref(c)))
        return a.value, u.value, c.value


# ---------------------------------------------------------------- host layer

def problem(kind, nx, ny=None, nz=None, p0=1.0, p1=0.0, seed=0):
    """Generate a seeded CSR system: returns (row, col, val, rhs)."""
    ny = nx if ny is None else ny
    nz = (1 if kind == Q1_2D else nx) if nz is None else nz
    n, nnz = C.c_int(), C.c_longlong()
    rc = lmp().hfl_problem_size(kind, nx, ny, nz, p0, p1, seed, C.byref(n), C.byref(nnz))
    if rc:
        raise ValueError("invalid problem specification")
    row = np.empty(n.value + 1, np.int32)
    col = np.empty(nnz.value, np.int32)
    val = np.empty(nnz.value, np.float64)
    rhs = np.empty(n.value, np.float64)
    if lmp().hfl_problem_fill(kind, nx, ny, nz, p0, p1, seed, _p(row), _p(col), _p(val), _p(rhs)):
        raise ValueError("problem generation failed")
    return row, col, val, rhs


def assemble(ctx, kind, nx, ny=None, nz=None, p0=1.0, p1=0.0, seed=0, dtype=np.float64):
    """The system problem() generates on the host, assembled on the device:
    (CsrMatrix, right-hand side Vector), bit-identical to problem()."""
    ny = nx if ny is None else ny
    nz = (1 if kind == Q1_2D else nx) if nz is None else nz
    n, nnz = C.c_int(), C.c_longlong()
    if lmp().hfl_problem_size(kind, nx, ny, nz, p0, p1, seed, C.byref(n), C.byref(nnz)):
        raise ValueError("bad problem size")
    kc = None
    if kind == FV_DIFFUSION:
        kc = np.empty(n.value, np.float64)
        _lmp_ck(lmp().hfl_fv_cells(nx, ny, nz, p0, seed, _p(kc), None))
    A = CsrMatrix.__new__(CsrMatrix)
    A.ctx, A.nrow, A.ncol, A.nnz, A.dtype = ctx, n.value, n.value, nnz.value, np.dtype(dtype)
    h = C.c_void_p()
    _ck(hip().hfh_csr_create(ctx.h, n.value, n.value, nnz.value, _dt(dtype), C.byref(h)))
    A.h = h
    b = Vector(ctx, n.value, dtype)
    _ck(hip().hfh_problem_assemble(A.h, kind, nx, ny, nz, p0, p1, seed, _p(kc), b.h))
    return A, b


def mm_read(path):
    """MatrixMarket file -> (row, col, val, ncol), CSR with ascending columns."""
    nr, nc, nz = C.c_int(), C.c_int(), C.c_longlong()
    bpath = os.fsencode(path)
    _lmp_ck(lmp().hfl_mm_read(bpath, C.byref(nr), C.byref(nc), C.byref(nz), None, None, None))
    row = np.empty(nr.value + 1, np.int32)
    col = np.empty(nz.value, np.int32)
    val = np.empty(nz.value, np.float64)
    _lmp_ck(lmp().hfl_mm_read(bpath, C.byref(nr), C.byref(nc), C.byref(nz), _p(row), _p(col), _p(val)))
    return row, col, val, nc.value


def mm_write(path, row, col, val, ncol=None):
    row = np.ascontiguousarray(row, np.int32)
    n = len(row) - 1
    _lmp_ck(lmp().hfl_mm_write(os.fsencode(path), n, n if ncol is None else ncol, _p(row),
                               _p(np.ascontiguousarray(col, np.int32)), _p(np.ascontiguousarray(val, np.float64))))


def vec_read(path):
    n = C.c_int(0)
    bpath = os.fsencode(path)
    _lmp_ck(lmp().hfl_vec_read(bpath, C.byref(n), None))
    out = np.empty(n.value, np.float64)
    _lmp_ck(lmp().hfl_vec_read(bpath, C.byref(n), _p(out)))
    return out[: n.value]


def vec_write(path, values):
    values = np.ascontiguousarray(values, np.float64)
    _lmp_ck(lmp().hfl_vec_write(os.fsencode(path), len(values), _p(values)))


def fv_cells(nx, ny, nz, sigma, seed):
    """Cell coefficients k = exp(sigma g) and rhs of the FV diffusion problem."""
    n = nx * ny * nz
    k = np.empty(n, np.float64)
    b = np.empty(n, np.float64)
    if lmp().hfl_fv_cells(nx, ny, nz, sigma, seed, _p(k), _p(b)):
        raise ValueError("invalid grid")
    return k, b


def multicoloring(row, col):
    n = len(row) - 1
    nc = C.c_int()
    color = np.empty(n, np.int32)
    sizes = np.empty(max(n, 1), np.int32)
    perm = np.empty(n, np.int32)
    if lmp().hfl_multicoloring(n, _p(row), _p(col), C.byref(nc), _p(color), _p(sizes), _p(perm)):
        raise ValueError("multicoloring failed")
    return nc.value, color, sizes[: nc.value].copy(), perm


def levels(row, col, upper=False):
    row = np.ascontiguousarray(row, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    n = len(row) - 1
    nl = C.c_int()
    ptr = np.empty(n + 1, np.int32)
    rows = np.empty(n, np.int32)
    fn = lmp().hfl_levels_upper if upper else lmp().hfl_levels_lower
    if fn(n, _p(row), _p(col), C.byref(nl), _p(ptr), _p(rows)):
        raise ValueError("level scheduling failed")
    return ptr[: nl.value + 1].copy(), rows


def ilu0_factor(row, col, val):
    lu = np.array(val, np.float64)
    miss = lmp().hfl_ilu0(len(row) - 1, _p(np.ascontiguousarray(row, np.int32)),
                          _p(np.ascontiguousarray(col, np.int32)), _p(lu))
    if miss:
        raise HfhError(ERR_ZERO_PIVOT, f"ILU(0): {miss} missing diagonal entries")
    return lu


def csr_permute(row, col, val, perm):
    n = len(row) - 1
    row = np.ascontiguousarray(row, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    perm = np.ascontiguousarray(perm, np.int32)
    prow = np.empty(n + 1, np.int32)
    pcol = np.empty(len(col), np.int32)
    pval = np.empty(len(col), np.float64)
    if lmp().hfl_csr_permute(n, _p(row), _p(col), _p(val), _p(perm), _p(prow), _p(pcol), _p(pval)):
        raise ValueError("invalid permutation")
    return prow, pcol, pval


def _lmp_ck(rc):
    if rc != OK:
        raise HfhError(rc, lmp().hfl_last_error().decode())


def lmp_solve(row, col, val, b, x0=None, solver=SOLVER_CG, precond=PREC_NONE, m=30, omega=1.0, single=False,
              maxits=1000, atol=1e-15, rtol=1e-8, dtol=1e6):
    """Runs the C++ LinearSolver templates (CG / GMRES / FGMRES) on HIP vectors."""
    n = len(row) - 1
    row = np.ascontiguousarray(row, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.array(np.zeros(n) if x0 is None else x0, np.float64)
    hist = np.zeros(maxits + 2)
    res = LmpResult()
    res.hist = hist.ctypes.data_as(C.POINTER(C.c_double))
    res.hist_cap = len(hist)
    prm = LmpParams(maxits, atol, rtol, dtol)
    _lmp_ck(lmp().hfl_lmp_solve(solver, precond, m, omega, 1 if single else 0, n, _p(row), _p(col), _p(val),
                                _p(b), _p(x), C.byref(prm), C.byref(res)))
    return x, {"iters": res.iters, "ret": res.ret, "state": res.state, "res": res.res,
               "hist": hist[: res.nhist].copy()}


def lmp_mpir(row, col, val, b, x0=None, precond_low=PREC_JACOBI, max_ecl=20, abs_tol=1e-12, rel_tol=1e-12,
             div_tol=1e6, ecl=(1000, 0.0, 1e-3, 1e6)):
    n = len(row) - 1
    row = np.ascontiguousarray(row, np.int32)
    col = np.ascontiguousarray(col, np.int32)
    val = np.ascontiguousarray(val, np.float64)
    b = np.ascontiguousarray(b, np.float64)
    x = np.array(np.zeros(n) if x0 is None else x0, np.float64)
    hist = np.zeros(max_ecl + 2)
    res = LmpResult()
    res.hist = hist.ctypes.data_as(C.POINTER(C.c_double))
    res.hist_cap = len(hist)
    _lmp_ck(lmp().hfl_lmp_mpir(precond_low, n, _p(row), _p(col), _p(val), _p(b), _p(x), max_ecl, abs_tol, rel_tol,
                               div_tol, C.byref(LmpParams(*ecl)), C.byref(res)))
    return x, {"iters": res.iters, "ret": res.ret, "n_ecl": res.state, "res": res.res,
               "hist": hist[: res.nhist].copy()}
