"""The C-ABI libraries load and export every symbol their headers declare.
No compute call is made (this runs without a GPU)."""
import ctypes as C
import os
import re

from conftest import PKG_DIR, ROOT


def declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hf[hl]_[a-z0-9_]+)\s*\(", text)))


def test_hip_library_exports_header(pkg):
    lib = C.CDLL(os.path.join(PKG_DIR, "libhfhip.so"))
    names = declared("hfh.h")
    assert len(names) > 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_lmp_library_exports_header(pkg):
    pkg.hfh.hip()
    lib = C.CDLL(os.path.join(PKG_DIR, "libhflmp.so"))
    names = declared("hflmp.h")
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_errors_are_status_codes(pkg):
    L = pkg.hfh.hip()
    # null arguments are rejected with a code and a message, never an exit
    assert L.hfh_vec_zeros(None) == pkg.ERR_INVALID
    assert b"null" in L.hfh_last_error()
    assert L.hfh_dot(None, None, None) == pkg.ERR_INVALID
    assert L.hfh_pcg_solve(None, None, None, None, None, None, None) == pkg.ERR_INVALID
    assert L.hfh_version().startswith(b"hfh-gfx950")
