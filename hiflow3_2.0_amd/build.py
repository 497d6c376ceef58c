"""Build recipe of the gfx950 HIP platform (libhfhip.so) and the C++ host
layer (libhflmp.so).

Both libraries are built in-tree next to their sources so that they travel
with the repository snapshot.  Run as a script or call build_all().
"""
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LMP = os.path.join(HERE, "lmp")
ROOT = os.path.dirname(HERE)
INCLUDE = os.path.join(ROOT, "include")
BUILD = os.path.join(HERE, "_build")
HIP_LIB = os.path.join(HERE, "libhfhip.so")
LMP_LIB = os.path.join(HERE, "libhflmp.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("HFH_ARCH", "gfx950")

HIP_SOURCES = ["hfh_core.hip", "hfh_spmv.hip", "hfh_prec.hip", "hfh_krylov.hip", "hfh_batch.hip", "hfh_jacobi.hip",
               "hfh_analysis.hip", "hfh_assemble.hip", "hfh_vecops.hip", "hfh_matops.hip"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall", "-Wno-unused-result",
    # packed FP32 instructions are kept out of the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
]

LMP_SOURCES = ["problems.cc", "analysis.cc", "hflmp_capi.cc", "lmp_host.cc", "solve_capi.cc", "mmio.cc"]
CXX = os.environ.get("CXX", "g++")
CXX_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-Wextra", "-Wno-unused-parameter", f"-I{INCLUDE}",
             f"-I{LMP}"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError(f"build step failed: {os.path.basename(cmd[-1])}")
    return r.stdout


def _stale(out, deps):
    if not os.path.exists(out):
        return True
    t = os.path.getmtime(out)
    return any(os.path.getmtime(d) > t for d in deps)


def build_hip(force=True, jobs=4):
    os.makedirs(BUILD, exist_ok=True)
    hdrs = [os.path.join(CSRC, h) for h in os.listdir(CSRC) if h.endswith(".h")] + \
        [os.path.join(INCLUDE, "hfh.h")]
    objs, todo = [], []
    for s in HIP_SOURCES:
        src = os.path.join(CSRC, s)
        obj = os.path.join(BUILD, s.replace(".hip", ".o"))
        objs.append(obj)
        if force or _stale(obj, [src] + hdrs):
            todo.append([HIPCC] + HIP_FLAGS + ["-c", "-o", obj, src])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, todo))
    tmp = HIP_LIB + ".tmp"
    # every symbol resolved at link time: a missing definition fails the build
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,--no-undefined", "-o", tmp] + objs)
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


def build_lmp(force=True, jobs=4):
    os.makedirs(BUILD, exist_ok=True)
    hdrs = [os.path.join(LMP, h) for h in os.listdir(LMP) if h.endswith(".h")] + \
        [os.path.join(INCLUDE, f) for f in os.listdir(INCLUDE)]
    objs, todo = [], []
    for s in LMP_SOURCES:
        src = os.path.join(LMP, s)
        obj = os.path.join(BUILD, "lmp_" + s.replace(".cc", ".o"))
        objs.append(obj)
        if force or _stale(obj, [src] + hdrs):
            todo.append([CXX] + CXX_FLAGS + ["-c", "-o", obj, src])
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(_run, todo))
    tmp = LMP_LIB + ".tmp"
    # the host layer reaches the device only through libhfhip.so's C ABI
    _run([CXX, "-shared", "-fPIC", "-fopenmp", "-o", tmp] + objs +
         [f"-L{HERE}", "-lhfhip", "-Wl,-rpath,$ORIGIN", "-Wl,--no-undefined"])
    os.replace(tmp, LMP_LIB)
    return LMP_LIB


ORACLE = os.path.join(ROOT, "oracle")
TEST_DRIVER = os.path.join(ROOT, "tests", "bin", "lmp_iface")


def build_test_driver():
    """The C++ test of the lmp interface (tests/lmp_iface.cc), linked with
    both product libraries and the oracle it checks against."""
    os.makedirs(os.path.dirname(TEST_DRIVER), exist_ok=True)
    tmp = TEST_DRIVER + ".tmp"
    _run([CXX, "-O1", "-std=c++17", "-Wall", f"-I{INCLUDE}", f"-I{LMP}", f"-I{ORACLE}",
          "-o", tmp, os.path.join(ROOT, "tests", "lmp_iface.cc"),
          f"-L{HERE}", "-lhflmp", "-lhfhip", f"-L{ORACLE}", "-lhfo", "-fopenmp",
          f"-Wl,-rpath,$ORIGIN/../../hiflow3_2.0_amd", "-Wl,-rpath,$ORIGIN/../../oracle"])
    os.replace(tmp, TEST_DRIVER)
    return TEST_DRIVER


def build_all(force=True):
    build_hip(force)
    build_lmp(force)


if __name__ == "__main__":
    build_all(force="--incremental" not in sys.argv)
    print(HIP_LIB)
    print(LMP_LIB)
