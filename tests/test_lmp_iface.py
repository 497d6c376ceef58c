"""The lmp interface of the product (hiflow3_2.0_amd/lmp/lmp.h), driven by
the C++ test tests/lmp_iface.cc the way the reference's own C++ tests drive
theirs (test/LinearAlgebra/CPUsimple_CSR_lMatrix.cc): known answers of those
tests, the reference's loops restated in the oracle, and every lVector /
lMatrix virtual on the HIP platform through base-class pointers from the
platform factories.  The driver is built by __graft_entry__.build()."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "bin", "lmp_iface")


def run_driver(mode):
    assert os.path.exists(DRIVER), f"{DRIVER} is missing: run __graft_entry__.build()"
    r = subprocess.run([DRIVER, mode], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "0 failed" in r.stdout


def test_lmp_interface_host():
    run_driver("cpu")


@pytest.mark.gpu
def test_lmp_interface_hip():
    run_driver("hip")
