"""Runs the C++ host-API test program (tests/cpp/filter_api_test.cpp) on the
GPU: the reference's filter_api KAT through dtl::filter::filter, the class
API, error behaviour, the type-erased wrappers and the cuckoo filter."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "bloomfilter-bsd_amd", "bin", "filter_api_test")


def test_cpp_filter_api():
    assert os.path.exists(BIN), "C++ API test not built; run __graft_entry__.build()"
    r = subprocess.run([BIN], capture_output=True, text=True, timeout=300)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASS" in r.stdout
