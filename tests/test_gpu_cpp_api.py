"""Runs the C++ host-API tests (tests/cpp/reference_api_test.cpp, built by __graft_entry__.build()) on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "reference_api_test")


def test_cpp_reference_api():
    assert os.path.exists(BIN), "run __graft_entry__.build() first"
    p = subprocess.run([BIN], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=240)
    print(p.stdout)
    assert p.returncode == 0, p.stdout
    assert "all reference-API tests passed" in p.stdout
