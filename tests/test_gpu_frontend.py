"""Runs the C++ front-end test program (tests/cpp/test_functional_frontend.cpp),
which drives the HIP path through include/serac_amd/functional.hpp and checks
it against the oracle, FD convergence ratios, the affine patch test and the
Newton solver."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_cpp_functional_frontend():
    exe = os.path.join(ROOT, "tests", "cpp", "bin", "test_functional_frontend")
    assert os.path.exists(exe), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    print(r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
