"""fmma::FMMA<TYPE,DIM> from C++ (tests/cpp/dropin_test.cpp, built by
__graft_entry__.build()): user __host__ __device__ functors through set_fn,
built-in kernels, float and double, the time log keys."""
import os
import subprocess

import pytest

from tests.conftest import ROOT

pytestmark = pytest.mark.gpu


def test_cpp_dropin_program():
    exe = os.path.join(ROOT, "tests", "cpp", "dropin_test")
    assert os.path.exists(exe), "tests/cpp/dropin_test is not built"
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    print(r.stderr[-2000:])
    assert r.returncode == 0 and r.stdout.strip().endswith("OK")
