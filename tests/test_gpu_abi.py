"""The drop-in boundary from compiled callers: a C program built against
include/tblis.h and a C++ program built against include/tblis.hpp
(tests/abi/, compiled by __graft_entry__.build()) allocate with hipMalloc,
run contractions through tblis_tensor_mult / tblis::mult and compare with
the oracle themselves (the reference's install check, test/test_install.cxx).
Each exits 0 when its results are within 1e-12 relative Frobenius error."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ABI = os.path.join(os.path.dirname(os.path.abspath(__file__)), "abi")


@pytest.mark.parametrize("prog", ["install_c", "install_cpp"])
def test_compiled_caller(prog):
    exe = os.path.join(ABI, prog)
    assert os.path.exists(exe), "build the callers first (__graft_entry__.build())"
    print("running %s" % exe, flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=100)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0, (r.returncode, r.stdout, r.stderr)
