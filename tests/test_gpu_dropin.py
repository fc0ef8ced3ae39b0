"""The C++ drop-in layer (include/laamd/dropin.hpp), exercised by a C++
program written against the reference's interfaces (tests/cpp/
dropin_test.cpp, built by __graft_entry__.build())."""
import os
import subprocess

import pytest

from helpers import ROOT

pytestmark = pytest.mark.gpu


def test_cpp_dropin_program(dev):
    exe = os.path.join(ROOT, "tests", "cpp", "dropin_test")
    if not os.path.exists(exe):
        pytest.fail("tests/cpp/dropin_test not built; run __graft_entry__.build()")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "FAIL" not in r.stdout
