"""The host thread team behind the threaded RCM and ILU(k) symbolic phases
(linearalgebra_amd/csrc/team.hpp): an exception in any member's work ends
the run on every member and is rethrown by the caller, with no hang and no
use of the work after the run returns (tests/cpp/team_test.cpp, built here
with g++). No GPU is touched."""
import os
import shutil
import subprocess

import pytest

from helpers import ROOT


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_team_exceptions(tmp_path):
    exe = str(tmp_path / "team_test")
    src = os.path.join(ROOT, "tests", "cpp", "team_test.cpp")
    subprocess.run(["g++", "-std=c++17", "-O2", "-pthread", "-o", exe, src], check=True,
                   timeout=300)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "FAIL" not in r.stdout
