"""The C-ABI library loads and exports every entry point include/laamd.h
declares (no GPU call is made), and the oracle library builds and loads."""
import ctypes as C
import os
import re

import linearalgebra_amd as la
from linearalgebra_amd import _lib
from helpers import ROOT


def _declared():
    txt = open(os.path.join(ROOT, "include", "laamd.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return set(re.findall(r"\b(la_[a-z0-9_]+)\s*\(", txt))


def test_library_exports_every_declared_symbol():
    L = C.CDLL(_lib.LIB_PATH)
    declared = _declared()
    assert len(declared) >= 50
    missing = [s for s in sorted(declared) if not hasattr(L, s)]
    assert not missing, missing
    # the Python binding covers the same set
    assert declared == set(_lib.EXPORTED)


def test_error_reporting_without_gpu():
    L = la.lib()
    # an invalid argument is reported with a status and a message
    st = L.la_topo_sort2(-1, 0, None, None, 0, None, None, None)
    assert st == _lib.LA_ERR_ARG
    assert b"null" in L.la_last_error() or L.la_last_error()
    assert L.la_version() >= 100


def test_oracle_loads():
    from oracle import oracle as orc
    assert orc.max_threads() >= 1


def test_blas1_null_vectors_rejected():
    """la_dot / la_axpy_dot check their vectors before touching the device."""
    r = C.c_double()
    assert la.lib().la_dot(10, _lib.F64, None, None, C.byref(r), None) == _lib.LA_ERR_ARG
    assert la.lib().la_axpy_dot(10, _lib.F64, 1.0, None, None, None, C.byref(r),
                                None) == _lib.LA_ERR_ARG
