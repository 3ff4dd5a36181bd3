"""CPU checks of the C-ABI library: it loads, and exports every entry point include/torchannel.h
declares (no compute calls without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "torchannel.h")
LIB = os.path.join(ROOT, "torchannel_amd", "libtorchannel_hip.so")


def declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(tc_[a-z0-9_]+)\s*\(", text)))


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        from torchannel_amd import build
        build.build()
    return ctypes.CDLL(LIB)


def test_header_declares_entry_points():
    names = declared()
    assert "tc_solver_step" in names and "tc_poisson_solve" in names and len(names) >= 20


def test_library_exports_every_declared_symbol(lib):
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_strerror_and_version(lib):
    lib.tc_strerror.restype = ctypes.c_char_p
    assert b"invalid argument" in lib.tc_strerror(-1000)
    assert lib.tc_version() >= 10000


def test_python_binding_covers_header():
    from torchannel_amd import _lib
    for n in declared():
        if n not in ("tc_strerror",):
            assert n in _lib.SIGNATURES, n


def test_argument_errors_without_gpu(lib):
    # null handles are rejected before any device work
    assert lib.tc_apply_bc(None, None, None, None, None) == -1000
    assert lib.tc_poisson_solve(None, None, None, None) == -1000
    assert lib.tc_solver_step(None, ctypes.c_double(0.0), 0, None, None, None) == -1000
