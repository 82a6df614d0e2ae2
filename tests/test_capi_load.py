"""The C-ABI library loads and exports exactly what include/snapy_amd.h
declares (CPU: no compute calls)."""

import ctypes
import os
import re

import pytest

from snapy_amd import _capi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for fn in os.listdir(os.path.join(ROOT, "include")):
        if not fn.endswith(".h"):
            continue
        text = open(os.path.join(ROOT, "include", fn)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(snapy_[a-z0-9_]+)\s*\(", text):
            names.add(m.group(1))
    return names


def test_header_declares_the_binding():
    assert declared_symbols() == set(_capi.SIGNATURES)


@pytest.mark.skipif(not _capi.LIB_PATH.exists(), reason="library not built")
def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(str(_capi.LIB_PATH))
    for name in declared_symbols():
        assert hasattr(lib, name), name
    lib = _capi.load()
    assert lib.snapy_kernel_count() > 0
    assert lib.snapy_last_error() == b""


@pytest.mark.skipif(not _capi.LIB_PATH.exists(), reason="library not built")
def test_host_argument_checks_need_no_device():
    lib = _capi.load()
    # a NULL grid is refused on the host before any launch
    assert lib.snapy_sweep(1, None, None, 0, None, None, None, None) < 0
    assert b"NULL" in lib.snapy_last_error()
