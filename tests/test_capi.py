"""The C-ABI library loads and exports every symbol include/msceer.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

import pytest

from msceer_amd import native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            text = open(os.path.join(ROOT, "include", f)).read()
            text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
            names |= set(re.findall(r"\b(msceer_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_header_declares_entry_points():
    names = declared_symbols()
    assert "msceer_lstar_grad" in names and "msceer_create" in names


def test_library_exports_every_declared_symbol():
    if not os.path.exists(native.LIB_PATH):
        pytest.fail("libmsceer_hip.so not built (python -m msceer_amd.build)")
    lib = ctypes.CDLL(native.LIB_PATH)
    missing = [n for n in declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_every_declared_symbol():
    assert sorted(native.SIGNATURES) == declared_symbols()


def test_version_and_errors_without_gpu():
    lib = native.lib()
    assert lib.msceer_version() == 1
    assert lib.msceer_strerror(-1) == b"invalid argument"
    assert lib.msceer_lstar_grad(None, None, 4, 4, 4, None, 0) == -1
