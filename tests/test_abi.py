"""CPU-side checks of the C ABI library: it loads, and exports every entry point that
include/sx_hmmer.h declares (no device calls)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "sx_hmmer.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sx_[a-z0-9_]+)\s*\(", txt)))


def test_header_declares_the_binding_list(sxa):
    assert declared_symbols() == sorted(sxa.EXPORTS)


def test_library_loads_and_exports_every_declared_symbol(sxa):
    if not os.path.exists(sxa.LIB_PATH):
        pytest.skip("libsxhmmer.so not built")
    lib = ctypes.CDLL(sxa.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_create_without_device_fails_loudly(sxa):
    """On a host with no HIP device the context refuses to start (no CPU fallback)."""
    if not os.path.exists(sxa.LIB_PATH):
        pytest.skip("libsxhmmer.so not built")
    import torch
    if torch.cuda.is_available():
        pytest.skip("a device is present")
    with pytest.raises(sxa.SxError):
        sxa.Context(0)
