"""The C-ABI library loads and exports every entry point include/pmx.h declares
(no device work: this runs without a GPU)."""
import ctypes
import os
import re

from mymoonmd_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    with open(os.path.join(ROOT, "include", "pmx.h")) as f:
        text = f.read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(pmx_\w+)\s*\(", text, re.M)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 40
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_bindings_cover_the_header():
    assert set(declared_symbols()) == set(_lib.SIGNATURES)


def test_version_and_error_reporting_without_device():
    L = _lib.lib()
    assert L.pmx_version() >= 100
    h = ctypes.c_void_p()
    rc = L.pmx_domain_create(7, 1, ctypes.byref(h))      # unknown geometry
    assert rc == 1 and b"geometry" in L.pmx_last_error()
    rc = L.pmx_domain_create(2, 99, ctypes.byref(h))     # too many refinements
    assert rc == 1


def test_vector_arguments_are_checked_on_the_host():
    """Host memory and non-tensors are refused before the C ABI sees them."""
    import numpy as np
    import pytest
    import torch
    from mymoonmd_amd.parmoon import _ptr
    for bad in (torch.zeros(4, dtype=torch.float64), np.zeros(4), [0.0] * 4):
        with pytest.raises(ValueError):
            _ptr(bad, 4, 0, "x")
