"""The C ABI library loads on a CPU-only host and exports every symbol that
include/mlp.h declares (no compute calls here)."""
import ctypes
import os
import re

import numpy as np
import pytest

from ml_prec_amd import native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "mlp.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(mlp_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(N.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 30
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_symbols()) == set(N.EXPORTS), set(declared_symbols()) ^ set(N.EXPORTS)


def test_errors_cross_the_abi_as_status_codes():
    lib = N.lib()
    h = ctypes.c_void_p()
    ia = np.array([0, 1, 2], np.uint32)
    ja = np.array([1, 0], np.uint32)            # not Skyline: row 0 does not start with its diagonal
    a = np.array([1.0, 1.0])
    rc = lib.mlp_hier_build_uh_cheb(2, ia, ja, a, np.array([3.0]), 1, np.array([2], np.uint32), 1, 1,
                                    ctypes.byref(h))
    assert rc == 1
    assert b"diagonal" in lib.mlp_last_error()
