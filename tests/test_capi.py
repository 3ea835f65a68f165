"""The C-ABI library loads and exports every symbol include/hawassi_hip.h
declares (no compute calls, runs without a GPU)."""
import ctypes as ct
import re
from pathlib import Path

import pytest

from hawassi_c_amd import _lib

ROOT = Path(__file__).resolve().parent.parent


def declared_symbols():
    text = (ROOT / "include" / "hawassi_hip.h").read_text()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(hw(?:1d)?_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_the_boundary():
    syms = declared_symbols()
    for s in ("hw_ctx_create_2d", "hw_rhs", "hw_evolve_apply", "hw_integrate", "hw_get_fields",
              "hw1d_create", "hw1d_rhs", "hw1d_integrate"):
        assert s in syms


def test_library_exports_every_declared_symbol():
    lib = _lib.load()
    missing = [s for s in declared_symbols() if not hasattr(lib, s)]
    assert not missing, missing
    assert set(_lib.EXPORTS) == set(declared_symbols())


def test_supported_sizes_and_version():
    lib = _lib.load()
    assert lib.hw_supported(2048, 2048, _lib.HW_PREC_FP32) == 1
    assert lib.hw_supported(1024, 1024, _lib.HW_PREC_FP32) == 1
    assert lib.hw_supported(64, 64, _lib.HW_PREC_FP64) == 1
    assert lib.hw_supported(100, 64, _lib.HW_PREC_FP64) == 0
    # the full-grid path: sides 2^a 3^b 5^c (the reference's 300x250, 2000x2)
    assert lib.hw_supported_full(100, 64, _lib.HW_PREC_FP64) == 1
    assert lib.hw_supported_full(300, 250, _lib.HW_PREC_FP32) == 1
    assert lib.hw_supported_full(2000, 2, _lib.HW_PREC_FP64) == 1
    assert lib.hw_supported_full(70, 64, _lib.HW_PREC_FP64) == 0
    assert b"gfx950" in lib.hw_version()


def test_bad_arguments_are_errors_not_crashes():
    lib = _lib.load()
    assert lib.hw_ctx_create_2d(None, None) == 1
    c = _lib.HwConfig2D()
    c.nx, c.ny, c.precision = 70, 64, 0     # 7 is a factor: no path takes it
    h = ct.c_void_p()
    assert lib.hw_ctx_create_2d(ct.byref(c), ct.byref(h)) == 1
    assert b"unsupported" in lib.hw_last_error()
    assert lib.hw_rhs(0.0, None, None, None) == 1


def test_missing_library_fails_loudly(tmp_path):
    with pytest.raises(_lib.HipPathMissing):
        _lib.load(tmp_path / "nope.so")
