"""The C-ABI library loads and exports every entry point include/amrvac_fv.h
declares (no GPU needed; nothing here launches a kernel)."""
import ctypes
import os
import re

import pytest

from amrvac_amd import fv

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    hdr = open(os.path.join(ROOT, "include", "amrvac_fv.h")).read()
    return sorted(set(re.findall(r"\b(amrvac_fv_\w+)\s*\(", hdr)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(fv.EXPORTS)


def test_library_exports_every_symbol():
    if not os.path.exists(fv.LIB_PATH):
        pytest.skip("library not built")
    lib = ctypes.CDLL(fv.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_init_rejects_bad_configs_without_touching_the_gpu():
    if not os.path.exists(fv.LIB_PATH):
        pytest.skip("library not built")
    L = fv.load_library()
    cfg = fv.FVConfig()
    cfg.ndim, cfg.ndir, cfg.physics, cfg.gamma = 2, 2, 1, 1.4
    cfg.limiter, cfg.boundspeed, cfg.nghost, cfg.precision = 5, 1, 2, 0
    cfg.block_nx[0], cfg.block_nx[1], cfg.block_nx[2] = 8, 16, 1
    h = ctypes.c_void_p()
    assert L.amrvac_fv_init(ctypes.byref(cfg), ctypes.byref(h)) == -1
    assert b"block_nx" in L.amrvac_fv_last_error()
    cfg.block_nx[0] = 16
    cfg.limiter = 16  # weno5 needs 3 ghost cells
    assert L.amrvac_fv_init(ctypes.byref(cfg), ctypes.byref(h)) == -1
    assert b"weno5" in L.amrvac_fv_last_error()
