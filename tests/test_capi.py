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


def test_config_struct_layout_matches_header(tmp_path):
    """The ctypes mirror of amrvac_fv_config has the header's size and field
    offsets (a C program compiled against include/amrvac_fv.h prints them)."""
    import subprocess
    names = [f[0] for f in fv.FVConfig._fields_]
    src = tmp_path / "layout.c"
    body = "".join(f'  printf("{n} %zu\\n", offsetof(amrvac_fv_config, {n}));\n' for n in names)
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"amrvac_fv.h\"\nint main(void) {\n"
                   '  printf("size %zu\\n", sizeof(amrvac_fv_config));\n' + body + "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = dict(line.split() for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                        check=True).stdout.splitlines())
    assert int(out["size"]) == ctypes.sizeof(fv.FVConfig)
    for n in names:
        assert int(out[n]) == getattr(fv.FVConfig, n).offset, n


def test_init_rejects_unsupported_constrained_transport():
    if not os.path.exists(fv.LIB_PATH):
        pytest.skip("library not built")
    L = fv.load_library()
    cfg = fv.FVConfig()
    cfg.ndim, cfg.ndir, cfg.physics, cfg.gamma = 3, 3, 2, 5.0 / 3.0
    cfg.limiter, cfg.boundspeed, cfg.nghost, cfg.precision = 9, 1, 2, 0
    cfg.block_nx[0], cfg.block_nx[1], cfg.block_nx[2] = 16, 16, 8
    cfg.divb, cfg.type_ct = 8, 2
    h = ctypes.c_void_p()
    # 3-D constrained transport needs 16^3 blocks
    assert L.amrvac_fv_init(ctypes.byref(cfg), ctypes.byref(h)) == -1
    assert b"block_nx = 16" in L.amrvac_fv_last_error()
    cfg.block_nx[2], cfg.ndir = 16, 2
    cfg.ndim = 2
    assert L.amrvac_fv_init(ctypes.byref(cfg), ctypes.byref(h)) == -1
    assert b"ndir = 3" in L.amrvac_fv_last_error()
    cfg.ndir, cfg.block_nx[2], cfg.type_ct = 3, 1, 4
    assert L.amrvac_fv_init(ctypes.byref(cfg), ctypes.byref(h)) == -1
    assert b"type_ct" in L.amrvac_fv_last_error()
