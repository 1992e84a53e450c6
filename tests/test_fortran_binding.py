"""The Fortran interface module that INTEGRATION.md gives the reference's
side (ISO_C_BINDING) compiles with the image's Fortran compiler, its
derived type amrvac_fv_config has the C header's size and field offsets,
and a program calling the entry points links against the library.  No
GPU is needed: the program is linked, not run."""
from __future__ import annotations

import os
import re
import shutil
import subprocess

import pytest

from amrvac_amd import fv

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FLANG = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"

pytestmark = pytest.mark.skipif(not os.path.exists(FLANG), reason="no Fortran compiler in this image")


def module_source() -> str:
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    m = re.search(r"```fortran\n(module mod_amrvac_fv.*?end module mod_amrvac_fv)\n```", text, re.S)
    assert m, "INTEGRATION.md holds no Fortran interface module"
    return m.group(1)


def config_fields(src: str) -> list[str]:
    """Component names of the bind(C) type, in declaration order."""
    body = re.search(r"type, bind\(C\) :: amrvac_fv_config\n(.*?)\n\s*end type", src, re.S).group(1)
    names = []
    for line in body.splitlines():
        decl = line.split("!")[0]
        if "::" not in decl:
            continue
        for item in decl.split("::")[1].split(","):
            names.append(re.sub(r"\(.*\)", "", item).strip())
    return names


def c_layout(tmp_path) -> dict[str, int]:
    names = [f[0] for f in fv.FVConfig._fields_]
    src = tmp_path / "layout.c"
    body = "".join(f'  printf("{n} %zu\\n", offsetof(amrvac_fv_config, {n}));\n' for n in names)
    src.write_text("#include <stdio.h>\n#include <stddef.h>\n#include \"amrvac_fv.h\"\nint main(void) {\n"
                   '  printf("size %zu\\n", sizeof(amrvac_fv_config));\n' + body + "  return 0;\n}\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    return {k: int(v) for k, v in (line.split() for line in out.splitlines())}


def test_fortran_config_type_matches_the_header(tmp_path):
    src = module_source()
    fields = config_fields(src)
    assert fields == [f[0] for f in fv.FVConfig._fields_]
    prog = tmp_path / "layout.f90"
    lines = [src, "program layout", "  use iso_c_binding", "  use mod_amrvac_fv",
             "  type(amrvac_fv_config), target :: c", "  integer(c_intptr_t) :: b",
             "  b = transfer(c_loc(c), b)", '  print "(a,1x,i0)", "size", c_sizeof(c)']
    for n in fields:
        ref = f"c%{n}(1)" if n == "block_nx" else f"c%{n}"
        lines.append(f'  print "(a,1x,i0)", "{n}", transfer(c_loc({ref}), b) - b')
    lines.append("end program layout")
    prog.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout_f"
    subprocess.run([FLANG, "-o", str(exe), str(prog)], check=True, cwd=tmp_path)
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout
    got = {k: int(v) for k, v in (line.split() for line in out.splitlines())}
    assert got == c_layout(tmp_path)


def test_fortran_calls_link_against_the_library(tmp_path):
    """Every function of the module is called once in a program that links
    with libamrvac_fv.so (the symbols and their bind(C) names resolve)."""
    if not os.path.exists(fv.LIB_PATH):
        pytest.skip("library not built")
    src = module_source()
    funcs = re.findall(r"^\s*(?:integer\(c_int\)|type\(c_ptr\)) function (amrvac_fv_\w+)", src, re.M)
    assert set(funcs) <= set(fv.EXPORTS)
    assert {"amrvac_fv_init", "amrvac_fv_advect1", "amrvac_fv_getdt", "amrvac_fv_crash"} <= set(funcs)
    prog = tmp_path / "calls.f90"
    prog.write_text(src + """
program calls
  use iso_c_binding
  use mod_amrvac_fv
  type(amrvac_fv_config) :: cfg
  type(c_ptr) :: ctx, msg
  integer(c_int) :: rc, crash, ibuf(8)
  integer(c_int8_t) :: flags(1)
  real(c_double) :: dbuf(8), dt, cmax
  if (command_argument_count() < 99) stop 0   ! linked, not run
  rc = amrvac_fv_init(cfg, ctx)
  rc = amrvac_fv_finalize(ctx)
  msg = amrvac_fv_last_error()
  rc = amrvac_fv_set_mesh(ctx, 1, dbuf, ibuf, dbuf, ibuf)
  rc = amrvac_fv_upload(ctx, 0, dbuf)
  rc = amrvac_fv_download(ctx, 0, dbuf)
  rc = amrvac_fv_getdt(ctx, 0, 0.5d0, dt, cmax)
  rc = amrvac_fv_advect1(ctx, 1, 1d0, dt, cmax, 0, 1, -1, 0d0, 0d0, crash)
  rc = amrvac_fv_combine(ctx, 0, 1, ibuf, 0, dbuf)
  rc = amrvac_fv_copy(ctx, 1, 0)
  rc = amrvac_fv_crash(ctx, 1, crash)
  rc = amrvac_fv_volume_average(ctx, 0, dbuf)
  rc = amrvac_fv_getdt_state_noalloc(ctx, c_null_ptr, 0.5d0, dt, cmax)
  rc = amrvac_fv_set_mesh_amr(ctx, 1, dbuf, ibuf, ibuf, ibuf, ibuf, ibuf, dbuf, ibuf)
  rc = amrvac_fv_getbc(ctx, 0)
  rc = amrvac_fv_set_fix_conserve(ctx, 1)
  rc = amrvac_fv_errest(ctx, 0, 1, ibuf, dbuf, 1d-2, 0.1d0, 0.01d0, flags)
  rc = amrvac_fv_regrid(ctx, 1, ibuf, dbuf, ibuf, ibuf, ibuf, ibuf, ibuf, dbuf, ibuf)
end program calls
""")
    called = set(re.findall(r"= (amrvac_fv_\w+)\(", prog.read_text()))
    assert called == set(funcs)
    libdir = os.path.dirname(fv.LIB_PATH)
    exe = tmp_path / "calls"
    subprocess.run([FLANG, "-o", str(exe), str(prog), "-L" + libdir, "-lamrvac_fv", "-Wl,-rpath," + libdir],
                   check=True, cwd=tmp_path)
    assert exe.exists()
