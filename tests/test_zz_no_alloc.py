"""amrvac_fv_getdt_state_noalloc (the allocation-free dt reduction, scratch
in a file-scope device array) against amrvac_fv_getdt_state (the context's
own buffers), through the C-ABI: identical dt and cmax on two calls in a
row, the second on a smaller level; the scratch array keeps the first
call's partial maxima, so a stale entry would show in the second."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from amrvac_amd import problems
from amrvac_amd.host import Par
from amrvac_amd.mesh import UniformMesh

pytestmark = pytest.mark.gpu


def _dt_both(be):
    from amrvac_amd.fv import load_library
    L = load_library()
    ptr, _ = be.slot_pointer(0)
    out = {}
    for name in ("amrvac_fv_getdt_state", "amrvac_fv_getdt_state_noalloc"):
        dt, cm = ctypes.c_double(), ctypes.c_double()
        rc = getattr(L, name)(be.h, ctypes.c_void_p(ptr), 0.6, ctypes.byref(dt), ctypes.byref(cm))
        assert rc == 0, L.amrvac_fv_last_error()
        out[name] = (dt.value, cm.value)
    return out


def test_noalloc_getdt_matches_getdt_state_twice():
    from amrvac_amd.fv import DeviceFV
    par = Par(physics="mhd", ndim=3, ndir=3, gamma=5.0 / 3.0, typedivbfix="glm", time_stepper="threestep",
              flux_scheme="hll", limiter="cada3", courantpar=0.6)
    results = []
    # the larger level first, then a smaller one whose fastest cells lie
    # elsewhere: the second call reads only its own blocks' partials
    for n, amp in ((64, 1.0), (32, 0.25)):
        print(f"[noalloc] OT 3-D {n}^3", flush=True)
        mesh = UniformMesh(3, (n, n, n), (16, 16, 16), (0.0,) * 3, (1.0,) * 3, ("periodic",) * 6)
        be = DeviceFV(par, mesh)
        be.set_state(0, problems.orszag_tang_3d(mesh, b_amp=amp))
        r = _dt_both(be)
        assert r["amrvac_fv_getdt_state_noalloc"] == r["amrvac_fv_getdt_state"]
        results.append(r["amrvac_fv_getdt_state"])
        del be
    assert results[0] != results[1]


def test_noalloc_getdt_rejects_too_many_blocks():
    from amrvac_amd.fv import DeviceFV, load_library
    par = Par(physics="hd", ndim=2, ndir=2, gamma=1.4, flux_scheme="hll", limiter="minmod")
    mesh = UniformMesh(2, (16 * 256, 16 * 129), (16, 16), (0.0, 0.0), (1.0, 1.0), ("periodic",) * 4)
    assert mesh.nblocks == 33024
    be = DeviceFV(par, mesh)
    L = load_library()
    ptr, _ = be.slot_pointer(0)
    dt, cm = ctypes.c_double(), ctypes.c_double()
    rc = L.amrvac_fv_getdt_state_noalloc(be.h, ctypes.c_void_p(ptr), 0.6, ctypes.byref(dt), ctypes.byref(cm))
    assert rc == -1
    assert b"32768" in L.amrvac_fv_last_error()
