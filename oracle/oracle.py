"""ctypes binding of the FP64 CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke`` and bench.py's CPU-baseline leg may
import this module; the product path never does.  ``OracleFV`` implements
the backend protocol of :mod:`amrvac_amd.host` on the C restatement in
amrvac_oracle.c.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liborc.so")

PHYS = {"hd": 1, "mhd": 2}
DIVB = {"none": 0, "glm": 1, "linde": 4, "ct": 8}
TYPE_CT = {"average": 1, "uct_contact": 2, "uct_hll": 3}


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int), ("ndir", ctypes.c_int), ("phys", ctypes.c_int),
        ("gamma", ctypes.c_double), ("divb", ctypes.c_int),
        ("glm_alpha", ctypes.c_double), ("glm_extended", ctypes.c_int),
        ("divbdiff", ctypes.c_double), ("limiter", ctypes.c_int),
        ("boundspeed", ctypes.c_int), ("tvdlfeps", ctypes.c_double),
        ("cada3_radius", ctypes.c_double), ("small_pressure", ctypes.c_double),
        ("small_density", ctypes.c_double), ("check_small_values", ctypes.c_int),
        ("nghost", ctypes.c_int), ("block_nx", ctypes.c_int * 3),
        ("type_ct", ctypes.c_int), ("fix_small_values", ctypes.c_int),
    ]


def build(force: bool = False) -> str:
    """Compile liborc.so with the Makefile beside this file."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-B" if force else "-s"], check=True,
                       stdout=subprocess.DEVNULL)
    return _LIB_PATH


_libs: dict = {}


def lib(path: str | None = None):
    """The oracle library (default: liborc.so beside this file; a different
    path loads another build of the same source, e.g. a planted mutant)."""
    path = path or _LIB_PATH
    if path not in _libs:
        if path == _LIB_PATH:
            build()
        L = ctypes.CDLL(path)
        vp = ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.orc_create.argtypes = [ctypes.POINTER(OrcParams), ctypes.c_int, dp, ip, dp, ip,
                                 ctypes.POINTER(vp)]
        L.orc_destroy.argtypes = [vp]
        L.orc_set_w.argtypes = [vp, ctypes.c_int, dp]
        L.orc_get_w.argtypes = [vp, ctypes.c_int, dp]
        L.orc_get_x.argtypes = [vp, dp]
        L.orc_getdt_courant.argtypes = [vp, ctypes.c_int, ctypes.c_double, dp, dp]
        L.orc_advect1.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_int, ctypes.c_int]
        L.orc_axpby.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                ctypes.c_double, ctypes.c_int]
        L.orc_copy.argtypes = [vp, ctypes.c_int, ctypes.c_int]
        L.orc_combine.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                                  ctypes.POINTER(ctypes.c_double)]
        L.orc_getbc.argtypes = [vp, ctypes.c_int]
        L.orc_volume_average.argtypes = [vp, ctypes.c_int, dp]
        L.orc_crash_flag.argtypes = [vp]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_dwlimiter.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, dp, dp]
        L.orc_weno5_face.argtypes = [dp, dp, dp]
        L.orc_face_flux.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, dp, ctypes.c_double, dp, dp, dp]
        L.orc_cell_cmax.argtypes = [vp, dp, ctypes.c_int]
        L.orc_cell_cmax.restype = ctypes.c_double
        L.orc_create_amr.argtypes = [ctypes.POINTER(OrcParams), ctypes.c_int, dp, ip, ip, ip, ip, dp, ip,
                                     ctypes.POINTER(vp)]
        L.orc_set_fix_conserve.argtypes = [vp, ctypes.c_int]
        L.orc_coarsen_child.argtypes = [vp, dp, ip, dp]
        L.orc_prolong_2nd.argtypes = [vp, dp, ip, dp]
        L.orc_lohner_error.argtypes = [vp, dp, dp, ctypes.c_double, dp]
        _libs[path] = L
    return _libs[path]


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _iptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def make_params(par, block_nx) -> OrcParams:
    from amrvac_amd.host import BOUNDSPEED, LIMITERS
    p = OrcParams()
    p.ndim = par.ndim
    p.ndir = par.ndir
    p.phys = PHYS[par.physics]
    p.gamma = par.gamma
    p.divb = DIVB[par.typedivbfix] if par.physics == "mhd" else 0
    p.glm_alpha = par.mhd_glm_alpha
    p.glm_extended = 1 if par.mhd_glm_extended else 0
    p.divbdiff = par.divbdiff
    p.limiter = LIMITERS[par.limiter]
    p.boundspeed = BOUNDSPEED[par.typeboundspeed]
    p.tvdlfeps = par.tvdlfeps
    p.cada3_radius = par.cada3_radius
    p.small_pressure = par.small_pressure
    p.small_density = par.small_density
    p.check_small_values = 1 if par.check_small_values else 0
    p.nghost = par.nghost
    for d in range(3):
        p.block_nx[d] = block_nx[d] if d < par.ndim else 1
    p.type_ct = TYPE_CT[par.type_ct] if par.physics == "mhd" and par.typedivbfix == "ct" else 0
    p.fix_small_values = 1 if par.fix_small_values else 0
    return p


class OracleFV:
    """Backend protocol of amrvac_amd.host on the C oracle."""

    def __init__(self, par, mesh, threads: int = 0, lib_path: str | None = None):
        self.par = par
        self.mesh = mesh
        self.L = lib(lib_path)
        self.L.orc_set_threads(threads)
        p = make_params(par, mesh.block_nx)
        h = ctypes.c_void_p()
        xl = np.ascontiguousarray(mesh.xlohi, dtype=np.float64)
        bc = np.ascontiguousarray(mesh.bc_codes, dtype=np.int32)
        if hasattr(mesh, "ntype"):  # AMR forest
            dxb = np.ascontiguousarray(mesh.dxb, dtype=np.float64)
            nt = np.ascontiguousarray(mesh.ntype, dtype=np.int32)
            nbr = np.ascontiguousarray(mesh.nbr, dtype=np.int32)
            nch = np.ascontiguousarray(mesh.nchild, dtype=np.int32)
            icp = np.ascontiguousarray(mesh.icpos, dtype=np.int32)
            rc = self.L.orc_create_amr(ctypes.byref(p), mesh.nblocks, _dptr(dxb), _iptr(nt), _iptr(nbr),
                                       _iptr(nch), _iptr(icp), _dptr(xl), _iptr(bc), ctypes.byref(h))
        else:
            dx = np.ascontiguousarray(mesh.dx, dtype=np.float64)
            nbr = np.ascontiguousarray(mesh.neighbor, dtype=np.int32)
            rc = self.L.orc_create(ctypes.byref(p), mesh.nblocks, _dptr(dx), _iptr(nbr), _dptr(xl),
                                   _iptr(bc), ctypes.byref(h))
        if rc != 0:
            raise RuntimeError(f"orc_create failed: {rc}")
        self.h = h
        self.nw = par.nw
        self.shape = (mesh.nblocks, par.nwt) + mesh.nxg[::-1]

    def __del__(self):
        if getattr(self, "h", None):
            self.L.orc_destroy(self.h)
            self.h = None

    def set_state(self, slot: int, w: np.ndarray) -> None:
        a = np.ascontiguousarray(w, dtype=np.float64)
        assert a.shape == self.shape, (a.shape, self.shape)
        self.L.orc_set_w(self.h, slot, _dptr(a))

    def get_state(self, slot: int) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        self.L.orc_get_w(self.h, slot, _dptr(out))
        return out

    def getbc(self, slot: int) -> None:
        self.L.orc_getbc(self.h, slot)

    def getdt_courant(self, slot: int, courantpar: float):
        dt = ctypes.c_double()
        cm = ctypes.c_double()
        self.L.orc_getdt_courant(self.h, slot, courantpar, ctypes.byref(dt), ctypes.byref(cm))
        return dt.value, cm.value

    def copy(self, dst: int, src: int) -> None:
        self.L.orc_copy(self.h, dst, src)

    def combine(self, dst: int, expr) -> None:
        expr.check()
        ops = np.ascontiguousarray(expr.ops, dtype=np.int32)
        k = np.ascontiguousarray(expr.consts + [0.0], dtype=np.float64)
        self.L.orc_combine(self.h, dst, len(expr.ops), _iptr(ops), _dptr(k))

    def advect1(self, method, dtfactor, dt, cmax_global, psa, psb, comb) -> None:
        if comb is not None:
            if comb[0] == "copy":
                self.L.orc_copy(self.h, psb, comb[1])
            else:
                _, a, x, b = comb
                self.L.orc_axpby(self.h, psb, a, x, b, psa)
        rc = self.L.orc_advect1(self.h, method, dtfactor, dt, cmax_global, psa, psb)
        if rc != 0:
            raise RuntimeError(f"orc_advect1 failed: {rc}")

    def volume_average(self, slot: int) -> np.ndarray:
        out = np.zeros(2 * self.nw, dtype=np.float64)
        self.L.orc_volume_average(self.h, slot, _dptr(out))
        return out

    def crash(self) -> bool:
        return bool(self.L.orc_crash_flag(self.h))

    # -- AMR -------------------------------------------------------------------
    def set_fix_conserve(self, on: bool) -> None:
        self.L.orc_set_fix_conserve(self.h, 1 if on else 0)

    def coarsen_child(self, wfi: np.ndarray, ic, wco: np.ndarray) -> None:
        a = np.ascontiguousarray(wfi, dtype=np.float64)
        icv = np.ascontiguousarray(ic, dtype=np.int32)
        assert wco.flags.c_contiguous and wco.dtype == np.float64
        self.L.orc_coarsen_child(self.h, _dptr(a), _iptr(icv), _dptr(wco))

    def prolong_2nd(self, wco: np.ndarray, ic, wfi: np.ndarray) -> None:
        a = np.ascontiguousarray(wco, dtype=np.float64)
        icv = np.ascontiguousarray(ic, dtype=np.int32)
        assert wfi.flags.c_contiguous and wfi.dtype == np.float64
        self.L.orc_prolong_2nd(self.h, _dptr(a), _iptr(icv), _dptr(wfi))

    def lohner_error(self, w: np.ndarray, weights: dict, wavefilter: float) -> np.ndarray:
        a = np.ascontiguousarray(w, dtype=np.float64)
        wt = np.zeros(self.nw)
        for k, v in weights.items():
            wt[k] = v
        nint = tuple(n - 2 * self.par.nghost for n in w.shape[1:])
        err = np.zeros(nint)
        self.L.orc_lohner_error(self.h, _dptr(a), _dptr(wt), wavefilter, _dptr(err))
        return err

    # -- single face / cell hooks for closed-form checks ----------------------
    def limiter(self, a: float, b: float, idim: int = 0):
        l, r = ctypes.c_double(), ctypes.c_double()
        self.L.orc_dwlimiter(self.h, idim, a, b, ctypes.byref(l), ctypes.byref(r))
        return l.value, r.value

    def face_flux(self, method: int, idim: int, prims: np.ndarray, ch: float = 0.0):
        """prims[k, iw] for the cells around one face (face between the
        middle two); returns (flux[nw], cmax, cmin)."""
        lw = np.ascontiguousarray(np.asarray(prims, dtype=np.float64).T)
        f = np.zeros(self.nw)
        cmax, cmin = ctypes.c_double(), ctypes.c_double()
        rc = self.L.orc_face_flux(self.h, method, idim, lw.shape[1], _dptr(lw), ch, _dptr(f),
                                  ctypes.byref(cmax), ctypes.byref(cmin))
        if rc != 0:
            raise RuntimeError("orc_face_flux failed")
        return f, cmax.value, cmin.value

    def cell_cmax(self, wcons: np.ndarray, idim: int) -> float:
        a = np.ascontiguousarray(wcons, dtype=np.float64)
        return self.L.orc_cell_cmax(self.h, _dptr(a), idim)


def weno5_face(w6, lib_path: str | None = None):
    a = np.ascontiguousarray(w6, dtype=np.float64)
    l, r = ctypes.c_double(), ctypes.c_double()
    lib(lib_path).orc_weno5_face(_dptr(a), ctypes.byref(l), ctypes.byref(r))
    return l.value, r.value
