"""ctypes binding of libamrvac_fv.so (include/amrvac_fv.h).

``DeviceFV`` implements the backend protocol of :mod:`amrvac_amd.host`
(getdt_courant / advect1 / volume_average / ...) on the HIP kernels, so the
same reference-shaped time loop drives either the device path or, in tests,
the CPU oracle.  There is no CPU fallback: constructing a ``DeviceFV``
without the built library or without a GPU raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from .host import BOUNDSPEED, DIVB, LIMITERS, TYPE_CT, Par
from .mesh import UniformMesh

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libamrvac_fv.so")

PREC = {"fp64": 0, "mixed": 1, "fp64_exact": 2}

# every symbol include/amrvac_fv.h declares
EXPORTS = ("amrvac_fv_init", "amrvac_fv_finalize", "amrvac_fv_set_stream", "amrvac_fv_set_mesh",
           "amrvac_fv_upload", "amrvac_fv_download", "amrvac_fv_getdt", "amrvac_fv_advect1",
           "amrvac_fv_copy", "amrvac_fv_combine", "amrvac_fv_volume_average", "amrvac_fv_slot_pointer",
           "amrvac_fv_sync", "amrvac_fv_sweep_timing", "amrvac_fv_last_error",
           "amrvac_fv_set_mesh_amr", "amrvac_fv_getbc", "amrvac_fv_set_fix_conserve", "amrvac_fv_adjoint1",
           "amrvac_fv_slot_copy", "amrvac_fv_stage", "amrvac_fv_getdt_state", "amrvac_fv_errest",
           "amrvac_fv_regrid", "amrvac_fv_crash", "amrvac_fv_getdt_state_noalloc")


class FVConfig(ctypes.Structure):
    _fields_ = [
        ("ndim", ctypes.c_int), ("ndir", ctypes.c_int), ("physics", ctypes.c_int),
        ("gamma", ctypes.c_double), ("divb", ctypes.c_int), ("glm_alpha", ctypes.c_double),
        ("glm_extended", ctypes.c_int), ("divbdiff", ctypes.c_double), ("limiter", ctypes.c_int),
        ("boundspeed", ctypes.c_int), ("tvdlfeps", ctypes.c_double), ("cada3_radius", ctypes.c_double),
        ("small_pressure", ctypes.c_double), ("small_density", ctypes.c_double),
        ("check_small_values", ctypes.c_int), ("nghost", ctypes.c_int),
        ("block_nx", ctypes.c_int * 3), ("precision", ctypes.c_int), ("type_ct", ctypes.c_int),
        ("fix_small_values", ctypes.c_int),
    ]


_lib = None


def load_library(path: str = LIB_PATH) -> ctypes.CDLL:
    """Load the in-tree library (build it with ``amrvac_amd.build``)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is missing: run `python -m amrvac_amd.build` first")
    # One HIP runtime per process: PyTorch ships its own libamdhip64 and
    # libhsa-runtime64.  Loaded first, it also serves this library (same
    # sonames); loaded after /opt/rocm's, both runtimes would claim the
    # device and PyTorch's fails to initialise.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(path)
    vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int)
    L.amrvac_fv_init.argtypes = [ctypes.POINTER(FVConfig), ctypes.POINTER(vp)]
    L.amrvac_fv_finalize.argtypes = [vp]
    L.amrvac_fv_set_stream.argtypes = [vp, vp]
    L.amrvac_fv_set_mesh.argtypes = [vp, ctypes.c_int, dp, ip, dp, ip]
    L.amrvac_fv_upload.argtypes = [vp, ctypes.c_int, dp]
    L.amrvac_fv_download.argtypes = [vp, ctypes.c_int, dp]
    L.amrvac_fv_getdt.argtypes = [vp, ctypes.c_int, ctypes.c_double, dp, dp]
    L.amrvac_fv_advect1.argtypes = [vp, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                    ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                    ctypes.c_double, ip]
    L.amrvac_fv_copy.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    L.amrvac_fv_combine.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_double)]
    L.amrvac_fv_volume_average.argtypes = [vp, ctypes.c_int, dp]
    L.amrvac_fv_slot_pointer.argtypes = [vp, ctypes.c_int, ctypes.POINTER(vp), ctypes.POINTER(ctypes.c_longlong)]
    L.amrvac_fv_sync.argtypes = [vp]
    L.amrvac_fv_sweep_timing.argtypes = [vp, ctypes.c_int, dp, ip]
    L.amrvac_fv_last_error.restype = ctypes.c_char_p
    L.amrvac_fv_set_mesh_amr.argtypes = [vp, ctypes.c_int, dp, ip, ip, ip, ip, ip, dp, ip]
    L.amrvac_fv_getbc.argtypes = [vp, ctypes.c_int]
    L.amrvac_fv_set_fix_conserve.argtypes = [vp, ctypes.c_int]
    dbl = ctypes.c_double
    L.amrvac_fv_adjoint1.argtypes = [vp, ctypes.c_int, dbl, dbl, dbl, vp, vp, dbl, vp, dbl, vp, dbl, vp]
    L.amrvac_fv_stage.argtypes = [vp, ctypes.c_int, dbl, dbl, dbl, vp, vp, dbl, dbl, vp]
    L.amrvac_fv_getdt_state.argtypes = [vp, vp, dbl, ctypes.POINTER(dbl), ctypes.POINTER(dbl)]
    L.amrvac_fv_errest.argtypes = [vp, ctypes.c_int, ctypes.c_int, ip, dp, dbl, dbl, dbl,
                                   ctypes.POINTER(ctypes.c_ubyte)]
    L.amrvac_fv_regrid.argtypes = [vp, ctypes.c_int, ip, dp, ip, ip, ip, ip, ip, dp, ip]
    L.amrvac_fv_slot_copy.argtypes = [vp, ctypes.c_int, vp, ctypes.c_int]
    L.amrvac_fv_crash.argtypes = [vp, ctypes.c_int, ip]
    L.amrvac_fv_getdt_state_noalloc.argtypes = [vp, vp, dbl, ctypes.POINTER(dbl), ctypes.POINTER(dbl)]
    _lib = L
    return L


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))


def make_config(par: Par, mesh: UniformMesh, precision: str | None = None) -> FVConfig:
    c = FVConfig()
    c.ndim = par.ndim
    c.ndir = par.ndir
    c.physics = 1 if par.physics == "hd" else 2
    c.gamma = par.gamma
    c.divb = DIVB[par.typedivbfix] if par.physics == "mhd" else 0
    c.glm_alpha = par.mhd_glm_alpha
    c.glm_extended = 1 if par.mhd_glm_extended else 0
    c.divbdiff = par.divbdiff
    c.limiter = LIMITERS[par.limiter]
    c.boundspeed = BOUNDSPEED[par.typeboundspeed]
    c.tvdlfeps = par.tvdlfeps
    c.cada3_radius = par.cada3_radius
    c.small_pressure = par.small_pressure
    c.small_density = par.small_density
    c.check_small_values = 1 if par.check_small_values else 0
    c.nghost = par.nghost
    for d in range(3):
        c.block_nx[d] = mesh.block_nx[d] if d < par.ndim else 1
    c.precision = PREC[precision or par.precision]
    c.type_ct = TYPE_CT[par.type_ct] if par.stagger else 0
    c.fix_small_values = 1 if par.fix_small_values else 0
    return c


class FVError(RuntimeError):
    pass


class DeviceFV:
    """Backend protocol of amrvac_amd.host on the HIP kernels."""

    def __init__(self, par: Par, mesh: UniformMesh, precision: str | None = None):
        self.L = load_library()
        self.par = par
        self.mesh = mesh
        self.nw = par.nw
        self.shape = (mesh.nblocks, par.nwt) + mesh.nxg[::-1]
        cfg = make_config(par, mesh, precision)
        h = ctypes.c_void_p()
        self._check(self.L.amrvac_fv_init(ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h
        xl = np.ascontiguousarray(mesh.xlohi, dtype=np.float64)
        bc = np.ascontiguousarray(mesh.bc_codes, dtype=np.int32)
        self.amr = hasattr(mesh, "ntype")
        if self.amr:  # an AMR forest (amrvac_amd.amr.AMRForest)
            dx1 = np.ascontiguousarray(mesh.dx[1], dtype=np.float64)
            lev = np.ascontiguousarray(mesh.level, dtype=np.int32)
            nt = np.ascontiguousarray(mesh.ntype, dtype=np.int32)
            nbr = np.ascontiguousarray(mesh.nbr, dtype=np.int32)
            nch = np.ascontiguousarray(mesh.nchild, dtype=np.int32)
            icp = np.ascontiguousarray(mesh.icpos, dtype=np.int32)
            self._check(self.L.amrvac_fv_set_mesh_amr(self.h, mesh.nblocks, _dp(dx1), _ip(lev), _ip(nt), _ip(nbr),
                                                      _ip(nch), _ip(icp), _dp(xl), _ip(bc)))
        else:
            dx = np.ascontiguousarray(mesh.dx, dtype=np.float64)
            nbr = np.ascontiguousarray(mesh.neighbor, dtype=np.int32)
            self._check(self.L.amrvac_fv_set_mesh(self.h, mesh.nblocks, _dp(dx), _ip(nbr), _dp(xl), _ip(bc)))

    def _check(self, rc: int) -> None:
        if rc != 0:
            msg = self.L.amrvac_fv_last_error().decode()
            raise FVError(f"amrvac_fv error {rc}: {msg}")

    def close(self) -> None:
        if getattr(self, "h", None):
            self.L.amrvac_fv_finalize(self.h)
            self.h = None

    def __del__(self):
        self.close()

    # -- backend protocol ---------------------------------------------------
    def set_state(self, slot: int, w: np.ndarray) -> None:
        a = np.ascontiguousarray(w, dtype=np.float64)
        if a.shape != self.shape:
            raise ValueError(f"state shape {a.shape} != {self.shape}")
        self._check(self.L.amrvac_fv_upload(self.h, slot, _dp(a)))

    def get_state(self, slot: int) -> np.ndarray:
        out = np.empty(self.shape, dtype=np.float64)
        self._check(self.L.amrvac_fv_download(self.h, slot, _dp(out)))
        return out

    def getbc(self, slot: int) -> None:
        """On a uniform level ghost cells are read from the neighbours by
        every sweep and materialised by get_state (nothing to do); on an AMR
        mesh the ghost cells across level jumps are filled on the device."""
        self._check(self.L.amrvac_fv_getbc(self.h, slot))

    # -- device arrays (zero-copy plumbing for PyTorch) ------------------------
    def set_stream(self, stream_handle: int | None) -> None:
        """Order all work of this context on the given hipStream_t (e.g.
        torch.cuda.current_stream().cuda_stream); None restores its own."""
        self._check(self.L.amrvac_fv_set_stream(self.h, ctypes.c_void_p(stream_handle) if stream_handle else None))

    def slot_to(self, slot: int, dptr: int) -> None:
        self._check(self.L.amrvac_fv_slot_copy(self.h, slot, ctypes.c_void_p(dptr), 0))

    def slot_from(self, slot: int, dptr: int) -> None:
        self._check(self.L.amrvac_fv_slot_copy(self.h, slot, ctypes.c_void_p(dptr), 1))

    def adjoint1(self, method: int, dtfactor: float, dt: float, u_in: int, lam_out: int, coef: float,
                 lam_in: int, add_a: int = 0, ca: float = 0.0, add_b: int = 0, cb: float = 0.0,
                 ch: float = 0.0) -> None:
        """lam_in = coef*lam_out + ca*add_a + cb*add_b + (d stage/d u_in)^T lam_out
        (device pointers; 0 for an absent add term); ch: the GLM speed of
        the step."""
        v = ctypes.c_void_p
        self._check(self.L.amrvac_fv_adjoint1(self.h, method, dtfactor, dt, ch, v(u_in), v(lam_out), coef,
                                              v(add_a) if add_a else None, ca, v(add_b) if add_b else None, cb,
                                              v(lam_in)))

    def stage(self, method: int, dtfactor: float, dt: float, cmax_global: float, u_in: int, u_out: int,
              base: int = 0, comb_a: float = 0.0, comb_b: float = 0.0) -> None:
        """u_out = (comb_a*base + comb_b*u_in, or u_in) + qdt L(u_in) on device pointers."""
        v = ctypes.c_void_p
        self._check(self.L.amrvac_fv_stage(self.h, method, dtfactor, dt, cmax_global, v(u_in),
                                           v(base) if base else None, comb_a, comb_b, v(u_out)))

    def getdt_state(self, u: int, courantpar: float):
        dt = ctypes.c_double()
        cm = ctypes.c_double()
        self._check(self.L.amrvac_fv_getdt_state(self.h, ctypes.c_void_p(u), courantpar, ctypes.byref(dt),
                                                 ctypes.byref(cm)))
        return dt.value, cm.value

    # -- regrid on the device (AMR forests) ------------------------------------
    @property
    def amr_device_regrid(self) -> bool:
        return self.amr

    def errest_flags(self, forest, ap):
        """errest (Löhner) of slot 0 on the device: (refine, coarsen) flags
        per leaf with the level limits of errest applied; proper nesting
        is the caller's."""
        weights = sorted((iw, w) for iw, w in ap.w_refine_weight.items() if w != 0.0)
        iws = np.ascontiguousarray([iw for iw, _ in weights], dtype=np.int32)
        wts = np.ascontiguousarray([w for _, w in weights], dtype=np.float64)
        thr = ap.refine_threshold
        flags = np.zeros(forest.nblocks, dtype=np.uint8)
        self._check(self.L.amrvac_fv_errest(self.h, 0, len(iws), _ip(iws), _dp(wts), ap.amr_wavefilter, thr,
                                            ap.derefine_ratio * thr,
                                            flags.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte))))
        lev = np.asarray(forest.level)
        refine = ((flags & 1) != 0) & (lev < ap.refine_max_level)
        coarsen = ((flags & 2) != 0) & (lev > 1)
        return refine, coarsen

    def regrid(self, plan, forest) -> None:
        """Move slot 0 to the forest's new leaves on the device (plan from
        amr.apply_regrid), then install the new connectivity and ghost
        cells."""
        nd = forest.ndim
        nc = 1 << nd
        pad = [0] * (nc - 1)
        rows = []
        for org in plan:
            if org[0] == "copy":
                rows.append([0, org[1]] + pad + [0])
            elif org[0] == "coarsen":
                rows.append([1] + list(org[1]) + [0])
            else:
                rows.append([2, org[1]] + pad + [sum((org[2][d] - 1) << d for d in range(nd))])
        P = np.array(rows, dtype=np.int32).reshape(forest.nblocks, 2 + nc)
        xl = np.ascontiguousarray(forest.xlohi, dtype=np.float64)
        bc = np.ascontiguousarray(forest.bc_codes, dtype=np.int32)
        dx1 = np.ascontiguousarray(forest.dx[1], dtype=np.float64)
        lev = np.ascontiguousarray(forest.level, dtype=np.int32)
        nt = np.ascontiguousarray(forest.ntype, dtype=np.int32)
        nbr = np.ascontiguousarray(forest.nbr, dtype=np.int32)
        nch = np.ascontiguousarray(forest.nchild, dtype=np.int32)
        icp = np.ascontiguousarray(forest.icpos, dtype=np.int32)
        self._check(self.L.amrvac_fv_regrid(self.h, forest.nblocks, _ip(np.ascontiguousarray(P)), _dp(dx1),
                                            _ip(lev), _ip(nt), _ip(nbr), _ip(nch), _ip(icp), _dp(xl), _ip(bc)))
        self.mesh = forest
        self.shape = (forest.nblocks, self.nw) + forest.nxg[::-1]

    def set_fix_conserve(self, on: bool) -> None:
        self._check(self.L.amrvac_fv_set_fix_conserve(self.h, 1 if on else 0))

    def getdt_courant(self, slot: int, courantpar: float):
        dt = ctypes.c_double()
        cm = ctypes.c_double()
        self._check(self.L.amrvac_fv_getdt(self.h, slot, courantpar, ctypes.byref(dt), ctypes.byref(cm)))
        return dt.value, cm.value

    def copy(self, dst: int, src: int) -> None:
        self._check(self.L.amrvac_fv_copy(self.h, dst, src))

    def combine(self, dst: int, expr) -> None:
        """ps[dst] = expr (an amrvac_amd.host.Expr over the slots)."""
        expr.check()
        ops = np.ascontiguousarray(expr.ops, dtype=np.int32)
        k = np.ascontiguousarray(expr.consts + [0.0], dtype=np.float64)
        self._check(self.L.amrvac_fv_combine(self.h, dst, len(expr.ops), _ip(ops), len(expr.consts), _dp(k)))

    def advect1(self, method, dtfactor, dt, cmax_global, psa, psb, comb) -> None:
        if comb is None:
            cs, ca, cb = -1, 0.0, 0.0
        elif comb[0] == "copy":
            cs, ca, cb = comb[1], 1.0, 0.0
        else:
            _, ca, cs, cb = comb
        self._check(self.L.amrvac_fv_advect1(self.h, method, dtfactor, dt, cmax_global, psa, psb, cs, ca, cb,
                                             None))

    def volume_average(self, slot: int) -> np.ndarray:
        out = np.zeros(2 * self.nw, dtype=np.float64)
        self._check(self.L.amrvac_fv_volume_average(self.h, slot, _dp(out)))
        return out

    def crash(self, clear: bool = False) -> bool:
        """The device's small-value flag (the reference's crash): set by any
        stage or dt reduction since it was last cleared."""
        flag = ctypes.c_int(0)
        self._check(self.L.amrvac_fv_crash(self.h, 1 if clear else 0, ctypes.byref(flag)))
        return bool(flag.value)

    # -- extras ---------------------------------------------------------------
    def sync(self) -> None:
        self._check(self.L.amrvac_fv_sync(self.h))

    def slot_pointer(self, slot: int) -> tuple[int, int]:
        p = ctypes.c_void_p()
        n = ctypes.c_longlong()
        self._check(self.L.amrvac_fv_slot_pointer(self.h, slot, ctypes.byref(p), ctypes.byref(n)))
        return int(p.value), int(n.value)

    def sweep_timing(self, reset: bool = False) -> tuple[float, int]:
        ms = ctypes.c_double()
        cnt = ctypes.c_int()
        self._check(self.L.amrvac_fv_sweep_timing(self.h, 1 if reset else 0, ctypes.byref(ms), ctypes.byref(cnt)))
        return ms.value, cnt.value
