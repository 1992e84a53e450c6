"""Host-side time integration that drives a finite-volume backend.

This mirrors the reference's host loop: ``timeintegration``
(src/amrvac.t:189-400), ``setdt`` (src/mod_dt.t:10-288), ``advance`` /
``advect`` (src/mod_advance.t:17-567) and the regression log
(src/io/mod_input_output.t:2653-2729).  The per-stage work, ``advect1``
(src/mod_advance.t:611-666), is delegated to a backend object; the product
backend is :class:`amrvac_amd.fv.DeviceFV`, which calls the HIP kernels
through the C-ABI declared in include/amrvac_fv.h.

Backend protocol (all slots are 0=ps, 1=ps1, 2=ps2):
  getdt_courant(slot, courantpar) -> (dtmin, cmax)
  advect1(method, dtfactor, dt, cmax_global, psa, psb, comb)
      comb = None                 psb holds the initial wnew
           | ("copy", src)        psb = ps[src] first (ps1=ps in advect)
           | ("axpby", a, x, b)   psb = a*ps[x] + b*ps[psa] first
  volume_average(slot) -> ndarray[2*nw]
  crash() -> bool
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from .mesh import UniformMesh

SMALLDOUBLE = 1.0e-12  # src/mod_constants.t:8
BIGDOUBLE = 1.0e99     # src/mod_constants.t:11

# flux scheme codes, src/mod_global_parameters.t:457-469
FS = {"hll": 1, "hllc": 2, "tvdlf": 5, "hancock": 8}
# limiter codes, src/limiter/mod_limiter.t:15-40
LIMITERS = {"minmod": 1, "woodward": 2, "mcbeta": 3, "superbee": 4,
            "vanleer": 5, "albada": 6, "koren": 7, "cada": 8, "cada3": 9,
            "weno5": 16}
BOUNDSPEED = {"Einfeldt": 1, "cmaxmean": 2, "cmaxleftright": 3}
DIVB = {"none": 0, "glm": 1, "linde": 4}


@dataclass
class Par:
    """The subset of the reference's .par namelists that the block update
    reads (defaults from src/io/mod_input_output.t:292-537,
    src/hd/mod_hd_phys.t, src/mhd/mod_mhd_phys.t)."""

    physics: str = "hd"            # 'hd' | 'mhd'
    ndim: int = 2
    ndir: int = 2
    gamma: float = 5.0 / 3.0
    typedivbfix: str = "linde"     # mhd_list
    mhd_glm_alpha: float = 0.5
    mhd_glm_extended: bool = True
    divbdiff: float = 0.8
    time_stepper: str = "twostep"  # 'onestep' | 'twostep' | 'threestep'
    flux_scheme: str = "tvdlf"
    limiter: str = "minmod"
    typeboundspeed: str = "Einfeldt"
    tvdlfeps: float = 1.0
    cada3_radius: float = 0.1
    courantpar: float = 0.8
    slowsteps: int = -1
    dtmin: float = 1.0e-10
    time_max: float = BIGDOUBLE
    it_max: int = 2 ** 62
    dtsave_log: float = BIGDOUBLE
    small_pressure: float = 0.0
    small_density: float = 0.0
    check_small_values: bool = True
    precision: str = "fp64"        # 'fp64' | 'mixed' (device backend only)

    @property
    def nghost(self) -> int:
        # the reference raises nghostcells for wide stencils,
        # src/io/mod_input_output.t:1420-1422
        return 3 if self.limiter == "weno5" else 2

    @property
    def nw(self) -> int:
        if self.physics == "hd":
            return self.ndir + 2
        return 2 * self.ndir + 2 + (1 if self.typedivbfix == "glm" else 0)

    @property
    def need_global_cmax(self) -> bool:
        return self.physics == "mhd" and self.typedivbfix == "glm"

    @property
    def typepred1(self) -> str:
        # src/io/mod_input_output.t:788-810
        return "hancock" if self.flux_scheme == "tvdlf" else self.flux_scheme

    @property
    def nstages(self) -> int:
        return {"onestep": 1, "twostep": 2, "threestep": 3}[self.time_stepper]


def ssprk3_coefficients() -> dict[str, float]:
    """SSPRK(3,3), src/io/mod_input_output.t:941-964."""
    c = dict(beta11=1.0, beta22=1.0 / 4.0, beta33=2.0 / 3.0,
             alfa21=3.0 / 4.0, alfa31=1.0 / 3.0)
    c["alfa22"] = 1.0 - c["alfa21"]
    c["alfa33"] = 1.0 - c["alfa31"]
    return c


@dataclass
class RunState:
    it: int = 0
    global_time: float = 0.0
    dt: float = 0.0
    cmax_global: float = 0.0
    crash: bool = False
    final_dt_exit: bool = False
    tsavelast_log: float = 0.0
    log: list = field(default_factory=list)
    dts: list = field(default_factory=list)


class TimeIntegrator:
    """Reference-shaped time loop over a backend (see module docstring)."""

    def __init__(self, par: Par, mesh: UniformMesh, backend):
        self.par = par
        self.mesh = mesh
        self.be = backend
        self.state = RunState()
        self.it_init = 0

    # -- setdt, src/mod_dt.t:10-135 ---------------------------------------
    def setdt(self) -> float:
        par, st = self.par, self.state
        dtmin_mype, cmax = self.be.getdt_courant(0, par.courantpar)
        if not math.isfinite(dtmin_mype):
            st.crash = True
            raise RuntimeError("Error: non-finite time step, run crash.")
        if dtmin_mype < par.dtmin:
            st.crash = True
        if par.slowsteps > st.it - self.it_init + 1:
            factor = 1.0 - (1.0 - float(st.it - self.it_init + 1) / float(par.slowsteps)) ** 2
            dtmin_mype = dtmin_mype * factor
        # final_dt_reduction (default .true. without local timestep)
        if par.time_max - st.global_time <= par.dtmin:
            st.final_dt_exit = True
        dtmin_mype = min(dtmin_mype, par.time_max - st.global_time)
        dt = dtmin_mype
        dtsave = [par.dtsave_log] + [BIGDOUBLE] * 4
        if any(d < BIGDOUBLE for d in dtsave):
            dtmax = min(float(math.ceil(st.global_time / d)) * d for d in dtsave) - st.global_time
            if dtmax > SMALLDOUBLE:
                dt = min(dt, dtmax)
            else:
                dt = min(dt, min(dtsave))
        st.dt = dt
        if par.need_global_cmax:
            st.cmax_global = cmax
        return dt

    # -- advect, src/mod_advance.t:56-567 ------------------------------------
    def advect(self) -> None:
        par, st = self.par, self.state
        fm = FS[par.flux_scheme]
        dt, ch = st.dt, st.cmax_global
        if par.time_stepper == "threestep":
            c = ssprk3_coefficients()
            self.be.advect1(fm, c["beta11"], dt, ch, 0, 1, ("copy", 0))
            self.be.advect1(fm, c["beta22"], dt, ch, 1, 2, ("axpby", c["alfa21"], 0, c["alfa22"]))
            self.be.advect1(fm, c["beta33"], dt, ch, 2, 0, ("axpby", c["alfa31"], 0, c["alfa33"]))
        elif par.time_stepper == "twostep":
            # Predictor_Corrector: predictor without refluxing, :104-111
            self.be.advect1(FS[par.typepred1], 0.5, dt, ch, 0, 1, ("copy", 0))
            self.be.advect1(fm, 1.0, dt, ch, 1, 0, None)
        elif par.time_stepper == "onestep":
            self.be.copy(1, 0)
            self.be.advect1(fm, 1.0, dt, ch, 1, 0, None)
        else:
            raise ValueError(f"unsupported time_stepper {par.time_stepper!r}")

    # -- timetosave for the log file, src/amrvac.t:405-441 ------------------
    def _log_due(self) -> bool:
        st = self.state
        if st.global_time >= 0.0 - SMALLDOUBLE:
            if st.global_time >= st.tsavelast_log + self.par.dtsave_log - SMALLDOUBLE:
                st.tsavelast_log = st.global_time
                return True
        return False

    def log_line(self) -> np.ndarray:
        modes = self.be.volume_average(0)
        row = np.concatenate([[self.state.global_time], modes])
        self.state.log.append(row)
        return row

    # -- timeintegration, src/amrvac.t:189-355 ------------------------------
    def run(self, nsteps: int | None = None) -> RunState:
        par, st = self.par, self.state
        steps = 0
        while True:
            self.setdt()
            if par.dtsave_log < BIGDOUBLE and self._log_due():
                self.log_line()
            if (st.it >= par.it_max or st.global_time >= par.time_max
                    or st.final_dt_exit or (nsteps is not None and steps >= nsteps)):
                break
            self.advect()
            st.dts.append(st.dt)
            if st.crash or self.be.crash():
                st.crash = True
                raise RuntimeError("Error: small value encountered, run crash.")
            st.it += 1
            st.global_time = st.global_time + st.dt
            steps += 1
        return st

    def step(self) -> None:
        """One time step (setdt + advance), no logging."""
        self.setdt()
        self.advect()
        self.state.dts.append(self.state.dt)
        self.state.it += 1
        self.state.global_time = self.state.global_time + self.state.dt


def read_log(path: str) -> np.ndarray:
    rows = []
    with open(path) as fh:
        for line in fh:
            s = line.strip()
            if not s or s[0] in "#!":
                continue
            rows.append([float(v) for v in s.split()])
    return np.array(rows, dtype=np.float64)


def logs_match(a: np.ndarray, b: np.ndarray, rtol: float = 1.0e-5,
               atol: float = 1.0e-8) -> bool:
    """compare_logs semantics, tools/fortran/compare_logs.f:47-50."""
    if a.shape != b.shape:
        return False
    return bool(np.all(np.abs(a - b) <= atol + 0.5 * rtol * (np.abs(a) + np.abs(b))))
