"""Host-side time integration that drives a finite-volume backend.

This mirrors the reference's host loop: ``timeintegration``
(src/amrvac.t:189-400), ``setdt`` (src/mod_dt.t:10-288), ``advance`` /
``advect`` (src/mod_advance.t:17-567) and the regression log
(src/io/mod_input_output.t:2653-2729).  The per-stage work, ``advect1``
(src/mod_advance.t:611-666), is delegated to a backend object; the product
backend is :class:`amrvac_amd.fv.DeviceFV`, which calls the HIP kernels
through the C-ABI declared in include/amrvac_fv.h.

Backend protocol (slots 0=ps, 1=ps1, 2=ps2, 3=ps3):
  getdt_courant(slot, courantpar) -> (dtmin, cmax)
  advect1(method, dtfactor, dt, cmax_global, psa, psb, comb)
      comb = None                 psb holds the initial wnew
           | ("copy", src)        psb = ps[src] first (ps1=ps in advect)
           | ("axpby", a, x, b)   psb = a*ps[x] + b*ps[psa] first
  combine(dst, ops, consts)       ps[dst] = a postfix expression over the
                                  slots (see Expr), one rounding per node
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
FS = {"hll": 1, "hllc": 2, "hlld": 3, "tvdlf": 5, "hancock": 8}
# limiter codes, src/limiter/mod_limiter.t:15-40
LIMITERS = {"minmod": 1, "woodward": 2, "mcbeta": 3, "superbee": 4,
            "vanleer": 5, "albada": 6, "koren": 7, "cada": 8, "cada3": 9,
            "weno5": 16}
BOUNDSPEED = {"Einfeldt": 1, "cmaxmean": 2, "cmaxleftright": 3}
DIVB = {"none": 0, "glm": 1, "linde": 4, "ct": 8}
TYPE_CT = {"average": 1, "uct_contact": 2, "uct_hll": 3}


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
    type_ct: str = "uct_contact"   # mhd_list, constrained transport with typedivbfix='ct'
    time_stepper: str = "twostep"  # 'onestep' | 'twostep' | 'threestep' | 'fourstep' | 'fivestep'
    time_integrator: str = "default"
    ssprk_order: int = 3
    rk3_switch: int = 3
    rk2_alfa: float = 0.5
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
    fix_small_values: bool = False
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
    def stagger(self) -> bool:
        """stagger_grid: constrained transport keeps B on cell faces."""
        return self.physics == "mhd" and self.typedivbfix == "ct"

    @property
    def nws(self) -> int:
        """Face-field components (nws = ndim, src/mhd/mod_mhd_phys.t:558)."""
        return self.ndim if self.stagger else 0

    @property
    def nwt(self) -> int:
        """Variables per block of a backend slot: nw cell variables, then
        the nws face-field components."""
        return self.nw + self.nws

    @property
    def need_global_cmax(self) -> bool:
        return self.physics == "mhd" and self.typedivbfix == "glm"

    @property
    def typepred1(self) -> str:
        # src/io/mod_input_output.t:788-810
        return "hancock" if self.flux_scheme == "tvdlf" else self.flux_scheme

    @property
    def nstages(self) -> int:
        return {"onestep": 1, "twostep": 2, "threestep": 3, "fourstep": 4, "fivestep": 5}[self.time_stepper]

    @property
    def integrator(self) -> str:
        """time_integrator with its per-stepper default
        (src/io/mod_input_output.t:830-1077)."""
        if self.time_integrator != "default":
            ti = self.time_integrator
        else:
            ti = {"onestep": "Forward_Euler", "twostep": "Predictor_Corrector", "threestep": "ssprk3",
                  "fourstep": "ssprk4", "fivestep": "ssprk5"}[self.time_stepper]
        allowed = {"onestep": ("Forward_Euler",), "twostep": ("Predictor_Corrector", "RK2_alfa", "ssprk2"),
                   "threestep": ("ssprk3", "RK3_BT"), "fourstep": ("ssprk4", "rk4"), "fivestep": ("ssprk5",)}
        if ti not in allowed[self.time_stepper]:
            raise ValueError(f"unsupported time_integrator {ti!r} for time_stepper {self.time_stepper!r}")
        return ti


def ssprk3_coefficients(order: int = 3) -> dict[str, float]:
    """SSPRK(3,3) (order 3) or SSP(3,2) (order 2),
    src/io/mod_input_output.t:941-964."""
    if order == 3:
        c = dict(beta11=1.0, beta22=1.0 / 4.0, beta33=2.0 / 3.0, alfa21=3.0 / 4.0, alfa31=1.0 / 3.0)
    elif order == 2:
        c = dict(beta11=1.0 / 2.0, beta22=1.0 / 2.0, beta33=1.0 / 3.0, alfa21=0.0, alfa31=1.0 / 3.0)
    else:
        raise ValueError(f"unknown ssprk_order {order}")
    c["alfa22"] = 1.0 - c["alfa21"]
    c["alfa33"] = 1.0 - c["alfa31"]
    return c


def rk3_bt_coefficients(switch: int = 3) -> dict[str, float]:
    """Three-stage Butcher tableaux of RK3_BT: Ralston (1), RK-Wray (2),
    Heun (3), Nystrom (4); src/io/mod_input_output.t:904-938."""
    tab = {1: (1.0 / 2.0, 0.0, 3.0 / 4.0, 2.0 / 9.0, 1.0 / 3.0),
           2: (8.0 / 15.0, 1.0 / 4.0, 5.0 / 12.0, 1.0 / 4.0, 0.0),
           3: (1.0 / 3.0, 0.0, 2.0 / 3.0, 1.0 / 4.0, 0.0),
           4: (2.0 / 3.0, 0.0, 2.0 / 3.0, 1.0 / 4.0, 3.0 / 8.0)}
    if switch not in tab:
        raise ValueError(f"unknown rk3_switch {switch}")
    a21, a31, a32, b1, b2 = tab[switch]
    return dict(a21=a21, a31=a31, a32=a32, b1=b1, b2=b2, b3=1.0 - b1 - b2)


def ssprk4_coefficients(order: int = 3) -> dict[str, float]:
    """SSPRK(4,3) (order 3) or SSP(4,2) (order 2),
    src/io/mod_input_output.t:1036-1064."""
    if order == 3:
        c = dict(beta11=1.0 / 2.0, beta22=1.0 / 2.0, beta33=1.0 / 6.0, beta44=1.0 / 2.0,
                 alfa21=0.0, alfa31=2.0 / 3.0, alfa41=0.0)
    elif order == 2:
        c = dict(beta11=1.0 / 3.0, beta22=1.0 / 3.0, beta33=1.0 / 3.0, beta44=1.0 / 4.0,
                 alfa21=0.0, alfa31=0.0, alfa41=1.0 / 4.0)
    else:
        raise ValueError(f"unknown ssprk_order {order}")
    c["alfa22"] = 1.0 - c["alfa21"]
    c["alfa33"] = 1.0 - c["alfa31"]
    c["alfa44"] = 1.0 - c["alfa41"]
    return c


def ssprk5_coefficients(order: int = 3) -> dict[str, float]:
    """SSPRK(5,4) coefficient sets: Gottlieb 2005 (order 3) or
    Spiteri-Ruuth (order 2), src/io/mod_input_output.t:1078-1136."""
    if order == 3:
        c = dict(beta11=0.391752226571890, beta22=0.368410593050371, beta33=0.251891774271694,
                 beta44=0.544974750228521, beta54=0.063692468666290, beta55=0.226007483236906,
                 alfa21=0.444370493651235, alfa31=0.620101851488403, alfa41=0.178079954393132,
                 alfa53=0.517231671970585, alfa54=0.096059710526147)
        c["alfa22"] = 1.0 - c["alfa21"]
        c["alfa33"] = 1.0 - c["alfa31"]
        c["alfa44"] = 1.0 - c["alfa41"]
        c["alfa55"] = 1.0 - c["alfa53"] - c["alfa54"]
    elif order == 2:
        c = dict(beta11=0.39175222700392, beta22=0.36841059262959, beta33=0.25189177424738,
                 beta44=0.54497475021237, alfa21=0.44437049406734, alfa31=0.62010185138540,
                 alfa41=0.17807995410773, alfa53=0.51723167208978)
        c["alfa22"] = 1.0 - c["alfa21"]
        c["alfa33"] = 1.0 - c["alfa31"]
        c["alfa44"] = 1.0 - c["alfa41"]
        rka51, rka54, rkb54 = 0.00683325884039, 0.12759831133288, 0.08460416338212
        c["beta54"] = rkb54 - c["beta44"] * rka51 / c["alfa41"]
        c["alfa54"] = rka54 - c["alfa44"] * rka51 / c["alfa41"]
        c["alfa55"] = 1.0 - c["alfa53"] - c["alfa54"]
        c2 = c["beta11"]
        c3 = c["alfa22"] * c2 + c["beta22"]
        c4 = c["alfa33"] * c3 + c["beta33"]
        c5 = c["alfa44"] * c4 + c["beta44"]
        c["beta55"] = 1.0 - c["beta54"] - c["alfa53"] * c3 - c["alfa54"] * c4 - c["alfa55"] * c5
    else:
        raise ValueError(f"unknown ssprk_order {order}")
    return c


# postfix operation codes of a state combination (shared with the C-ABI's
# amrvac_fv_combine: 0..3 push slot ps..ps3, then the constant and the
# arithmetic operators)
OP_SLOT0, OP_CONST, OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_NEG = 0, 8, 9, 10, 11, 12, 13
COMBINE_MAX_OPS, COMBINE_MAX_CONSTS, COMBINE_MAX_DEPTH = 32, 16, 8


class Expr:
    """A state combination as a postfix program.  Python's operator
    precedence and left-to-right association equal Fortran's for these
    expressions, so an update written here as in the reference's loops
    (e.g. ``K(a21)*S(0) + K(a22)*S(1)``) evaluates the same operations in the
    same order, each rounded once."""

    def __init__(self, ops, consts):
        self.ops, self.consts = list(ops), list(consts)

    @staticmethod
    def _lift(v):
        return v if isinstance(v, Expr) else K(float(v))

    def _bin(self, other, op, swap=False):
        o = Expr._lift(other)
        a, b = (o, self) if swap else (self, o)
        return Expr(a.ops + b.ops + [op], a.consts + b.consts)

    def __add__(self, o): return self._bin(o, OP_ADD)
    def __radd__(self, o): return self._bin(o, OP_ADD, True)
    def __sub__(self, o): return self._bin(o, OP_SUB)
    def __rsub__(self, o): return self._bin(o, OP_SUB, True)
    def __mul__(self, o): return self._bin(o, OP_MUL)
    def __rmul__(self, o): return self._bin(o, OP_MUL, True)
    def __truediv__(self, o): return self._bin(o, OP_DIV)
    def __rtruediv__(self, o): return self._bin(o, OP_DIV, True)
    def __neg__(self): return Expr(self.ops + [OP_NEG], self.consts)

    def check(self) -> None:
        depth, nc = 0, 0
        for op in self.ops:
            if OP_SLOT0 <= op <= OP_SLOT0 + 3 or op == OP_CONST:
                depth += 1
                nc += op == OP_CONST
            elif op == OP_NEG:
                if depth < 1:
                    raise ValueError("stack underflow")
            elif op in (OP_ADD, OP_SUB, OP_MUL, OP_DIV):
                if depth < 2:
                    raise ValueError("stack underflow")
                depth -= 1
            else:
                raise ValueError(f"bad op {op}")
            if depth > COMBINE_MAX_DEPTH:
                raise ValueError("combination too deep")
        if depth != 1 or nc != len(self.consts):
            raise ValueError("malformed combination")
        if len(self.ops) > COMBINE_MAX_OPS or nc > COMBINE_MAX_CONSTS:
            raise ValueError("combination too long")

    def evaluate(self, slots):
        """Reference evaluation on arrays (numpy): one operation per node."""
        st, kc = [], 0
        for op in self.ops:
            if OP_SLOT0 <= op <= OP_SLOT0 + 3:
                st.append(slots[op - OP_SLOT0])
            elif op == OP_CONST:
                st.append(self.consts[kc]); kc += 1
            elif op == OP_NEG:
                st.append(-st.pop())
            else:
                b, a = st.pop(), st.pop()
                if op == OP_ADD:
                    st.append(a + b)
                elif op == OP_SUB:
                    st.append(a - b)
                elif op == OP_MUL:
                    st.append(a * b)
                else:
                    st.append(a / b)
        return st[0]


def S(slot: int) -> Expr:
    """State slot ps (0), ps1, ps2, ps3."""
    if not 0 <= slot <= 3:
        raise ValueError("slot must be 0..3")
    return Expr([OP_SLOT0 + slot], [])


def K(value: float) -> Expr:
    return Expr([OP_CONST], [float(value)])


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
    chs: list = field(default_factory=list)  # cmax_global of each step (GLM c_h)


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
        ti = par.integrator
        be = self.be

        def adv(dtfactor, psa, psb, comb=None):
            be.advect1(fm, dtfactor, dt, ch, psa, psb, comb)

        # ps1 = ps at the start of advect (:69-74) is folded into the first
        # stage as a "copy" combination; an update a*ps + b*ps1 written just
        # before a stage whose input is ps1 is folded into that stage as an
        # "axpby" combination; every other update is a combine pass
        if ti == "ssprk3":  # :251-267
            c = ssprk3_coefficients(par.ssprk_order)
            adv(c["beta11"], 0, 1, ("copy", 0))
            adv(c["beta22"], 1, 2, ("axpby", c["alfa21"], 0, c["alfa22"]))
            adv(c["beta33"], 2, 0, ("axpby", c["alfa31"], 0, c["alfa33"]))
        elif ti == "Predictor_Corrector":
            # predictor without refluxing, :104-111
            be.advect1(FS[par.typepred1], 0.5, dt, ch, 0, 1, ("copy", 0))
            adv(1.0, 1, 0)
        elif ti == "Forward_Euler":  # :76-82
            be.copy(1, 0)
            adv(1.0, 1, 0)
        elif ti == "ssprk2":  # :124-133
            adv(1.0, 0, 1, ("copy", 0))
            adv(0.5, 1, 0, ("axpby", 0.5, 0, 0.5))
        elif ti == "RK2_alfa":  # :113-122
            a21 = par.rk2_alfa
            if a21 < SMALLDOUBLE or a21 > 1.0:
                raise ValueError("set rk2_alfa within [0,1]")
            b2 = 1.0 / (2.0 * a21)
            b1 = 1.0 - b2
            adv(a21, 0, 1, ("copy", 0))
            be.combine(0, S(0) + K(b1) * (S(1) - S(0)) / K(a21))
            adv(b2, 1, 0)
        elif ti == "RK3_BT":  # :269-291
            c = rk3_bt_coefficients(par.rk3_switch)
            adv(c["a21"], 0, 1, ("copy", 0))
            be.combine(3, (S(1) - S(0)) / K(c["a21"]))
            be.combine(2, S(0) + K(c["a31"]) * S(3))
            adv(c["a32"], 1, 2)
            be.combine(0, S(0) + K(c["b1"]) * S(3)
                       + K(c["b2"]) * (S(2) - (S(0) + K(c["a31"]) * S(3))) / K(c["a32"]))
            adv(c["b3"], 2, 0)
        elif ti == "ssprk4":  # :456-483
            c = ssprk4_coefficients(par.ssprk_order)
            adv(c["beta11"], 0, 1, ("copy", 0))
            adv(c["beta22"], 1, 2, ("axpby", c["alfa21"], 0, c["alfa22"]))
            adv(c["beta33"], 2, 1, ("axpby", c["alfa31"], 0, c["alfa33"]))
            adv(c["beta44"], 1, 0, ("axpby", c["alfa41"], 0, c["alfa44"]))
        elif ti == "rk4":  # :485-507 (ps2 = ps3 = ps first: "copy" into both)
            adv(0.5, 0, 1, ("copy", 0))
            adv(0.5, 1, 2, ("copy", 0))
            adv(1.0, 2, 3, ("copy", 0))
            be.combine(0, K(1.0 / 3.0) * (-S(0) + S(1) + K(2.0) * S(2) + S(3)))
            adv(1.0 / 6.0, 3, 0)
        elif ti == "ssprk5":  # :515-557
            c = ssprk5_coefficients(par.ssprk_order)
            adv(c["beta11"], 0, 1, ("copy", 0))
            adv(c["beta22"], 1, 2, ("axpby", c["alfa21"], 0, c["alfa22"]))
            adv(c["beta33"], 2, 1, ("axpby", c["alfa31"], 0, c["alfa33"]))
            be.combine(3, K(c["alfa53"]) * S(2) + K(c["alfa54"]) * S(1))
            adv(c["beta44"], 1, 2, ("axpby", c["alfa41"], 0, c["alfa44"]))
            be.combine(0, S(3) + K(c["alfa55"]) * S(2)
                       + K(c["beta54"] / c["beta44"]) * (S(2) - (K(c["alfa41"]) * S(0) + K(c["alfa44"]) * S(1))))
            adv(c["beta55"], 2, 0)
        else:
            raise ValueError(f"unsupported time_integrator {ti!r}")

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
            st.chs.append(st.cmax_global)
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
        self.state.chs.append(self.state.cmax_global)
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
