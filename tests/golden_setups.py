"""Setups of the reference's maintained AMR regression tests (directories in
the reference's tests/Makefile HD_DIRS / MHD_DIRS), for runs through the
host's AMR time loop on any backend.

Each setup is the test's own .par file followed by the scheme file of
tests/schemes/ (read in that order, so the scheme file wins: cmaxmean bound
speeds unless it names another, CFL 0.9), with the initial condition of the
test's mod_usr.t:

* ``kh2d``  tests/hd/Kelvin_Helmholtz_2D (kh_2d.par): 2-D HD, 32^2 in 16^2
  blocks, periodic, 3 levels, slowsteps 10, Löhner on rho.
* ``kh3d``  tests/mhd/Kelvin_Helmholtz_3D (kh_3d.par): 3-D MHD, 20x40x40
  in 10^3 blocks, periodic x, cont y/z, 3 levels, slowsteps 10, Löhner on
  rho (0.4) and momenta (0.2 each) at threshold 0.2.
* ``bw3d``  tests/mhd/blast_wave_Cartesian_3D (bw_3d.par): 3-D MHD, 16^3
  in 8^3 blocks, cont, 3 levels, Löhner on rho and e at threshold 0.2,
  small values 1e-14 with fix_small_values.
* ``ot2d``  tests/mhd/Orszag_Tang_2D (ot_2d.par): 2-D MHD with two vector
  components, 64^2 in 16^2 blocks, periodic, 3 levels, Löhner on rho and e
  at threshold 0.4.

Every MHD setup cleans div B with Linde's source (the default
typedivbfix) unless the scheme file selects GLM.
"""
from __future__ import annotations

import os

from amrvac_amd import problems
from amrvac_amd.amr import AMRForest, AMRPar
from amrvac_amd.host import Par, read_log

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

# tests/schemes/<name>.par
SCHEMES = {
    "2step_tvdlf_mm": dict(time_stepper="twostep", flux_scheme="tvdlf", limiter="minmod"),
    "3step_hll_cada": dict(time_stepper="threestep", flux_scheme="hll", limiter="cada3"),
    "3step_hll_ko": dict(time_stepper="threestep", flux_scheme="hll", limiter="koren"),
    "3step_hlld_cada": dict(time_stepper="threestep", flux_scheme="hlld", limiter="cada3", typedivbfix="glm",
                            typeboundspeed="Einfeldt"),
    "4step_hll_mc": dict(time_stepper="fourstep", time_integrator="rk4", flux_scheme="hll", limiter="woodward"),
    "4step_hllc_ko": dict(time_stepper="fourstep", time_integrator="ssprk4", flux_scheme="hllc", limiter="koren"),
    "rk4_tvdlf_cada": dict(time_stepper="fourstep", time_integrator="rk4", flux_scheme="tvdlf", limiter="cada3"),
}

# fixture name prefix per setup (tests/golden/<prefix><scheme>.log)
PREFIX = {"kh2d": "kh_2d_", "kh3d": "kh_3D_", "bw3d": "mhd_bw_3d_", "ot2d": "ot_2d_"}


def setup(case: str, scheme: str, time_max: float | None = None, **over):
    """(par, forest, amr_par, init_fn) of a maintained regression test;
    ``over`` replaces Par fields (e.g. the precision of a device run)."""
    kw = dict(typeboundspeed="cmaxmean")
    kw.update(SCHEMES[scheme])
    kw.update(courantpar=0.9, dtsave_log=1.0e-3)
    divb = kw.pop("typedivbfix", "linde")
    if case == "kh2d":
        kw.update(physics="hd", ndim=2, ndir=2, gamma=5.0 / 3.0, slowsteps=10, time_max=time_max or 5.0e-2)
        forest = AMRForest(2, (32, 32), (16, 16), (0.0, 0.0), (1.0, 1.0), ("periodic",) * 4, max_level=3)
        ap = AMRPar(refine_max_level=3)
        init = problems.kelvin_helmholtz_2d_blocks
    elif case == "kh3d":
        kw.update(physics="mhd", ndim=3, ndir=3, typedivbfix=divb, slowsteps=10, time_max=time_max or 1.0e-2)
        forest = AMRForest(3, (20, 40, 40), (10, 10, 10), (0.0, -1.0, -1.0), (1.0, 1.0, 1.0),
                           ("periodic", "periodic", "cont", "cont", "cont", "cont"), max_level=3)
        ap = AMRPar(refine_max_level=3, refine_threshold=0.2, w_refine_weight={0: 0.4, 1: 0.2, 2: 0.2, 3: 0.2})
        init = problems.mhd_kelvin_helmholtz_3d_blocks
    elif case == "bw3d":
        kw.update(physics="mhd", ndim=3, ndir=3, typedivbfix=divb, time_max=time_max or 2.0e-2,
                  small_pressure=1.0e-14, small_density=1.0e-14, fix_small_values=True)
        forest = AMRForest(3, (16, 16, 16), (8, 8, 8), (0.0,) * 3, (2.0,) * 3, ("cont",) * 6, max_level=3)
        ap = AMRPar(refine_max_level=3, refine_threshold=0.2, w_refine_weight={0: 0.5, 4: 0.5})
        init = problems.mhd_blast_wave_3d_blocks
    elif case == "ot2d":
        kw.update(physics="mhd", ndim=2, ndir=2, typedivbfix=divb, time_max=time_max or 0.1)
        forest = AMRForest(2, (64, 64), (16, 16), (0.0, 0.0), (1.0, 1.0), ("periodic",) * 4, max_level=3)
        ap = AMRPar(refine_max_level=3, refine_threshold=0.4, w_refine_weight={0: 0.5, 3: 0.5})
        init = problems.orszag_tang_2d_blocks
    else:
        raise ValueError(case)
    kw.update(over)
    par = Par(**kw)
    if par.physics == "mhd" and par.typedivbfix == "glm":
        base = init

        def init(f, b):
            return base(f, b, glm=True)
    return par, forest, ap, init


def golden(case: str, scheme: str):
    return read_log(os.path.join(GOLDEN, PREFIX[case] + scheme + ".log"))
