"""Every limiter, Riemann solver, bound-speed rule and time integrator of the
device path against the CPU oracle, on small meshes: the contraction-free
FP64 kernels must reproduce the oracle bit for bit (state and dt sequence),
covering the kernel instantiations the benchmark setups do not reach
(the runtime-switched limiters, cmaxleftright bounds, onestep, the 3-D MHD
kernels without divergence cleaning)."""
from __future__ import annotations

import numpy as np
import pytest

from amrvac_amd import problems
from amrvac_amd.host import Par, TimeIntegrator
from amrvac_amd.mesh import UniformMesh
from oracle.oracle import OracleFV

pytestmark = pytest.mark.gpu

LIMS = ["minmod", "woodward", "mcbeta", "superbee", "vanleer", "albada", "koren", "cada", "cada3", "weno5"]


def _cases():
    out = []
    for lim in LIMS:  # every limiter, 2-D HD, HLL, Einfeldt
        out.append(("hd2", lim, "hll", "Einfeldt", "threestep"))
    for fs in ("tvdlf", "hll", "hllc"):
        for bs in ("Einfeldt", "cmaxmean", "cmaxleftright"):
            out.append(("hd2", "vanleer", fs, bs, "threestep"))
    for ts in ("onestep", "twostep"):
        out.append(("hd2", "koren", "hll", "cmaxmean", ts))
    out.append(("hd2", "minmod", "tvdlf", "Einfeldt", "twostep"))  # Hancock predictor
    for lim in ("woodward", "superbee", "albada", "koren", "weno5"):  # 3-D MHD-GLM, runtime limiters
        out.append(("mhd3glm", lim, "hll", "Einfeldt", "threestep"))
    for lim, fs in (("vanleer", "hll"), ("minmod", "tvdlf"), ("cada3", "hll")):  # 3-D ideal MHD, no cleaning
        out.append(("mhd3", lim, fs, "Einfeldt", "threestep"))
    out.append(("hd3", "mcbeta", "hllc", "cmaxleftright", "threestep"))
    # every explicit integrator of advect (stepper:integrator:ssprk_order:rk3_switch)
    for ts in ("twostep:RK2_alfa:3:3", "twostep:ssprk2:3:3", "threestep:ssprk3:2:3", "threestep:RK3_BT:3:1",
               "threestep:RK3_BT:3:3", "fourstep:ssprk4:3:3", "fourstep:ssprk4:2:3", "fourstep:rk4:3:3",
               "fivestep:ssprk5:3:3", "fivestep:ssprk5:2:3"):
        out.append(("hd2", "vanleer", "hll", "Einfeldt", ts))
    out.append(("mhd3glm", "cada3", "hll", "Einfeldt", "fourstep:rk4:3:3"))
    # HLLD (3-D MHD, with and without GLM; every bound-speed rule)
    for lim, bs in (("cada3", "Einfeldt"), ("vanleer", "cmaxmean"), ("koren", "cmaxleftright")):
        out.append(("mhd3glm", lim, "hlld", bs, "threestep"))
    out.append(("mhd3", "minmod", "hlld", "Einfeldt", "threestep"))
    out.append(("mhd3glm", "weno5", "hlld", "Einfeldt", "threestep"))
    out.append(("mhd3", "vanleer", "hlld", "Einfeldt", "twostep"))
    # MHD HLLC (Li 2005 contact state), with and without GLM
    out.append(("mhd3glm", "koren", "hllc", "cmaxmean", "fourstep:ssprk4:3:3"))
    out.append(("mhd3glm", "cada3", "hllc", "Einfeldt", "threestep"))
    out.append(("mhd3", "vanleer", "hllc", "Einfeldt", "threestep"))
    out.append(("mhd3", "minmod", "hllc", "cmaxleftright", "twostep"))
    return out


def _stepper(ts):
    if ":" not in ts:
        return dict(time_stepper=ts)
    stepper, integ, order, sw = ts.split(":")
    return dict(time_stepper=stepper, time_integrator=integ, ssprk_order=int(order), rk3_switch=int(sw))


def _setup(phys, lim, fs, bs, ts):
    if phys == "hd2":
        par = Par(physics="hd", ndim=2, ndir=2, gamma=1.4, flux_scheme=fs, limiter=lim,
                  typeboundspeed=bs, courantpar=0.6, **_stepper(ts))
        mesh = UniformMesh(2, (32, 32), (16, 16), (0.0, 0.0), (1.0, 1.0),
                           ("periodic", "periodic", "symm", "cont"), nghost=par.nghost)
        w0 = problems.kelvin_helmholtz_2d(mesh, gamma=par.gamma)
    elif phys == "hd3":
        par = Par(physics="hd", ndim=3, ndir=3, gamma=5.0 / 3.0, flux_scheme=fs, limiter=lim,
                  typeboundspeed=bs, courantpar=0.6, **_stepper(ts))
        mesh = UniformMesh(3, (16, 16, 32), (16, 16, 16), (0.0,) * 3, (1.0, 1.0, 2.0), ("periodic",) * 6,
                           nghost=par.nghost)
        w0 = problems.blast_wave(mesh, gamma=par.gamma, p_in=1.0, radius=0.2)
    else:
        divb = "glm" if phys == "mhd3glm" else "none"
        par = Par(physics="mhd", ndim=3, ndir=3, gamma=5.0 / 3.0, typedivbfix=divb, flux_scheme=fs,
                  limiter=lim, typeboundspeed=bs, courantpar=0.6, **_stepper(ts))
        mesh = UniformMesh(3, (16, 16, 32), (16, 16, 16), (0.0,) * 3, (1.0, 1.0, 2.0), ("periodic",) * 6,
                           nghost=par.nghost)
        w0 = problems.orszag_tang_3d(mesh, gamma=par.gamma)
        if divb == "none":
            w0 = np.ascontiguousarray(w0[:, :8])
    return par, mesh, w0


def _run(par, mesh, w0, be, nsteps):
    be.set_state(0, w0)
    be.getbc(0)
    ti = TimeIntegrator(par, mesh, be)
    ti.run(nsteps=nsteps)
    return mesh.interior(be.get_state(0)), ti.state.dts


@pytest.mark.parametrize("phys,lim,fs,bs,ts", _cases(),
                         ids=lambda v: str(v))
def test_exact_kernels_match_oracle_bitwise(phys, lim, fs, bs, ts):
    from amrvac_amd.fv import DeviceFV
    print(f"[matrix] {phys} {lim} {fs} {bs} {ts}", flush=True)
    par, mesh, w0 = _setup(phys, lim, fs, bs, ts)
    nsteps = 3
    w_orc, dt_orc = _run(par, mesh, w0, OracleFV(par, mesh, threads=2), nsteps)
    w_dev, dt_dev = _run(par, mesh, w0, DeviceFV(par, mesh, precision="fp64_exact"), nsteps)
    assert dt_dev == dt_orc
    assert np.array_equal(w_dev, w_orc)
    w_p, _ = _run(par, mesh, w0, DeviceFV(par, mesh, precision="fp64"), nsteps)
    rel = np.abs(w_p - w_orc).sum() / np.abs(w_orc).sum()
    assert rel < 1e-10
