"""The small-value path on the device: the crash flag of the reference
(src/amrvac.t:311-318) set by the stage and dt kernels, read back by the
host loop, stops a device run at the same step as the oracle's, with the
same state.  The run is the oracle check's rarefaction (two smoothly
separating streams whose central pressure falls below small_pressure)."""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from amrvac_amd import problems
from amrvac_amd.amr import AMRForest, AMRIntegrator, AMRPar
from amrvac_amd.host import Par, TimeIntegrator
from amrvac_amd.mesh import UniformMesh
from oracle.oracle import OracleFV

pytestmark = pytest.mark.gpu

GAMMA = 5.0 / 3.0


def _mesh():
    return UniformMesh(2, (64, 16), (16, 16), (0.0, 0.0), (1.0, 0.25), ("cont", "cont", "periodic", "periodic"))


def _state(par, mesh):
    x = mesh.cell_centers()
    vx = 0.6 * np.tanh((x[:, 0] - 0.5) / 0.05)
    one = np.ones_like(vx)
    if par.physics == "hd":
        return problems.hd_to_conserved(np.stack([one, vx, 0 * one, 0.1 * one], axis=1), GAMMA, 2)
    return problems.mhd_to_conserved(np.stack([one, vx, 0 * one, 0 * one, 0.1 * one, 0.1 * one, 0 * one, 0 * one],
                                              axis=1), GAMMA, 3)


def _run(be, par, mesh, nmax=60):
    be.set_state(0, _state(par, mesh))
    be.getbc(0)
    ti = TimeIntegrator(par, mesh, be)
    try:
        ti.run(nsteps=nmax)
    except RuntimeError as e:
        assert "small value" in str(e)
        return True, ti.state.it, mesh.interior(be.get_state(0))
    return False, ti.state.it, mesh.interior(be.get_state(0))


CASES = {
    "hd_fix_small_values": Par(physics="hd", ndim=2, ndir=2, gamma=GAMMA, time_stepper="threestep",
                               flux_scheme="hll", limiter="vanleer", courantpar=0.4, small_pressure=0.02,
                               fix_small_values=True),
    "hd_cmaxmean_check_only": Par(physics="hd", ndim=2, ndir=2, gamma=GAMMA, time_stepper="threestep",
                                  flux_scheme="hllc", limiter="minmod", typeboundspeed="cmaxmean",
                                  courantpar=0.4, small_pressure=0.02),
    "mhd_fix_small_values": Par(physics="mhd", ndim=2, ndir=3, gamma=GAMMA, typedivbfix="linde",
                                time_stepper="twostep", flux_scheme="tvdlf", limiter="minmod", courantpar=0.4,
                                small_pressure=0.02, fix_small_values=True),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_device_stops_at_the_oracles_crash_step(case):
    from amrvac_amd.fv import DeviceFV
    print(f"[small values] {case}", flush=True)
    par = CASES[case]
    mesh = _mesh()
    crashed_o, it_o, w_o = _run(OracleFV(par, mesh), par, mesh)
    crashed_d, it_d, w_d = _run(DeviceFV(par, mesh, precision="fp64_exact"), par, mesh)
    assert crashed_o and 0 < it_o < 60
    assert crashed_d and it_d == it_o
    assert np.array_equal(w_d, w_o)
    # the production build stops at the same step
    crashed_p, it_p, _ = _run(DeviceFV(par, mesh, precision="fp64"), par, mesh)
    assert crashed_p and it_p == it_o


def test_no_floor_no_crash_and_the_flag_clears():
    from amrvac_amd.fv import DeviceFV
    par = dataclasses.replace(CASES["hd_fix_small_values"], small_pressure=0.0)
    mesh = _mesh()
    be = DeviceFV(par, mesh)
    crashed, it, _ = _run(be, par, mesh, nmax=40)
    assert not crashed and it == 40
    par2 = CASES["hd_fix_small_values"]
    be2 = DeviceFV(par2, mesh)
    crashed, _, _ = _run(be2, par2, mesh)
    assert crashed and be2.crash()
    assert be2.crash(clear=True) and not be2.crash()


def _amr_init(forest, blocks):
    x = forest.cell_centers(blocks)
    vx = 0.6 * np.tanh((x[:, 0] - 0.5) / 0.05)
    one = np.ones_like(vx)
    zero = 0 * one
    prim = [one, vx] + [zero] * (forest.ndim - 1) + [0.1 * one]
    return problems.hd_to_conserved(np.stack(prim, axis=1), GAMMA, forest.ndim)


def _amr_run(ndim, make_backend, nmax=300):
    """The rarefaction on a two-level forest refined on m1 (Löhner every
    step): the crash flag reaches the host with the estimate's flags."""
    par = dataclasses.replace(CASES["hd_fix_small_values"], ndim=ndim, ndir=ndim)
    forest = AMRForest(ndim, (64,) + (16,) * (ndim - 1), (16,) * ndim, (0.0,) * ndim, (1.0,) + (0.25,) * (ndim - 1),
                       ("cont", "cont") + ("periodic", "periodic") * (ndim - 1), max_level=2)
    ti = AMRIntegrator(par, forest, AMRPar(refine_max_level=2, w_refine_weight={1: 1.0}), make_backend, _amr_init)
    try:
        ti.run(nsteps=nmax)
    except RuntimeError as e:
        assert "small value" in str(e)
        return True, ti.state.it, [n.ig for n in forest.leaves], forest.interior(ti.be.get_state(0))
    return False, ti.state.it, [n.ig for n in forest.leaves], forest.interior(ti.be.get_state(0))


@pytest.mark.parametrize("ndim", [2, 3])
def test_amr_device_stops_at_the_oracles_crash_step(ndim):
    from amrvac_amd.fv import DeviceFV
    print(f"[small values] AMR, two levels, {ndim}-D", flush=True)
    crashed_o, it_o, leaves_o, w_o = _amr_run(ndim, lambda p, f: OracleFV(p, f, threads=4))
    assert crashed_o and 0 < it_o < 300
    crashed_d, it_d, leaves_d, w_d = _amr_run(ndim, lambda p, f: DeviceFV(p, f, precision="fp64_exact"))
    assert crashed_d and it_d == it_o and leaves_d == leaves_o
    assert np.array_equal(w_d, w_o)
    crashed_p, it_p, _, _ = _amr_run(ndim, lambda p, f: DeviceFV(p, f, precision="fp64"))
    assert crashed_p and it_p == it_o
