"""Edge cases of the mesh on the device against the CPU oracle: a level of a
single periodic block, which is its own neighbour in every direction (halo
cells and, with constrained transport, the shared faces come from the block
itself).  The contraction-free FP64 kernels must match the oracle bit for
bit."""
from __future__ import annotations

import numpy as np
import pytest

from amrvac_amd import problems
from amrvac_amd.host import Par, TimeIntegrator
from amrvac_amd.mesh import UniformMesh
from oracle.oracle import OracleFV

pytestmark = pytest.mark.gpu


def _compare(par, mesh, w0, nsteps=6):
    from amrvac_amd.fv import DeviceFV
    out = {}
    for name, mk in (("oracle", lambda: OracleFV(par, mesh, threads=2)),
                     ("device", lambda: DeviceFV(par, mesh, precision="fp64_exact"))):
        be = mk()
        be.set_state(0, w0)
        be.getbc(0)
        ti = TimeIntegrator(par, mesh, be)
        ti.run(nsteps=nsteps)
        out[name] = (mesh.interior(be.get_state(0)), list(ti.state.dts))
    assert out["device"][1] == out["oracle"][1]
    assert np.array_equal(out["device"][0], out["oracle"][0])


def test_single_block_2d_hd():
    print("[edge] 2-D HD, one periodic block", flush=True)
    par = Par(physics="hd", ndim=2, ndir=2, gamma=1.4, time_stepper="threestep", flux_scheme="hllc",
              limiter="koren", typeboundspeed="cmaxmean", courantpar=0.6)
    mesh = UniformMesh(2, (16, 16), (16, 16), (0.0, 0.0), (1.0, 1.0), ("periodic",) * 4)
    _compare(par, mesh, problems.kelvin_helmholtz_2d(mesh, gamma=par.gamma))


def test_single_block_3d_mhd_glm():
    print("[edge] 3-D MHD-GLM, one periodic block", flush=True)
    par = Par(physics="mhd", ndim=3, ndir=3, gamma=5.0 / 3.0, typedivbfix="glm", time_stepper="threestep",
              flux_scheme="hll", limiter="cada3", courantpar=0.6)
    mesh = UniformMesh(3, (16, 16, 16), (16, 16, 16), (0.0,) * 3, (1.0,) * 3, ("periodic",) * 6)
    _compare(par, mesh, problems.orszag_tang_3d(mesh, gamma=par.gamma), nsteps=4)


@pytest.mark.parametrize("type_ct", ["uct_contact", "uct_hll"])
def test_single_block_constrained_transport(type_ct):
    print(f"[edge] CT {type_ct}, one periodic block", flush=True)
    par = Par(physics="mhd", ndim=2, ndir=3, gamma=5.0 / 3.0, typedivbfix="ct", type_ct=type_ct,
              time_stepper="threestep", flux_scheme="hll", limiter="cada3", typeboundspeed="cmaxmean",
              courantpar=0.9)
    mesh = UniformMesh(2, (16, 16), (16, 16), (-0.5, -0.5), (0.5, 0.5), ("periodic",) * 4)
    _compare(par, mesh, problems.field_loop_ct(mesh))
