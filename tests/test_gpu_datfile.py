"""Snapshots written straight from the device state and restarts of device
runs from them: a restarted run continues bit for bit."""
from __future__ import annotations

import os

import numpy as np
import pytest

from amrvac_amd import datfile, problems
from amrvac_amd.host import Par, TimeIntegrator
from amrvac_amd.mesh import UniformMesh

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_restart_continues_bitwise(tmp_path, precision):
    from amrvac_amd.fv import DeviceFV
    print(f"[datfile] 3-D MHD-GLM restart, {precision}", flush=True)
    par = Par(physics="mhd", ndim=3, ndir=3, gamma=5.0 / 3.0, typedivbfix="glm", time_stepper="threestep",
              flux_scheme="hll", limiter="cada3", courantpar=0.6)
    mesh = UniformMesh(3, (32, 32, 32), (16, 16, 16), (0.0,) * 3, (1.0,) * 3, ("periodic",) * 6)
    w0 = problems.orszag_tang_3d(mesh)
    ref = DeviceFV(par, mesh, precision=precision)
    ref.set_state(0, w0)
    TimeIntegrator(par, mesh, ref).run(nsteps=6)
    be = DeviceFV(par, mesh, precision=precision)
    be.set_state(0, w0)
    ti = TimeIntegrator(par, mesh, be)
    ti.run(nsteps=3)
    path = os.path.join(tmp_path, "ot0000.dat")
    snap = datfile.write_snapshot(path, par, mesh, be.get_state(0), it=ti.state.it,
                                  global_time=ti.state.global_time)
    assert snap.w_names[-1] == "psi" and snap.geometry == "Cartesian_3D"
    be2 = DeviceFV(par, mesh, precision=precision)
    be2.set_state(0, datfile.state_from_snapshot(datfile.read_snapshot(path), mesh))
    ti2 = TimeIntegrator(par, mesh, be2)
    ti2.state.it, ti2.state.global_time = snap.it, snap.global_time
    ti2.run(nsteps=3)
    assert np.array_equal(mesh.interior(be2.get_state(0)), mesh.interior(ref.get_state(0)))


def test_device_amr_snapshot_restart(tmp_path):
    from amrvac_amd.amr import AMRForest, AMRIntegrator, AMRPar
    from amrvac_amd.fv import DeviceFV
    print("[datfile] 2-D AMR blast restart", flush=True)
    par = Par(physics="hd", ndim=2, ndir=2, gamma=5.0 / 3.0, time_stepper="twostep", flux_scheme="tvdlf",
              limiter="minmod", typeboundspeed="cmaxmean", courantpar=0.9)

    def setup():
        forest = AMRForest(2, (32, 32), (16, 16), (0.0, 0.0), (2.0, 2.0), ("cont",) * 4, max_level=3)
        ap = AMRPar(refine_max_level=3, w_refine_weight={0: 0.5, 3: 0.5})
        init = lambda f, b: problems.blast_wave_blocks(f, b)
        return AMRIntegrator(par, forest, ap, lambda p, f: DeviceFV(p, f, precision="fp64_exact"), init), forest, ap

    ref, f_ref, _ = setup()
    ref.run(nsteps=8)
    ti, forest, ap = setup()
    ti.run(nsteps=4)
    path = os.path.join(tmp_path, "bw0000.dat")
    snap = datfile.write_snapshot(path, par, forest, ti.be.get_state(0), it=ti.state.it,
                                  global_time=ti.state.global_time)
    f2 = datfile.forest_from_snapshot(datfile.read_snapshot(path), ("cont",) * 4, max_level=3)
    ti2 = AMRIntegrator(par, f2, ap, lambda p, f: DeviceFV(p, f, precision="fp64_exact"),
                        lambda f, b: datfile.state_from_snapshot(snap, f)[b], restart=True)
    assert [(n.level, n.ig) for n in f2.leaves] == [(n.level, n.ig) for n in forest.leaves]
    ti2.state.it, ti2.state.global_time = snap.it, snap.global_time
    ti2.run(nsteps=4)
    assert [(n.level, n.ig) for n in f2.leaves] == [(n.level, n.ig) for n in f_ref.leaves]
    assert np.array_equal(f2.interior(ti2.be.get_state(0)), f_ref.interior(ref.be.get_state(0)))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_staggered_restart_continues_bitwise(tmp_path, precision):
    """Constrained transport: snapshot of the device's cells and face field,
    restart, and a bit-for-bit continuation."""
    import sys
    sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
    from test_oracle_golden import field_loop_setup
    from amrvac_amd.fv import DeviceFV
    print(f"[datfile] 2.5-D CT field-loop restart, {precision}", flush=True)
    par, mesh = field_loop_setup("3step_hll_cada_ct_hll")
    w0 = problems.field_loop_ct(mesh)
    ref = DeviceFV(par, mesh, precision=precision)
    ref.set_state(0, w0)
    TimeIntegrator(par, mesh, ref).run(nsteps=6)
    be = DeviceFV(par, mesh, precision=precision)
    be.set_state(0, w0)
    ti = TimeIntegrator(par, mesh, be)
    ti.run(nsteps=3)
    path = os.path.join(tmp_path, "fl0000.dat")
    snap = datfile.write_snapshot(path, par, mesh, be.get_state(0), it=ti.state.it,
                                  global_time=ti.state.global_time)
    assert snap.staggered
    be2 = DeviceFV(par, mesh, precision=precision)
    be2.set_state(0, datfile.state_from_snapshot(datfile.read_snapshot(path), mesh))
    ti2 = TimeIntegrator(par, mesh, be2)
    ti2.state.it, ti2.state.global_time = snap.it, snap.global_time
    ti2.run(nsteps=3)
    assert np.array_equal(mesh.interior(be2.get_state(0)), mesh.interior(ref.get_state(0)))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_device_staggered_restart_continues_bitwise_in_3d(tmp_path, precision):
    """Constrained transport in 3-D (the Gardiner & Stone field loop):
    snapshot of the device's cells and three face components, restart, and
    a bit-for-bit continuation."""
    from amrvac_amd.fv import DeviceFV
    print(f"[datfile] 3-D CT field-loop restart, {precision}", flush=True)
    par = Par(physics="mhd", ndim=3, ndir=3, gamma=5.0 / 3.0, typedivbfix="ct", type_ct="uct_contact",
              time_stepper="threestep", flux_scheme="hll", limiter="cada3", typeboundspeed="cmaxmean",
              courantpar=0.4)
    mesh = UniformMesh(3, (32, 32, 64), (16, 16, 16), (-0.5, -0.5, -1.0), (0.5, 0.5, 1.0), ("periodic",) * 6)
    w0 = problems.field_loop_3d_ct(mesh)
    ref = DeviceFV(par, mesh, precision=precision)
    ref.set_state(0, w0)
    TimeIntegrator(par, mesh, ref).run(nsteps=4)
    be = DeviceFV(par, mesh, precision=precision)
    be.set_state(0, w0)
    ti = TimeIntegrator(par, mesh, be)
    ti.run(nsteps=2)
    path = os.path.join(tmp_path, "fl3d0000.dat")
    snap = datfile.write_snapshot(path, par, mesh, be.get_state(0), it=ti.state.it,
                                  global_time=ti.state.global_time)
    assert snap.staggered and snap.ws.shape[1:] == (3, 17, 17, 17)
    be2 = DeviceFV(par, mesh, precision=precision)
    be2.set_state(0, datfile.state_from_snapshot(datfile.read_snapshot(path), mesh))
    ti2 = TimeIntegrator(par, mesh, be2)
    ti2.state.it, ti2.state.global_time = snap.it, snap.global_time
    ti2.run(nsteps=2)
    assert np.array_equal(mesh.interior(be2.get_state(0)), mesh.interior(ref.get_state(0)))
