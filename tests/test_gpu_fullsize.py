"""The benchmark configurations at their full sizes, checked through
properties that hold at any size (the CPU oracle is too slow there):
the production FP64 kernels against the contraction-free FP64 kernels
(which equal the oracle bit for bit at every size the oracle runs),
the mixed form's density error, and conservation on periodic domains or
with refluxing across level jumps."""
from __future__ import annotations

import numpy as np
import pytest

from amrvac_amd import problems
from amrvac_amd.host import Par, TimeIntegrator
from amrvac_amd.mesh import UniformMesh

pytestmark = pytest.mark.gpu


def _run(par, mesh, w0, nsteps, precision):
    from amrvac_amd.fv import DeviceFV
    be = DeviceFV(par, mesh, precision=precision)
    be.set_state(0, w0)
    ti = TimeIntegrator(par, mesh, be)
    m0 = be.volume_average(0)
    ti.run(nsteps=nsteps)
    m1 = be.volume_average(0)
    w = mesh.interior(be.get_state(0))
    del be
    return w, m0, m1, ti.state.dts


def _rel_l1(a, b):
    return float(np.abs(a - b).sum() / np.abs(b).sum())


def test_kh_2d_256_hllc_vanleer_100_steps():
    """2-D HD Kelvin-Helmholtz 256^2 in 16^2 blocks, threestep / HLLC /
    vanleer, CFL 0.8, 100 steps."""
    print("[fullsize] KH 2-D 256^2, 100 steps", flush=True)
    mesh = UniformMesh(2, (256, 256), (16, 16), (0.0, 0.0), (1.0, 1.0), ("periodic",) * 4)
    par = Par(physics="hd", ndim=2, ndir=2, gamma=1.4, time_stepper="threestep", flux_scheme="hllc",
              limiter="vanleer", courantpar=0.8)
    w0 = problems.kelvin_helmholtz_2d(mesh)
    wx, m0, m1, dtx = _run(par, mesh, w0, 100, "fp64_exact")
    wp, _, _, _ = _run(par, mesh, w0, 100, "fp64")
    wm, _, _, _ = _run(par, mesh, w0, 100, "mixed")
    for iw in range(4):
        assert _rel_l1(wp[:, iw], wx[:, iw]) < 1e-10
    assert _rel_l1(wm[:, 0], wx[:, 0]) < 1e-5
    # conservative update on a periodic domain: mass, momenta, energy
    for iw in range(4):
        assert abs(m1[iw] - m0[iw]) <= 1e-13 * max(abs(m0[iw]), 1.0)


def test_orszag_tang_mhd_glm_3d_256():
    """3-D ideal MHD + GLM Orszag-Tang 256^3 in 16^3 blocks, threestep / HLL
    / cada3, CFL 0.6 (the benchmark workload), 5 steps."""
    print("[fullsize] OT 3-D MHD-GLM 256^3, 5 steps", flush=True)
    mesh = UniformMesh(3, (256, 256, 256), (16, 16, 16), (0.0,) * 3, (1.0,) * 3, ("periodic",) * 6)
    par = Par(physics="mhd", ndim=3, ndir=3, gamma=5.0 / 3.0, typedivbfix="glm", time_stepper="threestep",
              flux_scheme="hll", limiter="cada3", typeboundspeed="Einfeldt", courantpar=0.6)
    w0 = problems.orszag_tang_3d(mesh)
    wx, m0, m1, dtx = _run(par, mesh, w0, 5, "fp64_exact")
    wp, _, _, dtp = _run(par, mesh, w0, 5, "fp64")
    assert np.allclose(dtp, dtx, rtol=1e-12, atol=0.0)
    for iw in range(8):
        assert _rel_l1(wp[:, iw], wx[:, iw]) < 1e-10
    del wp
    wm, _, _, _ = _run(par, mesh, w0, 5, "mixed")
    assert _rel_l1(wm[:, 0], wx[:, 0]) < 1e-5
    # mass is conserved (the GLM sources act on momentum, energy, psi)
    assert abs(m1[0] - m0[0]) <= 1e-13 * abs(m0[0])


def test_amr_blast_2d_512_conserves_mass_with_refluxing():
    """2-D HD blast wave, 64^2 base in 16^2 blocks, 4 levels (512^2
    effective), Löhner regrid every step on the device, twostep / TVDLF /
    minmod, 200 steps without density noise (a tree with level jumps that
    follows the blast).  The blast stays inside the domain, so the total
    mass is conserved to round-off only if refluxing, restriction and the
    regrid's coarsening conserve it."""
    from amrvac_amd.amr import AMRForest, AMRIntegrator, AMRPar
    from amrvac_amd.fv import DeviceFV
    print("[fullsize] AMR blast 2-D 512^2 effective, 200 steps", flush=True)
    par = Par(physics="hd", ndim=2, ndir=2, gamma=5.0 / 3.0, time_stepper="twostep", flux_scheme="tvdlf",
              limiter="minmod", typeboundspeed="cmaxmean", courantpar=0.9)
    forest = AMRForest(2, (64, 64), (16, 16), (0.0, 0.0), (1.0, 1.0), ("cont",) * 4, max_level=4)
    ap = AMRPar(refine_max_level=4, w_refine_weight={0: 0.5, 3: 0.5})
    init = lambda f, b: problems.blast_wave_blocks(f, b, rho0=1.0, p_out=1.0e-2, p_in=10.0, radius=0.05)
    ti = AMRIntegrator(par, forest, ap, lambda p, f: DeviceFV(p, f), init)
    m0 = ti.be.volume_average(0)
    ti.run(nsteps=200)
    m1 = ti.be.volume_average(0)
    assert ti.regrids > 0 and forest.levmax == 4 and forest.levmin < 4
    assert abs(m1[0] - m0[0]) <= 1e-12 * abs(m0[0])
    assert abs(m1[3] - m0[3]) <= 1e-12 * abs(m0[3])
