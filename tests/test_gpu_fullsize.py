"""The benchmark configurations at their full sizes: the device against
the CPU oracle directly (KH 2-D 256^2 for its 100 steps, OT 3-D 256^3 for
two steps, the AMR blast at 512^2 effective for ten steps), and through
properties that hold at any size over the configured step counts: the
production FP64 kernels against the contraction-free ones (which equal the
oracle bit for bit), the mixed form's per-variable error, and conservation
on periodic domains or with refluxing across level jumps."""
from __future__ import annotations

import numpy as np
import pytest

from amrvac_amd import problems
from amrvac_amd.host import Par, TimeIntegrator
from amrvac_amd.mesh import UniformMesh

pytestmark = pytest.mark.gpu


def _run(par, mesh, w0, nsteps, precision, state=False):
    from amrvac_amd.fv import DeviceFV
    be = DeviceFV(par, mesh, precision=precision)
    be.set_state(0, w0)
    ti = TimeIntegrator(par, mesh, be)
    m0 = be.volume_average(0)
    ti.run(nsteps=nsteps)
    m1 = be.volume_average(0)
    w = mesh.interior(be.get_state(0))
    del be
    if state:
        return w, m0, m1, ti.state.dts, ti.state
    return w, m0, m1, ti.state.dts


def _rel_l1(a, b):
    return float(np.abs(a - b).sum() / np.abs(b).sum())


def _rel_l1_vec(a, b):
    """Relative L1 error of a vector field (components on axis 1): the
    summed component errors over the summed vector magnitude."""
    return float(np.abs(a - b).sum() / np.sqrt((b * b).sum(axis=1)).sum())


def _run_oracle(par, mesh, w0, nsteps):
    from oracle.oracle import OracleFV
    be = OracleFV(par, mesh, threads=16)
    be.set_state(0, w0)
    be.getbc(0)
    ti = TimeIntegrator(par, mesh, be)
    ti.run(nsteps=nsteps)
    w = mesh.interior(be.get_state(0))
    del be
    return w, ti.state.dts


def _kh_setup():
    mesh = UniformMesh(2, (256, 256), (16, 16), (0.0, 0.0), (1.0, 1.0), ("periodic",) * 4)
    par = Par(physics="hd", ndim=2, ndir=2, gamma=1.4, time_stepper="threestep", flux_scheme="hllc",
              limiter="vanleer", courantpar=0.8)
    return par, mesh, problems.kelvin_helmholtz_2d(mesh)


def test_kh_2d_256_against_oracle_100_steps():
    """KH 2-D 256^2 (the first benchmark configuration) for its configured
    100 steps against the oracle: the contraction-free build bit for bit
    (state and dt), the production FP64 build to 1e-10 relative L1 per
    variable."""
    print("[fullsize] KH 2-D 256^2, 100 steps vs oracle", flush=True)
    par, mesh, w0 = _kh_setup()
    wo, dto = _run_oracle(par, mesh, w0, 100)
    wx, _, _, dtx = _run(par, mesh, w0, 100, "fp64_exact")
    wp, _, _, dtp = _run(par, mesh, w0, 100, "fp64")
    assert dtx == dto
    assert np.array_equal(wx, wo)
    assert np.allclose(dtp, dto, rtol=1e-12, atol=0.0)
    for iw in range(4):
        assert _rel_l1(wp[:, iw], wo[:, iw]) < 1e-10


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


def _ot_setup():
    mesh = UniformMesh(3, (256, 256, 256), (16, 16, 16), (0.0,) * 3, (1.0,) * 3, ("periodic",) * 6)
    par = Par(physics="mhd", ndim=3, ndir=3, gamma=5.0 / 3.0, typedivbfix="glm", time_stepper="threestep",
              flux_scheme="hll", limiter="cada3", typeboundspeed="Einfeldt", courantpar=0.6)
    return par, mesh, problems.orszag_tang_3d(mesh)


def test_orszag_tang_256_against_oracle_2_steps():
    """The benchmark workload (OT 3-D MHD-GLM 256^3) against the oracle for
    two full steps: contraction-free build bit for bit, production FP64 to
    1e-10 relative L1 per variable."""
    print("[fullsize] OT 3-D MHD-GLM 256^3, 2 steps vs oracle", flush=True)
    par, mesh, w0 = _ot_setup()
    wx, _, _, dtx = _run(par, mesh, w0, 2, "fp64_exact")
    wp, _, _, dtp = _run(par, mesh, w0, 2, "fp64")
    wo, dto = _run_oracle(par, mesh, w0, 2)
    assert dtx == dto
    assert np.array_equal(wx, wo)
    del wx
    assert np.allclose(dtp, dto, rtol=1e-12, atol=0.0)
    for iw in range(9):
        assert _rel_l1(wp[:, iw], wo[:, iw]) < 1e-10


def test_orszag_tang_256_configured_50_steps():
    """The benchmark workload for its configured 50 steps: production FP64
    (FMA contraction, refined hardware reciprocals) against the
    contraction-free build (the oracle's rounding, bit for bit) and the
    same dt sequence to 1e-12; the mixed form's error per variable (density
    within 1e-5, the target; the others within the bounds measured on this
    workload: momenta and fields 1e-3, psi 5e-2) and its mass and energy
    drift.

    The FP64 bound is 1e-10 relative L1 for every variable of the state's
    own scale (rho, m1, m2, e, b1, b2; measured 1e-12).  The z components
    m3 and b3 hold only the seeded 1e-5-scale perturbation, and psi only
    the cleaning wave: rounding-level differences, amplified by the
    turbulent flow over 150 stages, reach 3e-8 (m3, b3) and 6e-7 (psi) of
    those small fields.  They are bounded against the scale of their own
    equation: m3 and b3 by the momentum and field vectors (summed component
    errors over the summed vector magnitude, measured 4e-12), psi by
    c_h |B| (psi carries units of field times speed)."""
    print("[fullsize] OT 3-D MHD-GLM 256^3, 50 steps", flush=True)
    par, mesh, w0 = _ot_setup()
    wx, m0, m1, dtx, stx = _run(par, mesh, w0, 50, "fp64_exact", state=True)
    wp, _, _, dtp = _run(par, mesh, w0, 50, "fp64")
    assert np.allclose(dtp, dtx, rtol=1e-12, atol=0.0)
    e64 = {iw: _rel_l1(wp[:, iw], wx[:, iw]) for iw in range(9)}
    vec = {3: _rel_l1_vec(wp[:, 1:4], wx[:, 1:4]), 7: _rel_l1_vec(wp[:, 5:8], wx[:, 5:8])}
    bmag = np.sqrt((wx[:, 5:8] ** 2).sum(axis=1)).sum()
    psi = float(np.abs(wp[:, 8] - wx[:, 8]).sum() / (stx.cmax_global * bmag))
    print("[fullsize] fp64 vs exact rel L1 per variable", ["%.2e" % e64[iw] for iw in range(9)],
          "momentum / field vectors %.2e %.2e, psi over c_h|B| %.2e" % (vec[3], vec[7], psi), flush=True)
    for iw in (0, 1, 2, 4, 5, 6):
        assert e64[iw] < 1e-10, iw
    assert vec[3] < 1e-10 and vec[7] < 1e-10 and psi < 1e-10
    del wp
    wm, mm0, mm1, _ = _run(par, mesh, w0, 50, "mixed")
    err = [_rel_l1(wm[:, iw], wx[:, iw]) for iw in range(9)]
    print("[fullsize] mixed rel L1 per variable", ["%.2e" % e for e in err], flush=True)
    assert err[0] < 1e-5
    assert max(err[1:8]) < 1e-3 and err[8] < 5e-2
    # mass is conserved to round-off in both forms (the volume sum over
    # 16.7 M cells itself rounds at about 1e-13 relative: measured drifts
    # 1.0e-13 in both forms); energy changes through the GLM sources only,
    # and the two forms' changes agree closely
    de64, demx = (m1[4] - m0[4]) / abs(m0[4]), (mm1[4] - mm0[4]) / abs(m0[4])
    print(f"[fullsize] relative energy drift fp64 {de64:.3e} mixed {demx:.3e}", flush=True)
    assert abs(m1[0] - m0[0]) <= 1e-12 * abs(m0[0]) and abs(mm1[0] - mm0[0]) <= 1e-12 * abs(mm0[0])
    assert abs(demx - de64) <= 1e-6


def _amr_blast_setup(noise):
    from amrvac_amd.amr import AMRForest, AMRPar
    par = Par(physics="hd", ndim=2, ndir=2, gamma=5.0 / 3.0, time_stepper="twostep", flux_scheme="tvdlf",
              limiter="minmod", typeboundspeed="cmaxmean", courantpar=0.9)
    forest = AMRForest(2, (64, 64), (16, 16), (0.0, 0.0), (1.0, 1.0), ("cont",) * 4, max_level=4)
    ap = AMRPar(refine_max_level=4, w_refine_weight={0: 0.5, 3: 0.5})

    def init(f, b):
        return problems.blast_wave_blocks(f, b, rho0=1.0, p_out=1.0e-2, p_in=10.0, radius=0.05, noise=noise)
    return par, forest, ap, init


def test_amr_blast_512_against_oracle_10_steps():
    """The AMR blast configuration (64^2 base, 4 levels, 1% seeded density
    noise, Löhner regrid every step) against the oracle for ten steps: the
    contraction-free build bit for bit, with the same leaves and dt."""
    from amrvac_amd.amr import AMRIntegrator
    from amrvac_amd.fv import DeviceFV
    from oracle.oracle import OracleFV
    print("[fullsize] AMR blast 512^2 effective, 10 steps vs oracle", flush=True)
    out = {}
    for name, mk in (("device", lambda p, f: DeviceFV(p, f, precision="fp64_exact")),
                     ("oracle", lambda p, f: OracleFV(p, f, threads=16))):
        par, forest, ap, init = _amr_blast_setup(0.01)
        ti = AMRIntegrator(par, forest, ap, mk, init)
        ti.run(nsteps=10)
        out[name] = ([n.ig for n in forest.leaves], ti.state.dts, forest.interior(ti.be.get_state(0)))
    assert out["device"][0] == out["oracle"][0]
    assert out["device"][1] == out["oracle"][1]
    assert np.array_equal(out["device"][2], out["oracle"][2])


def test_amr_blast_512_mixed_200_steps():
    """The AMR blast configuration as specified (1% noise: every block
    refines to level 4) for its 200 steps in the mixed form against FP64:
    per-variable relative L1 (density within 1e-5), and mass and energy
    conserved to round-off in both forms (the blast stays inside the
    domain)."""
    from amrvac_amd.amr import AMRIntegrator
    from amrvac_amd.fv import DeviceFV
    print("[fullsize] AMR blast 512^2 effective, 200 steps, mixed vs FP64", flush=True)
    res = {}
    for prec in ("fp64", "mixed"):
        par, forest, ap, init = _amr_blast_setup(0.01)
        ti = AMRIntegrator(par, forest, ap, lambda p, f: DeviceFV(p, f, precision=prec), init)
        m0 = ti.be.volume_average(0)
        ti.run(nsteps=200)
        m1 = ti.be.volume_average(0)
        res[prec] = ([n.ig for n in forest.leaves], forest.interior(ti.be.get_state(0)), m0, m1)
    assert res["fp64"][0] == res["mixed"][0]
    wp, wm = res["fp64"][1], res["mixed"][1]
    err = [_rel_l1(wm[:, iw], wp[:, iw]) for iw in range(4)]
    print("[fullsize] AMR mixed rel L1 per variable", ["%.2e" % e for e in err], flush=True)
    assert err[0] < 1e-5 and err[3] < 1e-5
    for prec in ("fp64", "mixed"):
        _, _, m0, m1 = res[prec]
        assert abs(m1[0] - m0[0]) <= 1e-12 * abs(m0[0])
        assert abs(m1[3] - m0[3]) <= 1e-12 * abs(m0[3])


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
