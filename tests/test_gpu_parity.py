"""Parity of the HIP stage kernels (through the C-ABI) with the CPU oracle.

Three precisions are checked against the oracle run of the same setup:
  fp64_exact  FP64 kernels built with -ffp-contract=off: bit-identical state;
  fp64        production FP64 kernels (FMA contraction): relative L1 <= 1e-10;
  mixed       FP32 reconstruction/Riemann, FP64 state: density L1 <= 1e-5.
Sizes are chosen so the oracle finishes in seconds.
"""
import os

import numpy as np
import pytest

from amrvac_amd import problems
from amrvac_amd.host import Par, TimeIntegrator, logs_match, read_log
from amrvac_amd.mesh import UniformMesh
from oracle.oracle import OracleFV

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def device_backend(par, mesh, precision):
    from amrvac_amd.fv import DeviceFV
    return DeviceFV(par, mesh, precision=precision)


def rel_l1(a, b):
    return float(np.abs(a - b).sum() / max(np.abs(b).sum(), 1e-300))


def run(par, mesh, w0, backend, nsteps):
    backend.set_state(0, w0)
    backend.getbc(0)
    ti = TimeIntegrator(par, mesh, backend)
    ti.run(nsteps=nsteps)
    return mesh.interior(backend.get_state(0)), ti.state


CASES = {
    # config 0 shape at reduced size: 2-D Euler KH, ssprk3 / HLLC / vanleer
    "kh2d_hllc_vanleer": dict(
        par=dict(physics="hd", ndim=2, ndir=2, gamma=1.4, time_stepper="threestep", flux_scheme="hllc",
                 limiter="vanleer", courantpar=0.8),
        mesh=dict(ndim=2, domain_nx=(64, 64), block_nx=(16, 16), xprobmin=(0.0, 0.0), xprobmax=(1.0, 1.0),
                  bc=("periodic",) * 4),
        ic="kh", steps=10),
    # predictor-corrector TVDLF / minmod with Hancock predictor, cont boundaries
    "blast2d_tvdlf_minmod": dict(
        par=dict(physics="hd", ndim=2, ndir=2, gamma=5.0 / 3.0, time_stepper="twostep", flux_scheme="tvdlf",
                 limiter="minmod", courantpar=0.8),
        mesh=dict(ndim=2, domain_nx=(64, 48), block_nx=(16, 16), xprobmin=(0.0, 0.0), xprobmax=(1.0, 0.75),
                  bc=("cont",) * 4),
        ic="blast", steps=10),
    "kh2d_hll_cada3_einfeldt": dict(
        par=dict(physics="hd", ndim=2, ndir=2, gamma=1.4, time_stepper="threestep", flux_scheme="hll",
                 limiter="cada3", courantpar=0.8),
        mesh=dict(ndim=2, domain_nx=(32, 64), block_nx=(16, 16), xprobmin=(0.0, 0.0), xprobmax=(0.5, 1.0),
                  bc=("periodic", "periodic", "symm", "symm")),
        ic="kh", steps=8),
    "kh2d_tvdlf_weno5": dict(
        par=dict(physics="hd", ndim=2, ndir=2, gamma=1.4, time_stepper="threestep", flux_scheme="tvdlf",
                 limiter="weno5", courantpar=0.6),
        mesh=dict(ndim=2, domain_nx=(64, 64), block_nx=(16, 16), xprobmin=(0.0, 0.0), xprobmax=(1.0, 1.0),
                  bc=("periodic",) * 4),
        ic="kh", steps=6),
    # config 1 shape at reduced size: 3-D ideal MHD + GLM, ssprk3 / HLL / cada3
    "ot3d_glm_hll_cada3": dict(
        par=dict(physics="mhd", ndim=3, ndir=3, typedivbfix="glm", time_stepper="threestep", flux_scheme="hll",
                 limiter="cada3", courantpar=0.6),
        mesh=dict(ndim=3, domain_nx=(32, 32, 32), block_nx=(16, 16, 16), xprobmin=(0.0,) * 3,
                  xprobmax=(1.0,) * 3, bc=("periodic",) * 6),
        ic="ot3d", steps=4),
    "ot3d_glm_tvdlf_vanleer_cmaxmean": dict(
        par=dict(physics="mhd", ndim=3, ndir=3, typedivbfix="glm", time_stepper="threestep",
                 flux_scheme="tvdlf", limiter="vanleer", typeboundspeed="cmaxmean", courantpar=0.6),
        mesh=dict(ndim=3, domain_nx=(32, 16, 32), block_nx=(16, 16, 8), xprobmin=(0.0,) * 3,
                  xprobmax=(1.0, 0.5, 1.0), bc=("periodic",) * 6),
        ic="ot3d", steps=3),
    "alfven2d_linde_hll_cada3": dict(
        par=dict(physics="mhd", ndim=2, ndir=3, typedivbfix="linde", time_stepper="threestep",
                 flux_scheme="hll", limiter="cada3", typeboundspeed="cmaxmean", courantpar=0.9),
        mesh=dict(ndim=2, domain_nx=(64, 32), block_nx=(16, 16), xprobmin=(0.0, 0.0), xprobmax=(1.0, 0.5),
                  bc=("periodic",) * 4),
        ic="alfven", steps=6),
    "ot2d_glm_hll_vanleer": dict(
        par=dict(physics="mhd", ndim=2, ndir=3, typedivbfix="glm", time_stepper="threestep",
                 flux_scheme="hll", limiter="vanleer", courantpar=0.8),
        mesh=dict(ndim=2, domain_nx=(64, 64), block_nx=(16, 16), xprobmin=(0.0, 0.0), xprobmax=(1.0, 1.0),
                  bc=("periodic",) * 4),
        ic="ot2d", steps=6),
    "hd3d_hllc_minmod": dict(
        par=dict(physics="hd", ndim=3, ndir=3, gamma=5.0 / 3.0, time_stepper="threestep", flux_scheme="hllc",
                 limiter="minmod", courantpar=0.6),
        mesh=dict(ndim=3, domain_nx=(32, 32, 16), block_nx=(16, 16, 16), xprobmin=(0.0,) * 3,
                  xprobmax=(1.0, 1.0, 0.5), bc=("periodic", "periodic", "cont", "cont", "periodic", "periodic")),
        ic="blast3d", steps=4),
}


def make_ic(kind, par, mesh):
    if kind == "kh":
        return problems.kelvin_helmholtz_2d(mesh, gamma=par.gamma)
    if kind == "ot3d":
        return problems.orszag_tang_3d(mesh, gamma=par.gamma)
    if kind == "alfven":
        return problems.alfven_wave_2p5d(mesh, gamma=par.gamma)
    if kind == "ot2d":
        return problems.orszag_tang_2p5d(mesh, gamma=par.gamma)
    if kind == "blast":
        return problems.blast_wave(mesh, gamma=par.gamma)
    if kind == "blast3d":
        return problems.blast_wave(mesh, gamma=par.gamma, p_in=1.0, radius=0.2)
    raise ValueError(kind)


def setup(name):
    c = CASES[name]
    par = Par(**c["par"])
    mesh = UniformMesh(nghost=par.nghost, **c["mesh"])
    return par, mesh, make_ic(c["ic"], par, mesh), c["steps"]


@pytest.fixture(scope="module")
def oracle_results():
    return {}


def oracle_run(cache, name):
    if name not in cache:
        par, mesh, w0, steps = setup(name)
        cache[name] = run(par, mesh, w0, OracleFV(par, mesh), steps)
    return cache[name]


@pytest.mark.parametrize("name", sorted(CASES))
def test_fp64_exact_is_bit_identical(name, oracle_results):
    print(f"[gpu] exact parity {name}", flush=True)
    par, mesh, w0, steps = setup(name)
    ref, rst = oracle_run(oracle_results, name)
    got, gst = run(par, mesh, w0, device_backend(par, mesh, "fp64_exact"), steps)
    assert gst.dts == rst.dts, "dt sequence differs"
    diff = np.abs(got - ref)
    assert np.array_equal(got, ref), f"max |diff| {diff.max():.3e}, rel L1 {rel_l1(got, ref):.3e}"


@pytest.mark.parametrize("name", sorted(CASES))
def test_fp64_within_1e10(name, oracle_results):
    print(f"[gpu] fp64 parity {name}", flush=True)
    par, mesh, w0, steps = setup(name)
    ref, _ = oracle_run(oracle_results, name)
    got, _ = run(par, mesh, w0, device_backend(par, mesh, "fp64"), steps)
    for iw in range(par.nw):
        if np.abs(ref[:, iw]).sum() == 0.0:
            assert np.abs(got[:, iw]).max() < 1e-14
            continue
        assert rel_l1(got[:, iw], ref[:, iw]) <= 1e-10, f"variable {iw}"


@pytest.mark.parametrize("name", sorted(CASES))
def test_mixed_density_within_1e5(name, oracle_results):
    print(f"[gpu] mixed parity {name}", flush=True)
    par, mesh, w0, steps = setup(name)
    ref, _ = oracle_run(oracle_results, name)
    got, _ = run(par, mesh, w0, device_backend(par, mesh, "mixed"), steps)
    assert np.all(np.isfinite(got))
    assert rel_l1(got[:, 0], ref[:, 0]) <= 1e-5


def test_alfven_golden_on_device():
    """The reference's golden log of tests/mhd/Alfven_wave_2.5D
    (2step_tvdlf_mm) reproduced by the device path."""
    from test_oracle_golden import run_alfven
    from amrvac_amd.fv import DeviceFV
    log = run_alfven("2step_tvdlf_mm", lambda p, m: DeviceFV(p, m, precision="fp64"))
    gold = read_log(os.path.join(GOLD, "aw_2d_2step_tvdlf_mm.log"))
    assert logs_match(log, gold)


@pytest.mark.parametrize("scheme", ["3step_hll_cada", "3step_hlld_cada"])
def test_field_loop_golden_on_device(scheme):
    """tests/mhd/field_loop_2.5D: HLL with Linde cleaning, HLLD with GLM;
    the contraction-free kernels equal the oracle's log exactly, the
    production kernels match the reference's golden log."""
    from test_oracle_golden import run_field_loop
    from amrvac_amd.fv import DeviceFV
    print(f"[parity] field loop {scheme}", flush=True)
    log = run_field_loop(lambda p, m: DeviceFV(p, m, precision="fp64"), scheme=scheme)
    gold = read_log(os.path.join(GOLD, f"fl_2d_{scheme}.log"))
    assert logs_match(log, gold)
    exact = run_field_loop(lambda p, m: DeviceFV(p, m, precision="fp64_exact"), scheme=scheme)
    orc = run_field_loop(lambda p, m: OracleFV(p, m, threads=4), scheme=scheme)
    assert np.allclose(exact, orc, rtol=1e-12, atol=1e-15)


def test_download_ghosts_match_oracle_getbc():
    par, mesh, w0, _ = setup("blast2d_tvdlf_minmod")
    o = OracleFV(par, mesh)
    o.set_state(0, w0)
    o.getbc(0)
    d = device_backend(par, mesh, "fp64")
    d.set_state(0, w0)
    assert np.array_equal(d.get_state(0), o.get_state(0))


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_fused_cfl_partials_leave_the_dt_sequence_unchanged(precision):
    """On a uniform level the mixed-form stage that writes ps also leaves the
    per-block CFL maxima of its output, and the next getdt only reduces
    them (the FP64 form keeps the separate pass).  The dt sequence and the
    state must equal those of the separate dt pass bit for bit (a context
    whose ps was handed out by amrvac_fv_slot_pointer keeps to the separate
    pass)."""
    from amrvac_amd.fv import DeviceFV
    print(f"[parity] fused CFL partials, 3-D MHD-GLM {precision}", flush=True)
    par = Par(physics="mhd", ndim=3, ndir=3, gamma=5.0 / 3.0, typedivbfix="glm", time_stepper="threestep",
              flux_scheme="hll", limiter="cada3", courantpar=0.6)
    mesh = UniformMesh(3, (64, 64, 48), (16, 16, 16), (0.0,) * 3, (1.0, 1.0, 0.75), ("periodic",) * 6)
    w0 = problems.orszag_tang_3d(mesh)
    out = {}
    for fused in (True, False):
        be = DeviceFV(par, mesh, precision=precision)
        if not fused:
            be.slot_pointer(0)
        w, st = run(par, mesh, w0, be, 6)
        out[fused] = (w, st.dts, st.cmax_global)
    assert out[True][1] == out[False][1]
    assert out[True][2] == out[False][2]
    assert np.array_equal(out[True][0], out[False][0])
