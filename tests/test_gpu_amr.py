"""AMR on the device against the CPU oracle: ghost cells across level jumps
(restriction, prolongation), refluxing, level-dependent cell sizes, with the
reference's regrid every step.  The contraction-free FP64 build must match
the oracle bit for bit, state and log; the production builds to the stated
tolerances."""
from __future__ import annotations

import os

import numpy as np
import pytest

from amrvac_amd import problems
from amrvac_amd.amr import AMRForest, AMRIntegrator, AMRPar
from amrvac_amd.host import Par, logs_match, read_log
from oracle.oracle import OracleFV

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")

SCHEMES = {
    "2step_tvdlf_mm": dict(time_stepper="twostep", flux_scheme="tvdlf", limiter="minmod"),
    "3step_hll_cada": dict(time_stepper="threestep", flux_scheme="hll", limiter="cada3"),
    "4step_hll_mc": dict(time_stepper="fourstep", time_integrator="rk4", flux_scheme="hll", limiter="woodward"),
    "4step_hllc_ko": dict(time_stepper="fourstep", time_integrator="ssprk4", flux_scheme="hllc", limiter="koren"),
    "rk4_tvdlf_cada": dict(time_stepper="fourstep", time_integrator="rk4", flux_scheme="tvdlf", limiter="cada3"),
}


def _device(precision):
    from amrvac_amd.fv import DeviceFV
    return lambda p, f: DeviceFV(p, f, precision=precision)


def _blast(scheme, time_max):
    par = Par(physics="hd", ndim=2, ndir=2, gamma=5.0 / 3.0, typeboundspeed="cmaxmean", courantpar=0.9,
              dtsave_log=1.0e-3, time_max=time_max, **SCHEMES[scheme])
    forest = AMRForest(2, (32, 32), (16, 16), (0.0, 0.0), (2.0, 2.0), ("cont",) * 4, max_level=3)
    ap = AMRPar(refine_max_level=3, w_refine_weight={0: 0.5, 3: 0.5})
    return par, forest, ap


def _init(forest, blocks):
    return problems.blast_wave_blocks(forest, blocks)


@pytest.mark.parametrize("scheme", sorted(SCHEMES))
def test_amr_blast_exact_matches_oracle_bitwise(scheme):
    print(f"[amr] blast 2-D {scheme} fp64_exact", flush=True)
    par, f_dev, ap = _blast(scheme, 0.05)
    ti_dev = AMRIntegrator(par, f_dev, ap, _device("fp64_exact"), _init)
    st_dev = ti_dev.run()
    par, f_orc, ap = _blast(scheme, 0.05)
    ti_orc = AMRIntegrator(par, f_orc, ap, lambda p, f: OracleFV(p, f, threads=4), _init)
    st_orc = ti_orc.run()
    assert st_dev.dts == st_orc.dts
    assert [n.ig for n in f_dev.leaves] == [n.ig for n in f_orc.leaves]
    assert np.array_equal(f_dev.interior(ti_dev.be.get_state(0)), f_orc.interior(ti_orc.be.get_state(0)))
    # the log's volume sums reduce in another order on the device
    assert np.allclose(np.array(st_dev.log), np.array(st_orc.log), rtol=1e-12, atol=1e-12)
    ref = read_log(os.path.join(GOLDEN, f"bw_2d_{scheme}.log"))
    assert logs_match(np.array(st_dev.log), ref, rtol=1e-7, atol=1e-8)


@pytest.mark.parametrize("scheme", sorted(SCHEMES))
def test_amr_blast_fp64_matches_golden_log(scheme):
    print(f"[amr] blast 2-D {scheme} fp64", flush=True)
    par, forest, ap = _blast(scheme, 0.05)
    ti = AMRIntegrator(par, forest, ap, _device("fp64"), _init)
    st = ti.run()
    ref = read_log(os.path.join(GOLDEN, f"bw_2d_{scheme}.log"))
    assert logs_match(np.array(st.log), ref, rtol=1e-7, atol=1e-8)


def test_amr_blast_mixed_matches_golden_log_loosely():
    print("[amr] blast 2-D 3step_hll_cada mixed", flush=True)
    par, forest, ap = _blast("3step_hll_cada", 0.05)
    ti = AMRIntegrator(par, forest, ap, _device("mixed"), _init)
    st = ti.run()
    ref = read_log(os.path.join(GOLDEN, "bw_2d_3step_hll_cada.log"))
    got = np.array(st.log)
    assert got.shape == ref.shape
    assert logs_match(got, ref, rtol=1e-4, atol=1e-8)


def test_amr_periodic_conservation_on_device():
    """Refluxing on the device: an advected density bump crosses coarse/fine
    faces; the volume integrals stay constant to round-off."""
    print("[amr] periodic bump fp64", flush=True)
    par = Par(physics="hd", ndim=2, ndir=2, gamma=1.4, time_stepper="threestep", flux_scheme="hllc",
              limiter="vanleer", courantpar=0.8)
    forest = AMRForest(2, (64, 64), (16, 16), (0.0, 0.0), (1.0, 1.0), ("periodic",) * 4, max_level=3)
    ap = AMRPar(refine_max_level=3, refine_threshold=0.1, w_refine_weight={0: 1.0})

    def init(f, blocks):
        x = f.cell_centers(blocks)
        r2 = (x[:, 0] - 0.5) ** 2 + (x[:, 1] - 0.5) ** 2
        rho = 1.0 + 0.5 * np.exp(-r2 / 0.01)
        one = np.ones_like(rho)
        return problems.hd_to_conserved(np.stack([rho, 1.0 * one, 0.5 * one, one], axis=1), 1.4, 2)

    ti = AMRIntegrator(par, forest, ap, _device("fp64"), init)
    assert forest.levmin == 1 and forest.levmax == 3

    def totals():
        w = ti.be.get_state(0)
        vol = forest.cell_volume()
        return (forest.interior(w).sum(axis=(2, 3)) * vol[:, None]).sum(axis=0)

    t0 = totals()
    ti.run(nsteps=12)
    t1 = totals()
    assert abs(t1[0] - t0[0]) < 1e-13 * abs(t0[0])
    assert abs(t1[3] - t0[3]) < 1e-13 * abs(t0[3])


def test_amr_3d_exact_matches_oracle_bitwise():
    """3-D level jumps (faces, edges, corners) with 16^3 blocks."""
    print("[amr] blast 3-D fp64_exact", flush=True)

    def setup():
        par = Par(physics="hd", ndim=3, ndir=3, gamma=5.0 / 3.0, typeboundspeed="cmaxmean", courantpar=0.9,
                  time_stepper="threestep", flux_scheme="hll", limiter="cada3")
        forest = AMRForest(3, (64, 64, 64), (16, 16, 16), (0.0,) * 3, (2.0,) * 3, ("cont",) * 6, max_level=2)
        ap = AMRPar(refine_max_level=2, w_refine_weight={0: 0.5, 4: 0.5})
        return par, forest, ap

    def init(f, blocks):
        return problems.blast_wave_blocks(f, blocks)

    par, f_dev, ap = setup()
    ti_dev = AMRIntegrator(par, f_dev, ap, _device("fp64_exact"), init)
    par, f_orc, ap = setup()
    ti_orc = AMRIntegrator(par, f_orc, ap, lambda p, f: OracleFV(p, f, threads=4), init)
    assert f_dev.levmin == 1 and f_dev.levmax == 2
    ti_dev.run(nsteps=4)
    ti_orc.run(nsteps=4)
    assert ti_dev.state.dts == ti_orc.state.dts
    assert np.array_equal(f_dev.interior(ti_dev.be.get_state(0)), f_orc.interior(ti_orc.be.get_state(0)))


@pytest.mark.parametrize("ndim,bx", [(2, 16), (3, 16), (3, 8), (3, 10)])
def test_device_errest_and_regrid_match_host(ndim, bx):
    """The device Löhner flags and remap (coarsen_grid, prolong_2nd) against
    the host's on the same ghosted state: flags identical, new leaves bit
    for bit (3-D: block_nx 16, 8 and 10, each estimated from an LDS tile of
    the block and its stencil rim)."""
    from amrvac_amd.amr import apply_regrid, coarsen_child, errest_flags, prolong_2nd, proper_nesting
    from amrvac_amd.fv import DeviceFV
    print(f"[amr] device errest/regrid vs host, {ndim}-D, block_nx {bx}", flush=True)
    if ndim == 2:
        par, forest, ap = _blast("3step_hll_cada", 0.02)
    else:
        par = Par(physics="hd", ndim=3, ndir=3, gamma=5.0 / 3.0, typeboundspeed="cmaxmean", courantpar=0.9,
                  time_stepper="twostep", flux_scheme="tvdlf", limiter="minmod")
        n = 2 * bx if bx == 16 else 4 * bx
        forest = AMRForest(3, (n, n, n), (bx, bx, bx), (0.0,) * 3, (2.0,) * 3, ("cont",) * 6, max_level=3)
        ap = AMRPar(refine_max_level=3, w_refine_weight={0: 0.5, 4: 0.5})
    init = lambda f, b: problems.blast_wave_blocks(f, b, radius=0.3)
    ti = AMRIntegrator(par, forest, ap, lambda p, f: DeviceFV(p, f, precision="fp64_exact"), init)
    checked = 0
    for _ in range(80):
        if checked >= 2:
            break
        ti.setdt()
        ti.advect()
        ti.state.it += 1
        ti.state.global_time += ti.state.dt
        be = ti.be
        w = be.get_state(0)
        r_dev, c_dev = be.errest_flags(forest, ap)
        r_host, c_host = errest_flags(forest, w, ap)
        assert np.array_equal(r_dev, r_host) and np.array_equal(c_dev, c_host)
        proper_nesting(forest, ap, r_dev, c_dev)
        if not r_dev.any() and not c_dev.any():
            continue
        plan = apply_regrid(forest, r_dev, c_dev)
        nd, ng = forest.ndim, forest.nghost
        new = np.zeros((forest.nblocks,) + w.shape[1:])
        for b, org in enumerate(plan):
            if org[0] == "copy":
                new[b] = w[org[1]]
            elif org[0] == "coarsen":
                for j, ob in enumerate(org[1]):
                    coarsen_child(w[ob], tuple(1 + ((j >> d) & 1) for d in range(nd)), new[b], nd, ng)
            else:
                prolong_2nd(w[org[1]], org[2], new[b], nd, ng)
        be.regrid(plan, forest)
        got = be.get_state(0)
        assert np.array_equal(forest.interior(got), forest.interior(new))
        checked += 1
    assert checked > 0
