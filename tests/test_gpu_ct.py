"""Constrained transport on the device (staggered face field, 2.5-D field
loop, uniform periodic level) against the CPU oracle: the contraction-free
FP64 kernels reproduce the oracle bit for bit (cells, face field, dt
sequence) for every edge-field average and Riemann solver; the FMA build
reproduces the reference's regression logs."""
from __future__ import annotations

import dataclasses
import os

import numpy as np
import pytest

from amrvac_amd import problems
from amrvac_amd.host import TimeIntegrator, logs_match, read_log
from oracle.oracle import OracleFV

from test_oracle_golden import GOLD, FieldLoopIntegrator, field_loop_setup

pytestmark = pytest.mark.gpu

CASES = [
    # (scheme file, flux scheme override, limiter override, time stepper override)
    ("3step_hll_cada_ct", None, None, None),
    ("3step_hll_cada_ct_average", None, None, None),
    ("3step_hll_cada_ct_hll", None, None, None),
    ("3step_hlld_cada_ct", None, None, None),
    ("3step_hll_cada_ct", "tvdlf", "minmod", None),
    ("3step_hll_cada_ct_hll", "tvdlf", "vanleer", None),
    ("3step_hll_cada_ct_average", "hllc", "koren", None),
    ("3step_hll_cada_ct_hll", "hlld", "woodward", "fourstep"),
    ("3step_hll_cada_ct", "hll", "mcbeta", "fourstep:rk4"),
    ("3step_hll_cada_ct_average", "tvdlf", "superbee", "fivestep"),
    ("3step_hll_cada_ct", "hll", "vanleer", "twostep"),          # predictor-corrector, HLL predictor
    ("3step_hlld_cada_ct", "hlld", "minmod", "onestep"),
    ("3step_hll_cada_ct_average", "tvdlf", "minmod", "twostep"),  # Hancock predictor
]


def _par(scheme, fs, lim, ts):
    par, mesh = field_loop_setup(scheme)
    kw = {}
    if fs:
        kw["flux_scheme"] = fs
    if lim:
        kw["limiter"] = lim
    if ts:
        kw["time_stepper"] = ts.split(":")[0]
        if ":" in ts:
            kw["time_integrator"] = ts.split(":")[1]
    return dataclasses.replace(par, **kw), mesh


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(x for x in c if x))
def test_ct_exact_kernels_match_oracle_bitwise(case):
    from amrvac_amd.fv import DeviceFV
    print(f"[ct] {case}", flush=True)
    par, mesh = _par(*case)
    w0 = problems.field_loop_ct(mesh)
    runs = {}
    for name, mk in (("oracle", lambda: OracleFV(par, mesh, threads=4)),
                     ("device", lambda: DeviceFV(par, mesh, precision="fp64_exact"))):
        be = mk()
        be.set_state(0, w0)
        be.getbc(0)
        ti = TimeIntegrator(par, mesh, be)
        ti.run(nsteps=8)
        runs[name] = (be.get_state(0), list(ti.state.dts))
    (wo, dto), (wd, dtd) = runs["oracle"], runs["device"]
    assert dto == dtd
    assert np.array_equal(mesh.interior(wd), mesh.interior(wo))


@pytest.mark.parametrize("scheme", ["3step_hll_cada_ct", "3step_hll_cada_ct_average", "3step_hll_cada_ct_hll",
                                    "3step_hlld_cada_ct"])
def test_ct_device_reproduces_reference_log(scheme):
    from amrvac_amd.fv import DeviceFV
    print(f"[ct] golden log {scheme}", flush=True)
    par, mesh = field_loop_setup(scheme)
    be = DeviceFV(par, mesh, precision="fp64")
    be.set_state(0, problems.field_loop_ct(mesh))
    be.getbc(0)
    log = np.array(FieldLoopIntegrator(par, mesh, be).run().log)
    gold = read_log(os.path.join(GOLD, f"fl_2d_{scheme}.log"))
    assert log.shape == gold.shape
    assert logs_match(log, gold)


@pytest.mark.parametrize("scheme", ["3step_hll_cada_ct", "3step_hll_cada_ct_hll", "3step_hlld_cada_ct"])
def test_ct_mixed_form(scheme):
    """The mixed form (FP32 reconstruction and Riemann solver, FP64 state,
    face field and edge fields): per-variable L1 within 1e-5 of FP64 after
    20 steps, and the face-field divergence stays at round-off."""
    from amrvac_amd.fv import DeviceFV
    print(f"[ct] mixed {scheme}", flush=True)
    par, mesh = field_loop_setup(scheme)
    out = {}
    for prec in ("fp64", "mixed"):
        be = DeviceFV(par, mesh, precision=prec)
        be.set_state(0, problems.field_loop_ct(mesh))
        be.getbc(0)
        TimeIntegrator(par, mesh, be).run(nsteps=20)
        out[prec] = be.get_state(0)
    a, b = mesh.interior(out["fp64"]), mesh.interior(out["mixed"])
    for iw in (0, 1, 2, 4, 5, 6):
        l1 = np.abs(b[:, iw] - a[:, iw]).sum() / np.abs(a[:, iw]).sum()
        assert l1 <= 1e-5, (iw, l1)
    ex = problems.field_loop_output_vars_ct(mesh, out["mixed"])
    divb = np.abs(mesh.interior(ex)[:, 0]).max()
    assert divb <= 1e-12 * np.abs(b[:, 5:7]).max() / min(mesh.dx)
