"""Constrained transport in 3-D on the device (staggered face field on a
uniform periodic level, 16^3 blocks) against the CPU oracle, whose 3-D form
is pinned in test_oracle_ct3d.py.

* The contraction-free FP64 kernels reproduce the oracle bit for bit (cells,
  face field, dt sequence) on a fully 3-D sheared, compressible state with a
  field from an edge vector potential: every edge-field type, every Riemann
  solver, the TVD limiters (compile-time and run-time switched), every kind
  of stepper, and a box with different cell sizes per direction.
* A level of one block, which is its own neighbour in every direction.
* The FMA build within 1e-12 of the exact one; the mixed form (FP32
  reconstruction and Riemann solver, FP64 state, faces and edge fields)
  within 1e-5 relative L1 per variable, with div B at round-off in both."""
from __future__ import annotations

import dataclasses

import numpy as np
import pytest

from amrvac_amd.host import TimeIntegrator
from amrvac_amd.mesh import UniformMesh
from oracle.oracle import OracleFV

from test_oracle_ct3d import ct_par, div_faces, from_global, sheared_state

pytestmark = pytest.mark.gpu

CASES = [
    # (type_ct, flux scheme, limiter, time stepper[:integrator])
    ("uct_contact", "hll", "cada3", "threestep"),
    ("average", "hll", "cada3", "threestep"),
    ("uct_hll", "hll", "cada3", "threestep"),
    ("uct_contact", "hlld", "cada3", "threestep"),
    ("uct_hll", "tvdlf", "vanleer", "threestep"),
    ("average", "hllc", "koren", "threestep"),
    ("uct_contact", "tvdlf", "minmod", "fourstep:rk4"),
    ("uct_hll", "hlld", "woodward", "fourstep"),
    ("average", "hll", "vanleer", "twostep"),       # predictor-corrector, HLL predictor
    ("uct_contact", "tvdlf", "minmod", "twostep"),  # Hancock predictor
    ("uct_hll", "hll", "minmod", "onestep"),
]


def _par(ct, fs, lim, ts):
    kw = dict(flux_scheme=fs, limiter=lim, time_stepper=ts.split(":")[0])
    if ":" in ts:
        kw["time_integrator"] = ts.split(":")[1]
    return dataclasses.replace(ct_par(3, ct), **kw)


def _runs(par, mesh, w0, backends, nsteps):
    out = {}
    for name, mk in backends:
        be = mk()
        be.set_state(0, w0)
        be.getbc(0)
        ti = TimeIntegrator(par, mesh, be)
        ti.run(nsteps=nsteps)
        out[name] = (be.get_state(0), list(ti.state.dts))
    return out


def _exact_vs_oracle(par, mesh, g, nsteps):
    from amrvac_amd.fv import DeviceFV
    w0 = from_global(mesh, g)
    runs = _runs(par, mesh, w0, (("oracle", lambda: OracleFV(par, mesh, threads=8)),
                                 ("device", lambda: DeviceFV(par, mesh, precision="fp64_exact"))), nsteps)
    (wo, dto), (wd, dtd) = runs["oracle"], runs["device"]
    assert dto == dtd
    assert np.array_equal(mesh.interior(wd), mesh.interior(wo))


@pytest.mark.parametrize("case", CASES, ids=lambda c: "-".join(c).replace(":", "_"))
def test_ct3d_exact_kernels_match_oracle_bitwise(case):
    print(f"[ct3d] {case}", flush=True)
    par = _par(*case)
    mesh = UniformMesh(3, (32, 32, 32), (16, 16, 16), (0.0,) * 3, (1.0,) * 3, ("periodic",) * 6)
    _exact_vs_oracle(par, mesh, sheared_state(32, 0.3), nsteps=4)


@pytest.mark.parametrize("ct", ["uct_contact", "average", "uct_hll"])
def test_ct3d_exact_kernels_with_unequal_cell_sizes(ct):
    print(f"[ct3d] unequal cells {ct}", flush=True)
    n, lengths = (32, 16, 48), (1.0, 0.75, 2.0)
    mesh = UniformMesh(3, n, (16, 16, 16), (0.0,) * 3, lengths, ("periodic",) * 6)
    _exact_vs_oracle(ct_par(3, ct), mesh, sheared_state(n, 0.3, lengths), nsteps=4)


@pytest.mark.parametrize("ct", ["uct_contact", "uct_hll"])
def test_ct3d_single_block_level(ct):
    """One periodic block: its own neighbour in every direction, so its
    lower faces are its own upper faces."""
    print(f"[ct3d] single block {ct}", flush=True)
    mesh = UniformMesh(3, (16, 16, 16), (16, 16, 16), (0.0,) * 3, (1.0,) * 3, ("periodic",) * 6)
    _exact_vs_oracle(ct_par(3, ct), mesh, sheared_state(16, 0.3), nsteps=3)


@pytest.mark.parametrize("ct", ["uct_contact", "uct_hll"])
def test_ct3d_fma_and_mixed_forms(ct):
    from amrvac_amd.fv import DeviceFV
    print(f"[ct3d] fma / mixed {ct}", flush=True)
    par = ct_par(3, ct)
    mesh = UniformMesh(3, (32, 32, 32), (16, 16, 16), (0.0,) * 3, (1.0,) * 3, ("periodic",) * 6)
    w0 = from_global(mesh, sheared_state(32, 0.3))
    runs = _runs(par, mesh, w0, [(p, (lambda p=p: DeviceFV(par, mesh, precision=p)))
                                 for p in ("fp64_exact", "fp64", "mixed")], nsteps=10)
    ex = mesh.to_global(runs["fp64_exact"][0])
    h = mesh.dx
    for prec, bound in (("fp64", 1e-12), ("mixed", 1e-5)):
        g = mesh.to_global(runs[prec][0])
        for iw in range(11):
            l1 = np.abs(g[iw] - ex[iw]).sum() / np.abs(ex[iw]).sum()
            assert l1 <= bound, (prec, iw, l1)
        bmax = np.abs(g[8:]).max()
        assert np.abs(div_faces(g[8:], h)).max() * min(h) / bmax < 1e-13, prec


def test_ct3d_field_loop_exact_kernels_match_oracle_bitwise():
    """The Gardiner & Stone 3-D field loop (the bench workload, 32 x 32 x 64)."""
    from amrvac_amd import problems
    from amrvac_amd.fv import DeviceFV
    print("[ct3d] field loop", flush=True)
    par = ct_par(3, "uct_contact")
    mesh = UniformMesh(3, (32, 32, 64), (16, 16, 16), (-0.5, -0.5, -1.0), (0.5, 0.5, 1.0), ("periodic",) * 6)
    w0 = problems.field_loop_3d_ct(mesh)
    runs = _runs(par, mesh, w0, (("oracle", lambda: OracleFV(par, mesh, threads=8)),
                                 ("device", lambda: DeviceFV(par, mesh, precision="fp64_exact"))), 5)
    (wo, dto), (wd, dtd) = runs["oracle"], runs["device"]
    assert dto == dtd
    assert np.array_equal(mesh.interior(wd), mesh.interior(wo))


def test_ct3d_exact_kernels_64_cubed_10_steps():
    """A larger level (64^3, 64 blocks) and a longer run (10 steps): uct_hll
    with HLLD, whose normal field comes from ps, bit for bit."""
    print("[ct3d] 64^3, 10 steps, uct_hll + hlld", flush=True)
    par = dataclasses.replace(ct_par(3, "uct_hll"), flux_scheme="hlld")
    mesh = UniformMesh(3, (64, 64, 64), (16, 16, 16), (0.0,) * 3, (1.0,) * 3, ("periodic",) * 6)
    _exact_vs_oracle(par, mesh, sheared_state(64, 0.3), nsteps=10)
