"""The PyTorch restatement used as the autograd reference of the adjoint
kernels must reproduce the CPU oracle's forward run."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from amrvac_amd import adjoint
from amrvac_amd.host import Par, TimeIntegrator
from amrvac_amd.mesh import UniformMesh
from oracle.oracle import OracleFV

import torch_fv_ref


@pytest.mark.parametrize("riemann,lim", [("hllc", "vanleer"), ("hll", "minmod"), ("tvdlf", "vanleer")])
def test_torch_reference_matches_oracle(riemann, lim):
    mesh = UniformMesh(2, (32, 32), (16, 16), (0.0, 0.0), (1.0, 1.0), ("periodic",) * 4)
    par = Par(physics="hd", ndim=2, ndir=2, gamma=1.4, time_stepper="threestep", flux_scheme=riemann,
              limiter=lim, courantpar=0.8)
    rho, vx, vy, p = adjoint.kh_fields(mesh, 20261202)
    u0 = adjoint.hd_conserved(rho, [vx, vy], p, par.gamma)
    idx = torch.as_tensor(adjoint.block_index(mesh))
    blocks = adjoint.to_blocks(u0, idx).numpy()
    w = np.zeros((mesh.nblocks, 4) + mesh.nxg[::-1])
    mesh.interior(w)[:] = blocks
    be = OracleFV(par, mesh, threads=2)
    be.set_state(0, w)
    be.getbc(0)
    ti = TimeIntegrator(par, mesh, be)
    ti.run(nsteps=8)
    got_orc = torch.as_tensor(mesh.to_global(be.get_state(0)))
    got_t = torch_fv_ref.run(u0, ti.state.dts, mesh.dx, par.gamma, lim, riemann)
    assert torch.allclose(got_t, got_orc, rtol=1e-12, atol=1e-13)


@pytest.mark.parametrize("physics,riemann,lim", [("hd", "hllc", "vanleer"), ("mhd", "hll", "vanleer"),
                                                 ("mhd", "tvdlf", "minmod")])
def test_torch_reference_3d_matches_oracle(physics, riemann, lim):
    mesh = UniformMesh(3, (16, 16, 32), (16, 16, 16), (0.0,) * 3, (1.0,) * 3, ("periodic",) * 6)
    if physics == "hd":
        par = Par(physics="hd", ndim=3, ndir=3, gamma=1.4, time_stepper="threestep", flux_scheme=riemann,
                  limiter=lim, courantpar=0.8)
        rho, v, p, b = adjoint.orszag_tang_fields(mesh, 7)
        u0 = adjoint.hd_conserved(rho, v, p, par.gamma)
    else:
        par = Par(physics="mhd", ndim=3, ndir=3, gamma=5.0 / 3.0, typedivbfix="none", time_stepper="threestep",
                  flux_scheme=riemann, limiter=lim, courantpar=0.6)
        rho, v, p, b = adjoint.orszag_tang_fields(mesh, 7)
        u0 = adjoint.mhd_conserved(rho, v, p, b, par.gamma)
    idx = torch.as_tensor(adjoint.block_index(mesh))
    w = np.zeros((mesh.nblocks, par.nw) + mesh.nxg[::-1])
    mesh.interior(w)[:] = adjoint.to_blocks(u0, idx).numpy()
    be = OracleFV(par, mesh, threads=2)
    be.set_state(0, w)
    be.getbc(0)
    ti = TimeIntegrator(par, mesh, be)
    ti.run(nsteps=3)
    got_orc = torch.as_tensor(mesh.to_global(be.get_state(0)))
    sch = torch_fv_ref.Scheme(3, physics, par.gamma, lim, riemann)
    got_t = sch.run(u0, ti.state.dts, mesh.dx)
    assert torch.allclose(got_t, got_orc, rtol=1e-12, atol=1e-13)


def test_torch_reference_glm_cada3_matches_oracle():
    """3-D MHD with GLM (interface states, psi / B_n fluxes, decay and
    extended sources) and the cada3 limiter: the benchmark's scheme."""
    mesh = UniformMesh(3, (16, 16, 32), (16, 16, 16), (0.0,) * 3, (1.0,) * 3, ("periodic",) * 6)
    par = Par(physics="mhd", ndim=3, ndir=3, gamma=5.0 / 3.0, typedivbfix="glm", time_stepper="threestep",
              flux_scheme="hll", limiter="cada3", courantpar=0.6)
    rho, v, p, b = adjoint.orszag_tang_fields(mesh, 7)
    gen = torch.Generator().manual_seed(11)
    psi = 1.0e-2 * torch.randn(rho.shape, generator=gen, dtype=torch.float64)
    u0 = adjoint.mhd_conserved(rho, v, p, b, par.gamma, psi=psi)
    idx = torch.as_tensor(adjoint.block_index(mesh))
    w = np.zeros((mesh.nblocks, par.nw) + mesh.nxg[::-1])
    mesh.interior(w)[:] = adjoint.to_blocks(u0, idx).numpy()
    be = OracleFV(par, mesh, threads=2)
    be.set_state(0, w)
    be.getbc(0)
    ti = TimeIntegrator(par, mesh, be)
    ti.run(nsteps=3)
    got_orc = torch.as_tensor(mesh.to_global(be.get_state(0)))
    sch = torch_fv_ref.Scheme(3, "mhd", par.gamma, "cada3", "hll", glm=True)
    got_t = sch.run(u0, ti.state.dts, mesh.dx, ti.state.chs)
    assert ti.state.chs[0] > 0.0
    assert torch.allclose(got_t, got_orc, rtol=1e-12, atol=1e-13)


def test_cada3_gradient_is_ill_conditioned():
    """The gradient of a cada3 stage moves by far more than rounding under a
    1e-14 relative change of the state (theta = x / (d0 + sign(eps, d0))
    with eps = 1.5e-8 has curvature of order 1/eps^2 where the differences
    are small), a vanleer stage's by about the change itself: the bound on
    how closely two implementations' cada3 gradients can agree."""
    mesh = UniformMesh(3, (16, 16, 32), (16, 16, 16), (0.0,) * 3, (1.0,) * 3, ("periodic",) * 6)
    rho, v, p, b = adjoint.orszag_tang_fields(mesh, 7)
    gen = torch.Generator().manual_seed(3)
    noise = lambda q: q + 1.0e-3 * torch.randn(q.shape, generator=gen, dtype=torch.float64)
    v, b, p = [noise(c) for c in v], [noise(c) for c in b], noise(p)
    u = adjoint.mhd_conserved(rho, v, p, b, 5.0 / 3.0)
    lam = torch.randn(u.shape, generator=gen, dtype=torch.float64)
    pert = 1.0e-14 * torch.randn(u.shape, generator=gen, dtype=torch.float64)
    change = {}
    for lim in ("vanleer", "cada3"):
        sch = torch_fv_ref.Scheme(3, "mhd", 5.0 / 3.0, lim, "hll")
        gs = []
        for uu in (u, u * (1.0 + pert)):
            x = uu.clone().requires_grad_(True)
            o = sch.stage(x, torch.zeros_like(x), 2.0e-3, mesh.dx)
            g, = torch.autograd.grad((o * lam).sum(), x)
            gs.append(g)
        change[lim] = ((gs[0] - gs[1]).norm() / gs[0].norm()).item()
    assert change["vanleer"] < 1e-10
    assert change["cada3"] > 1e-8
