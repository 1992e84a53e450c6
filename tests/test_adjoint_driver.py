"""The checkpoint-and-recompute driver of the differentiable runs
(amrvac_amd.adjoint.SSPRK3Run) on the CPU, with a backend that applies a
state-dependent map through the same stage / adjoint1 entry points:

  stage:    u_out = init + qdt * K(u_in),   K(u) = S u + c u^2  (elementwise c),
  adjoint1: lam_in = coef*lam_out + ca*add_a + cb*add_b + qdt * (S^T + 2 c u_in) lam_out,

with S a fixed periodic stencil along x.  The checkpointed run's gradient must equal
torch.autograd through the same ssprk3 steps, for any checkpoint spacing:
the RK reverse, the stage tape, the segment recompute and the forward's last
segment on the tape are all exercised (a wrong stage state changes the
2 c u term)."""
from __future__ import annotations

import ctypes
import math

import numpy as np
import pytest
import torch

from amrvac_amd import adjoint
from amrvac_amd.host import Par, ssprk3_coefficients
from amrvac_amd.mesh import UniformMesh

W = (0.3, -0.5, 0.2)  # S u = 0.3 u(i-1) - 0.5 u(i) + 0.2 u(i+1) along x (last axis)
C = 0.05


def _view(ptr: int, shape):
    n = int(np.prod(shape))
    return np.ctypeslib.as_array((ctypes.c_double * n).from_address(ptr)).reshape(shape)


def _S(u):
    return W[0] * np.roll(u, 1, axis=-1) + W[1] * u + W[2] * np.roll(u, -1, axis=-1)


def _ST(u):
    return W[0] * np.roll(u, -1, axis=-1) + W[1] * u + W[2] * np.roll(u, 1, axis=-1)


class FakeBackend:
    def __init__(self, shape):
        self.shape = shape
        self.calls = {"stage": 0, "adjoint1": 0}

    def set_stream(self, s):
        pass

    def getdt_state(self, u, courantpar):
        return 0.1 / (1.0 + float(np.abs(_view(u, self.shape)).max())), 0.0

    def stage(self, method, dtfactor, dt, cmax, u_in, u_out, base=0, comb_a=0.0, comb_b=0.0):
        self.calls["stage"] += 1
        qdt = dtfactor * dt
        x = _view(u_in, self.shape).copy()
        init = comb_a * _view(base, self.shape) + comb_b * x if base else x
        _view(u_out, self.shape)[...] = init + qdt * (_S(x) + C * x * x)

    def adjoint1(self, method, dtfactor, dt, u_in, lam_out, coef, lam_in, add_a=0, ca=0.0, add_b=0, cb=0.0,
                 ch=0.0):
        self.calls["adjoint1"] += 1
        qdt = dtfactor * dt
        x = _view(u_in, self.shape)
        lo = _view(lam_out, self.shape).copy()
        r = coef * lo + qdt * (_ST(lo) + 2.0 * C * x * lo)
        if add_a:
            r = r + ca * _view(add_a, self.shape)
        if add_b:
            r = r + cb * _view(add_b, self.shape)
        _view(lam_in, self.shape)[...] = r


def _torch_run(u0, dts):
    c = ssprk3_coefficients()
    S = lambda u: W[0] * torch.roll(u, 1, -1) + W[1] * u + W[2] * torch.roll(u, -1, -1)
    L = lambda u, q: q * (S(u) + C * u * u)
    u = u0
    for dt in dts:
        u1 = u + L(u, c["beta11"] * dt)
        u2 = c["alfa21"] * u + c["alfa22"] * u1 + L(u1, c["beta22"] * dt)
        u = c["alfa31"] * u + c["alfa33"] * u2 + L(u2, c["beta33"] * dt)
    return u


@pytest.mark.parametrize("nsteps,every", [(1, 1), (7, 3), (12, 4), (10, 20), (9, 1)])
def test_driver_gradient_equals_autograd(nsteps, every):
    mesh = UniformMesh(2, (32, 16), (16, 16), (0.0, 0.0), (1.0, 0.5), ("periodic",) * 4)
    par = Par(physics="hd", ndim=2, ndir=2, time_stepper="threestep", flux_scheme="hllc", limiter="vanleer")
    shape = (mesh.nblocks, par.nw) + tuple(mesh.block_nx[::-1])
    be = FakeBackend(shape)
    run = adjoint.SSPRK3Run(par, mesh, nsteps, every=every, backend=be)
    gen = torch.Generator().manual_seed(nsteps * 100 + every)
    u0 = 1.0 + 0.1 * torch.rand(shape, generator=gen, dtype=torch.float64)
    w = torch.randn(shape, generator=gen, dtype=torch.float64)
    x = u0.clone().requires_grad_(True)
    out = adjoint.run_differentiable(x, run)
    g, = torch.autograd.grad((out * w).sum(), x)
    xr = u0.clone().requires_grad_(True)
    ref = _torch_run(xr, run.dts)
    assert torch.allclose(out.detach(), ref.detach(), rtol=1e-14, atol=1e-14)
    gr, = torch.autograd.grad((ref * w).sum(), xr)
    assert torch.allclose(g, gr, rtol=1e-12, atol=1e-13)
    # forward: 3 stages per step; backward: 3 adjoint stages per step, and
    # the recompute of every segment but the last (without its last stage)
    nseg = (nsteps + every - 1) // every
    last = nsteps - (nseg - 1) * every
    assert be.calls["adjoint1"] == 3 * nsteps
    assert be.calls["stage"] == 3 * nsteps + 3 * (nsteps - last) - (nseg - 1)


def test_whole_run_tape_when_it_fits():
    """The default schedule keeps every stage state when the tape fits in
    tape_bytes: no recompute, and the gradient bit for bit that of the
    checkpointed schedule (the recompute replays the same stages)."""
    mesh = UniformMesh(2, (32, 16), (16, 16), (0.0, 0.0), (1.0, 0.5), ("periodic",) * 4)
    par = Par(physics="hd", ndim=2, ndir=2, time_stepper="threestep", flux_scheme="hllc", limiter="vanleer")
    shape = (mesh.nblocks, par.nw) + tuple(mesh.block_nx[::-1])
    nsteps = 9
    tape = (nsteps + 1) * 3 * math.prod(shape) * 8
    assert adjoint.SSPRK3Run(par, mesh, nsteps, backend=FakeBackend(shape), tape_bytes=tape - 1).every == 3
    be = FakeBackend(shape)
    run = adjoint.SSPRK3Run(par, mesh, nsteps, backend=be, tape_bytes=tape)
    assert run.every == nsteps
    gen = torch.Generator().manual_seed(5)
    u0 = 1.0 + 0.1 * torch.rand(shape, generator=gen, dtype=torch.float64)
    w = torch.randn(shape, generator=gen, dtype=torch.float64)
    x = u0.clone().requires_grad_(True)
    g, = torch.autograd.grad((adjoint.run_differentiable(x, run) * w).sum(), x)
    assert be.calls["stage"] == 3 * nsteps  # the forward's stages only
    run3 = adjoint.SSPRK3Run(par, mesh, nsteps, every=3, backend=FakeBackend(shape))
    x3 = u0.clone().requires_grad_(True)
    g3, = torch.autograd.grad((adjoint.run_differentiable(x3, run3) * w).sum(), x3)
    assert torch.equal(g, g3)


def test_driver_replays_given_dts():
    mesh = UniformMesh(2, (16, 16), (16, 16), (0.0, 0.0), (1.0, 1.0), ("periodic",) * 4)
    par = Par(physics="hd", ndim=2, ndir=2, time_stepper="threestep", flux_scheme="hllc", limiter="vanleer")
    shape = (mesh.nblocks, par.nw) + tuple(mesh.block_nx[::-1])
    run = adjoint.SSPRK3Run(par, mesh, 4, every=2, backend=FakeBackend(shape))
    u0 = torch.ones(shape, dtype=torch.float64)
    dts = [0.01, 0.02, 0.03, 0.04]
    out = run.forward(u0, dts)
    assert run.dts == dts
    assert torch.allclose(out, _torch_run(u0, dts), rtol=1e-14, atol=1e-14)
    with pytest.raises(ValueError):
        run.forward(u0, dts[:3])
