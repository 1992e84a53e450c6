"""Adjoint kernels on the device: gradients of a loss on the final density
w.r.t. the initial rho, vx, vy of a periodic Kelvin-Helmholtz run (threestep
/ HLLC / vanleer), against the FP64 PyTorch-autograd restatement at 32x32
and against FP64 central finite differences at 128x128 / 200 steps."""
from __future__ import annotations

import time

import numpy as np
import pytest
import torch

from amrvac_amd import adjoint
from amrvac_amd.host import Par
from amrvac_amd.mesh import UniformMesh

import torch_fv_ref

pytestmark = pytest.mark.gpu

GAMMA = 1.4
# seeded density noise: keeps every limiter argument off its kink at t=0
# (with flat density bands the run has one-sided derivatives only there)
RHO_AMP = 1.0e-3


def _setup(n, riemann="hllc", lim="vanleer"):
    mesh = UniformMesh(2, (n, n), (16, 16), (0.0, 0.0), (1.0, 1.0), ("periodic",) * 4)
    par = Par(physics="hd", ndim=2, ndir=2, gamma=GAMMA, time_stepper="threestep", flux_scheme=riemann,
              limiter=lim, courantpar=0.8)
    return mesh, par


class _Problem:
    """loss(rho0, vx0, vy0) = sum((rho_final - rho_target)^2) / ncells"""

    def __init__(self, n, nsteps, every, precision="fp64", riemann="hllc", lim="vanleer", rho_amp=RHO_AMP):
        self.mesh, self.par = _setup(n, riemann, lim)
        self.dev = torch.device("cuda")
        self.idx = torch.as_tensor(adjoint.block_index(self.mesh), device=self.dev)
        self.run = adjoint.SSPRK3Run(self.par, self.mesh, nsteps, every=every, precision=precision)
        self.shape = tuple(self.mesh.domain_nx[::-1])
        rho, vx, vy, p = adjoint.kh_fields(self.mesh, 20261202, device=self.dev, rho_amp=rho_amp)
        self.x0 = (rho, vx, vy)
        self.p = p
        rt, vxt, vyt, pt = adjoint.kh_fields(self.mesh, 20261203, device=self.dev, rho_amp=rho_amp)
        ut = adjoint.to_blocks(adjoint.hd_conserved(rt, [vxt, vyt], pt, GAMMA), self.idx)
        with torch.no_grad():
            self.target = adjoint.from_blocks(self.run.forward(ut), self.idx, self.shape)[0].clone()
        self.dts = None

    def loss(self, rho, vx, vy, dts=None):
        u0 = adjoint.to_blocks(adjoint.hd_conserved(rho, [vx, vy], self.p, GAMMA), self.idx)
        uf = adjoint.run_differentiable(u0, self.run, dts)
        r = adjoint.from_blocks(uf, self.idx, self.shape)[0]
        return ((r - self.target) ** 2).sum() / r.numel()

    def grad(self):
        xs = [x.clone().requires_grad_(True) for x in self.x0]
        L = self.loss(*xs)
        self.dts = list(self.run.dts)
        gs = torch.autograd.grad(L, xs)
        return L.item(), [g.detach() for g in gs]


def test_adjoint_matches_torch_autograd_32():
    print("[adjoint] 32x32 vs torch autograd", flush=True)
    pb = _Problem(32, 20, 5)
    L, g_dev = pb.grad()
    # the same scheme and dt sequence in FP64 torch on the CPU
    xs = [x.detach().cpu().clone().requires_grad_(True) for x in pb.x0]
    u0 = adjoint.hd_conserved(xs[0], [xs[1], xs[2]], pb.p.cpu(), GAMMA)
    uf = torch_fv_ref.run(u0, pb.dts, pb.mesh.dx, GAMMA, "vanleer", "hllc")
    Lr = ((uf[0] - pb.target.cpu()) ** 2).sum() / uf[0].numel()
    g_ref = torch.autograd.grad(Lr, xs)
    assert abs(Lr.item() - L) <= 1e-12 * abs(L)
    for a, b in zip(g_dev, g_ref):
        rel = (a.cpu() - b).norm() / b.norm()
        print(f"[adjoint] rel L2 grad difference {rel:.3e}", flush=True)
        assert rel < 1e-9


@pytest.mark.parametrize("riemann,lim", [("hll", "minmod"), ("tvdlf", "vanleer")])
def test_adjoint_other_schemes_match_torch_autograd(riemann, lim):
    print(f"[adjoint] 32x32 {riemann}/{lim} vs torch autograd", flush=True)
    pb = _Problem(32, 10, 5, riemann=riemann, lim=lim)
    L, g_dev = pb.grad()
    xs = [x.detach().cpu().clone().requires_grad_(True) for x in pb.x0]
    u0 = adjoint.hd_conserved(xs[0], [xs[1], xs[2]], pb.p.cpu(), GAMMA)
    uf = torch_fv_ref.run(u0, pb.dts, pb.mesh.dx, GAMMA, lim, riemann)
    Lr = ((uf[0] - pb.target.cpu()) ** 2).sum() / uf[0].numel()
    g_ref = torch.autograd.grad(Lr, xs)
    for a, b in zip(g_dev, g_ref):
        assert (a.cpu() - b).norm() / b.norm() < 1e-9


def _fd_directions(pb, g, ndir=16, eps=1.0e-6):
    """(fd, adjoint) directional derivatives along ``ndir`` seeded unit
    directions of (rho, vx, vy), central differences of the FP64 loss."""
    gen = torch.Generator(device="cpu").manual_seed(20261202)
    fd, an = [], []
    with torch.no_grad():
        for _ in range(ndir):
            d = [torch.randn(pb.shape, generator=gen, dtype=torch.float64).to(pb.dev) for _ in range(3)]
            nrm = torch.sqrt(sum((v * v).sum() for v in d))
            d = [v / nrm for v in d]
            lp = pb.loss(*[x + eps * v for x, v in zip(pb.x0, d)], dts=pb.dts).item()
            lm = pb.loss(*[x - eps * v for x, v in zip(pb.x0, d)], dts=pb.dts).item()
            fd.append((lp - lm) / (2.0 * eps))
            an.append(sum((gi * v).sum() for gi, v in zip(g, d)).item())
    return np.array(fd), np.array(an)


def test_adjoint_finite_differences_kh_128():
    """128^2, 200 steps, checkpoints every 20; FP64 central differences
    (eps 1e-6) along 16 seeded unit directions.  The run is only piecewise
    smooth (limiter and HLLC branches), so a difference step crosses branch
    switches somewhere in the 600 stages; that error grows linearly with
    eps.  The criterion is normwise over the directions:
    ||fd - adjoint||_2 / ||fd||_2 <= 1e-4."""
    print("[adjoint] 128x128 200 steps vs finite differences", flush=True)
    pb = _Problem(128, 200, 20)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    L, g = pb.grad()
    torch.cuda.synchronize()
    t_grad = time.perf_counter() - t0
    t0 = time.perf_counter()
    with torch.no_grad():
        pb.loss(*pb.x0)
    torch.cuda.synchronize()
    t_fwd = time.perf_counter() - t0
    print(f"[adjoint] forward {t_fwd:.3f} s, forward+backward {t_grad:.3f} s, "
          f"backward/forward {(t_grad - t_fwd) / t_fwd:.2f}", flush=True)
    fd, an = _fd_directions(pb, g)
    for k in range(len(fd)):
        print(f"[adjoint] direction {k}: fd {fd[k]:.10e} adjoint {an[k]:.10e} "
              f"rel {abs(fd[k] - an[k]) / abs(fd[k]):.2e}", flush=True)
    rel = np.linalg.norm(fd - an) / np.linalg.norm(fd)
    print(f"[adjoint] normwise relative difference {rel:.3e}", flush=True)
    assert rel < 1.0e-4


def test_adjoint_mixed_precision_close_to_fp64():
    """The mixed-precision gradient (FP32 faces, FP64 state) is the gradient
    of a slightly different discrete map: against the FP64 one it agrees to
    a few 1e-3 normwise, and against FP64 central differences likewise."""
    print("[adjoint] mixed vs fp64 gradients, 128x128 200 steps", flush=True)
    p64 = _Problem(128, 200, 20)
    g64 = p64.grad()[1]
    p32 = _Problem(128, 200, 20, precision="mixed")
    g32 = p32.grad()[1]
    # The mixed adjoint is the exact reverse of the FP32-face map, not of
    # the FP64 one: the two gradients differ by 3.7e-4 to 1.1e-3 in L2
    # (rho, vx, vy), and that difference, not the adjoint, bounds the
    # agreement with FP64 differences (measured 1.2e-3 normwise).  The FP64
    # adjoint meets the 1e-4 target against the same differences
    # (test_adjoint_finite_differences_kh_128).
    for a, b in zip(g32, g64):
        rel = ((a - b).norm() / b.norm()).item()
        print(f"[adjoint] mixed vs fp64 rel L2 {rel:.3e}", flush=True)
        assert rel < 3e-3
    p64.dts = p32.dts  # FP64 differences along the mixed run's dt sequence
    fd, an = _fd_directions(p64, g32, ndir=8)
    rel = np.linalg.norm(fd - an) / np.linalg.norm(fd)
    print(f"[adjoint] mixed adjoint vs FP64 differences, normwise {rel:.3e}", flush=True)
    assert rel < 3e-3
