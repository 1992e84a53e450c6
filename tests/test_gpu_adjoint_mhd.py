"""Adjoint of 3-D ideal MHD on the device (threestep / HLL / vanleer,
Einfeldt bounds, no divergence cleaning, periodic): gradients of the final
magnetic energy w.r.t. the Orszag-Tang field amplitude b0 and the initial
density field, against FP64 central finite differences."""
from __future__ import annotations

import time

import pytest
import torch

import torch_fv_ref

from amrvac_amd import adjoint
from amrvac_amd.host import Par
from amrvac_amd.mesh import UniformMesh

pytestmark = pytest.mark.gpu

GAMMA = 5.0 / 3.0
SEED = 20261202
# seeded noise on v, p and B: the Orszag-Tang fields are constant along
# some axes, where every limiter slope sits on its kink (the run then has
# one-sided derivatives only); the noise moves the evaluation point off them
FIELD_AMP = 1.0e-3


class _MHDProblem:
    """loss(b0, rho0) = sum over cells of |B_final|^2 / 2 * dV"""

    def __init__(self, n, nsteps, precision="fp64", every=None, field_amp=FIELD_AMP, divb="none", lim="vanleer"):
        self.mesh = UniformMesh(3, (n, n, n), (16, 16, 16), (0.0,) * 3, (1.0,) * 3, ("periodic",) * 6)
        self.par = Par(physics="mhd", ndim=3, ndir=3, gamma=GAMMA, typedivbfix=divb, time_stepper="threestep",
                       flux_scheme="hll", limiter=lim, courantpar=0.6)
        self.dev = torch.device("cuda")
        self.idx = torch.as_tensor(adjoint.block_index(self.mesh), device=self.dev)
        self.run = adjoint.SSPRK3Run(self.par, self.mesh, nsteps, every=every, precision=precision)
        self.rho, self.v, self.p, self.bshape = adjoint.orszag_tang_fields(self.mesh, SEED, device=self.dev,
                                                                           field_amp=field_amp)
        self.b0 = torch.tensor(1.0, dtype=torch.float64, device=self.dev)
        self.dV = 1.0 / n ** 3
        self.dts = None

    def loss(self, b0, rho, dts=None):
        psi = torch.zeros_like(rho) if self.par.typedivbfix == "glm" else None
        u0 = adjoint.to_blocks(adjoint.mhd_conserved(rho, self.v, self.p, [b0 * c for c in self.bshape], GAMMA,
                                                     psi=psi), self.idx)
        uf = adjoint.run_differentiable(u0, self.run, dts)
        bf = uf[:, 5:8]
        return 0.5 * (bf * bf).sum() * self.dV

    def grad(self):
        b0 = self.b0.clone().requires_grad_(True)
        rho = self.rho.clone().requires_grad_(True)
        L = self.loss(b0, rho)
        self.dts = list(self.run.dts)
        gb, gr = torch.autograd.grad(L, (b0, rho))
        return L.item(), gb.item(), gr.detach()


@pytest.mark.parametrize("physics,riemann,lim", [("mhd", "hll", "vanleer"), ("mhd", "tvdlf", "minmod"),
                                                 ("hd", "hllc", "vanleer"), ("mhd-glm", "hll", "cada3"),
                                                 ("mhd-glm", "tvdlf", "vanleer"), ("mhd-glm", "hll", "vanleer"),
                                                 ("mhd", "hll", "cada3")])
def test_adjoint_3d_matches_torch_autograd(physics, riemann, lim):
    """3-D gradients of sum(u_final * r) (r seeded) w.r.t. the initial
    conserved state against torch autograd through the FP64 PyTorch
    restatement, on the CPU, for the dt (and GLM c_h) sequence of the
    device run.  mhd-glm with hll / cada3 is the benchmark's scheme."""
    print(f"[adjoint-3d] {physics} {riemann}/{lim} 32^3 vs torch autograd", flush=True)
    mesh = UniformMesh(3, (32, 32, 32), (16, 16, 16), (0.0,) * 3, (1.0,) * 3, ("periodic",) * 6)
    glm = physics == "mhd-glm"
    if physics == "hd":
        par = Par(physics="hd", ndim=3, ndir=3, gamma=1.4, time_stepper="threestep", flux_scheme=riemann,
                  limiter=lim, courantpar=0.8)
    else:
        par = Par(physics="mhd", ndim=3, ndir=3, gamma=GAMMA, typedivbfix="glm" if glm else "none",
                  time_stepper="threestep", flux_scheme=riemann, limiter=lim, courantpar=0.6)
    dev = torch.device("cuda")
    idx = torch.as_tensor(adjoint.block_index(mesh), device=dev)
    rho, v, p, b = adjoint.orszag_tang_fields(mesh, SEED, device=dev)
    gen = torch.Generator(device="cpu").manual_seed(SEED)
    # a generic point: the Orszag-Tang fields are constant along some axes
    # (every slope there sits on a limiter kink), so all fields get noise
    noise = lambda q: q + 1.0e-3 * torch.randn(q.shape, generator=gen, dtype=torch.float64).to(dev)
    v, b, p = [noise(c) for c in v], [noise(c) for c in b], noise(p)
    psi = 1.0e-3 * torch.randn(rho.shape, generator=gen, dtype=torch.float64).to(dev) if glm else None
    u0 = (adjoint.hd_conserved(rho, v, p, par.gamma) if physics == "hd"
          else adjoint.mhd_conserved(rho, v, p, b, par.gamma, psi=psi))
    r = torch.randn(u0.shape, generator=gen, dtype=torch.float64)
    run = adjoint.SSPRK3Run(par, mesh, 4, every=2)
    x = u0.clone().requires_grad_(True)
    uf = adjoint.from_blocks(adjoint.run_differentiable(adjoint.to_blocks(x, idx), run), idx, (32, 32, 32))
    g_dev, = torch.autograd.grad((uf * r.to(dev)).sum(), x)
    sch = torch_fv_ref.Scheme(3, "hd" if physics == "hd" else "mhd", par.gamma, lim, riemann, glm=glm)
    xc = u0.detach().cpu().clone().requires_grad_(True)
    ufc = sch.run(xc, run.dts, mesh.dx, run.chs)
    assert torch.allclose(ufc.detach(), uf.detach().cpu(), rtol=1e-10, atol=1e-12)
    g_ref, = torch.autograd.grad((ufc * r).sum(), xc)
    rel = ((g_dev.cpu() - g_ref).norm() / g_ref.norm()).item()
    worst = ((g_dev.cpu() - g_ref).abs().max() / g_ref.abs().max()).item()
    print(f"[adjoint-3d] rel L2 {rel:.3e}, max {worst:.3e}", flush=True)
    # cada3's gradient is ill-conditioned (theta = x / (d0 + sign(eps, d0))
    # with eps = 1.5e-8): a 1e-14 relative change of the state moves the
    # stage gradient by 3e-7 (tests/test_torch_ref.py), and the device and
    # torch forward runs differ at rounding level, so over four steps the
    # two gradients agree to 4e-8 (GLM) / 6e-9 (none); at one stage, at the
    # same state, to 2e-10 (test_adjoint_stage_matches_torch_autograd)
    assert rel < (1e-7 if lim == "cada3" else 1e-9)


def _fd(pb: _MHDProblem, gb: float, gr: torch.Tensor, ndir: int = 3, eps_random: float = 1.0e-5):
    """Adjoint against FP64 central differences.  Along b0 (eps 1e-6).  For
    the density field, along the adjoint's own gradient direction (eps
    1e-6: the difference must reproduce |grad|), and along ``ndir`` seeded
    unit directions (normwise over the directions).  At 128^3 random unit
    directions move the loss by little more than the rounding noise of a
    100-step run at eps 1e-6, while larger steps cross limiter kinks (the
    map is only piecewise smooth); eps 1e-5 balances the two there.  On a
    32^3 grid the same eps moves each cell 8x more: 1e-6 is used."""
    out = {}
    with torch.no_grad():
        eps = 1.0e-6
        lp = pb.loss(pb.b0 + eps, pb.rho, dts=pb.dts).item()
        lm = pb.loss(pb.b0 - eps, pb.rho, dts=pb.dts).item()
        fdb = (lp - lm) / (2.0 * eps)
        out["b0"] = abs(fdb - gb) / abs(fdb)
        print(f"[adjoint-mhd] b0: fd {fdb:.10e} adjoint {gb:.10e} rel {out['b0']:.2e}", flush=True)
        d = gr / gr.norm()
        lp = pb.loss(pb.b0, pb.rho + eps * d, dts=pb.dts).item()
        lm = pb.loss(pb.b0, pb.rho - eps * d, dts=pb.dts).item()
        fdg = (lp - lm) / (2.0 * eps)
        ang = gr.norm().item()
        out["grad_dir"] = abs(fdg - ang) / abs(fdg)
        print(f"[adjoint-mhd] density, gradient direction: fd {fdg:.10e} |grad| {ang:.10e} "
              f"rel {out['grad_dir']:.2e}", flush=True)
        gen = torch.Generator(device="cpu").manual_seed(SEED)
        eps = eps_random
        fd, an = [], []
        for k in range(ndir):
            d = torch.randn(pb.rho.shape, generator=gen, dtype=torch.float64).to(pb.dev)
            d = d / d.norm()
            lp = pb.loss(pb.b0, pb.rho + eps * d, dts=pb.dts).item()
            lm = pb.loss(pb.b0, pb.rho - eps * d, dts=pb.dts).item()
            fd.append((lp - lm) / (2.0 * eps))
            an.append((gr * d).sum().item())
            print(f"[adjoint-mhd] density dir {k}: fd {fd[-1]:.10e} adjoint {an[-1]:.10e}", flush=True)
    fd, an = torch.tensor(fd), torch.tensor(an)
    out["random"] = ((fd - an).norm() / fd.norm()).item()
    print(f"[adjoint-mhd] density, random unit directions, normwise {out['random']:.2e}", flush=True)
    return out


def test_mhd_adjoint_fp64_finite_differences_32():
    print("[adjoint-mhd] 32^3, 10 steps, FP64", flush=True)
    pb = _MHDProblem(32, 10, "fp64", every=3)
    _, gb, gr = pb.grad()
    r = _fd(pb, gb, gr, eps_random=1.0e-6)
    # b0 and the gradient direction meet the 1e-4 target (measured 4e-9 and
    # 5e-6).  A random unit direction cannot: it is nearly orthogonal to the
    # gradient, so its directional derivative is about |grad|/sqrt(N) (5e-8
    # against |grad| 4e-5 here) and the absolute error of the difference
    # quotient, from the limiter kinks of the noisy 32^3 field crossed within
    # +-eps, is magnified by about sqrt(N) relative to it (measured 3.0e-3).
    # The adjoint itself equals the autograd twin of the same map to 1e-13
    # at this size (test_adjoint_3d_matches_torch_autograd).
    assert r["b0"] < 1.0e-6 and r["grad_dir"] < 1.0e-4 and r["random"] < 1.0e-2


def test_mhd_adjoint_config_128():
    """128^3, 100 steps, sqrt(N) checkpoints: FP64 adjoint against FP64
    central differences, and the mixed-precision adjoint (FP32 faces)
    against the same differences along its own dt sequence."""
    print("[adjoint-mhd] 128^3, 100 steps, FP64 and mixed", flush=True)
    pb = _MHDProblem(128, 100, "fp64")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _, gb, gr = pb.grad()
    torch.cuda.synchronize()
    print(f"[adjoint-mhd] fp64 forward+backward {time.perf_counter() - t0:.3f} s", flush=True)
    r = _fd(pb, gb, gr)
    # the 1e-4 target in every check (measured: b0 1.5e-10, gradient
    # direction 4.7e-8, random directions 7.6e-5 normwise; the runs are
    # deterministic)
    assert r["b0"] < 1.0e-6 and r["grad_dir"] < 1.0e-4 and r["random"] < 1.0e-4
    pm = _MHDProblem(128, 100, "mixed")
    _, gbm, grm = pm.grad()
    pb.dts = pm.dts
    rm = _fd(pb, gbm, grm)
    print(f"[adjoint-mhd] mixed vs fp64 density gradient rel L2 {((grm - gr).norm() / gr.norm()).item():.3e}",
          flush=True)
    # The mixed adjoint is the exact reverse of the FP32-face map, and the
    # differences are of the FP64 map: the two maps' gradients differ by
    # 1.5e-4 in L2 (printed above), which bounds how well the random
    # directions can agree (measured 3.8e-4); b0 and the gradient direction
    # still meet 1e-4 (4.8e-9, 7.2e-7).
    assert rm["b0"] < 1.0e-6 and rm["grad_dir"] < 1.0e-4 and rm["random"] < 1.0e-3


@pytest.mark.parametrize("divb,riemann,lim", [("glm", "hll", "cada3"), ("none", "hll", "cada3"),
                                              ("glm", "hll", "vanleer")])
def test_adjoint_stage_matches_torch_autograd(divb, riemann, lim):
    """One stage at one point: the device adjoint (J^T lam at u_in, with the
    psi row of lam scaled by the stage's decay as the checkpointed run does) against
    torch autograd of the FP64 PyTorch stage at the same u_in, 32^3.  No
    trajectory lies between the two, so limiter and Riemann branches are
    those of the same state."""
    print(f"[adjoint-stage] {divb} {riemann}/{lim} 32^3", flush=True)
    mesh = UniformMesh(3, (32, 32, 32), (16, 16, 16), (0.0,) * 3, (1.0,) * 3, ("periodic",) * 6)
    par = Par(physics="mhd", ndim=3, ndir=3, gamma=GAMMA, typedivbfix=divb, time_stepper="threestep",
              flux_scheme=riemann, limiter=lim, courantpar=0.6)
    glm = divb == "glm"
    dev = torch.device("cuda")
    idx = torch.as_tensor(adjoint.block_index(mesh), device=dev)
    rho, v, p, b = adjoint.orszag_tang_fields(mesh, SEED, device=dev)
    gen = torch.Generator(device="cpu").manual_seed(SEED + 5)
    noise = lambda q: q + 1.0e-3 * torch.randn(q.shape, generator=gen, dtype=torch.float64).to(dev)
    v, b, p = [noise(c) for c in v], [noise(c) for c in b], noise(p)
    psi = 1.0e-3 * torch.randn(rho.shape, generator=gen, dtype=torch.float64).to(dev) if glm else None
    u = adjoint.mhd_conserved(rho, v, p, b, par.gamma, psi=psi)
    lam = torch.randn(u.shape, generator=gen, dtype=torch.float64).to(dev)
    run = adjoint.SSPRK3Run(par, mesh, 1)
    be = run.be
    ub = adjoint.to_blocks(u, idx).contiguous()
    dt, ch = be.getdt_state(ub.data_ptr(), par.courantpar)
    beta = 2.0 / 3.0
    lb = adjoint.to_blocks(lam, idx).contiguous()
    lam_scaled = lb.clone()
    run._psi_decay(lam_scaled, beta, dt, ch)
    out = torch.empty_like(lb)
    be.adjoint1(run.method, beta, dt, ub.data_ptr(), lam_scaled.data_ptr(), 0.0, out.data_ptr(), ch=ch)
    torch.cuda.synchronize()
    g_dev = adjoint.from_blocks(out, idx, (32, 32, 32)).cpu()
    sch = torch_fv_ref.Scheme(3, "mhd", par.gamma, lim, riemann, glm=glm)
    xc = u.detach().cpu().clone().requires_grad_(True)
    oc = sch.stage(xc, torch.zeros_like(xc), beta * dt, mesh.dx, ch)
    g_ref, = torch.autograd.grad((oc * lam.cpu()).sum(), xc)
    rel = ((g_dev - g_ref).norm() / g_ref.norm()).item()
    worst = ((g_dev - g_ref).abs().max() / g_ref.abs().max()).item()
    print(f"[adjoint-stage] rel L2 {rel:.3e}, max {worst:.3e}", flush=True)
    # the device's primitive states round differently from torch's (refined
    # hardware reciprocal, FMA), which the ill-conditioned cada3 gradient
    # turns into 1e-10 (measured 1.0e-10 GLM, 1.9e-10 none); vanleer 1.7e-14
    assert rel < (1e-9 if lim == "cada3" else 1e-12)



def test_glm_adjoint_finite_differences():
    """3-D MHD + GLM (interface states, psi / B_n fluxes, psi decay and the
    extended sources) with HLL, Einfeldt and vanleer at 64^3 for 10 steps,
    FP64: the adjoint against FP64 central differences with the dt and c_h
    sequences of the unperturbed run held fixed (measured: b0 5e-12, the
    gradient direction 1.0e-7, random directions 4e-4 normwise; the random
    directions as at 32^3 above).  The benchmark's cada3 limiter is not
    used here: limiter_cada3 switches between its third-order and limited
    forms at eta = 1 within a band of 2e-14, a jump that differences of a
    run across 1e8 limiter evaluations cross; its reverse is checked face by
    face against differences (tests/test_face_adjoint_cpu.py, jump points
    skipped) and stage by stage against torch autograd."""
    print("[adjoint-glm] 64^3, 10 steps, GLM / HLL / vanleer, FP64", flush=True)
    pb = _MHDProblem(64, 10, "fp64", every=3, divb="glm", lim="vanleer")
    _, gb, gr = pb.grad()
    r = _fd(pb, gb, gr, eps_random=1.0e-6)
    assert r["b0"] < 1.0e-4 and r["grad_dir"] < 1.0e-4 and r["random"] < 1.0e-2
