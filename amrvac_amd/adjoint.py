"""Differentiable runs of the device scheme: a torch.autograd.Function whose
forward is N ssprk3 steps of the HIP stage kernel and whose backward runs the
hand-written adjoint stage kernels (amrvac_fv_adjoint1) with
checkpoint-and-recompute in time, or, when the stage states of the whole run
fit in device memory (a 128^3 MHD run of 100 steps needs about 46 GB of the
GPU's 288 GB), a tape of all of them and no recompute.

The time-step sequence is taken from the forward pass (CFL condition of
src/mod_dt.t:160-287) and held fixed in the backward pass, like the limiter
and Riemann branches: the gradient is that of the scheme with the forward
run's dt sequence (and, with GLM, its sequence of cleaning speeds c_h =
cmax_global, src/mod_dt.t:128-131, which set the psi decay of each stage).  Checkpoints of the state every ``every`` steps are kept
in device memory; in the backward pass each segment is recomputed once,
keeping its stage states on a tape, and then reversed stage by stage (three
adjoint launches per step, the RK combination fused into the last).

PyTorch supplies device memory and the stream: the context's work is
ordered on torch's current stream.
"""
from __future__ import annotations

import math

import numpy as np
import torch

from .fv import DeviceFV
from .host import FS, Par, ssprk3_coefficients
from .mesh import UniformMesh


class SSPRK3Run:
    """N steps of threestep (SSPRK(3,3), src/mod_advance.t:251-267) on a
    uniform periodic level, differentiable w.r.t. the initial state.

    States live in torch device tensors in the block layout [nb, nw, (z,) y,
    x] and are handed to the stage and adjoint kernels by pointer
    (amrvac_fv_stage / amrvac_fv_adjoint1): no copies between steps."""

    def __init__(self, par: Par, mesh: UniformMesh, nsteps: int, every: int | None = None,
                 precision: str | None = None, backend=None, tape_bytes: int | None = None):
        """every: steps per checkpoint segment.  By default the whole run is
        one segment when its stage tape ((nsteps+1) x 3 states) fits in
        tape_bytes (default: half the device memory free at construction),
        so that the backward pass recomputes nothing; else sqrt(nsteps)."""
        if par.time_stepper != "threestep" or par.integrator != "ssprk3" or par.ssprk_order != 3:
            raise ValueError("the differentiable run integrates with threestep SSPRK(3,3)")
        if any(b != "periodic" for b in mesh.bc):
            raise ValueError("the differentiable run needs a periodic domain")
        self.par = par
        self.mesh = mesh
        self.nsteps = nsteps
        # the backend: DeviceFV (the HIP kernels), or an object with the same
        # stage / adjoint1 / getdt_state / set_stream methods on the
        # tensors' memory (the checkpoint-schedule tests use a CPU one)
        self.be = backend if backend is not None else DeviceFV(par, mesh, precision=precision)
        self.device_backend = backend is None
        self.precision = precision
        self.shape = (mesh.nblocks, par.nw) + tuple(mesh.block_nx[::-1])
        if every is None:
            tape = (nsteps + 1) * 3 * math.prod(self.shape) * 8
            if tape_bytes is None:
                tape_bytes = torch.cuda.mem_get_info()[0] // 2 if self.device_backend else 0
            every = nsteps if tape <= tape_bytes else max(1, int(math.sqrt(nsteps)))
        self.every = every
        self.method = FS[par.flux_scheme]
        self.c = ssprk3_coefficients()
        self.dts: list[float] | None = None
        self.chs: list[float] | None = None
        # GLM: the psi decay of a stage, exp(-qdt c_h alpha / min dx)
        # (add_source_glm, src/mhd/mod_mhd_phys.t:5491-5495), is a diagonal
        # factor on the psi row of the adjoint
        self.glm = par.physics == "mhd" and par.typedivbfix == "glm"
        self.psi = 2 * par.ndir + 2
        self.mindx = min(mesh.dx[:mesh.ndim])
        self.checkpoints: dict[int, torch.Tensor] = {}
        self._tape_buf: torch.Tensor | None = None
        self._tape_segment = -1

    def _bind_stream(self) -> None:
        if self.device_backend:
            self.be.set_stream(torch.cuda.current_stream().cuda_stream)

    def _step(self, u: int, u1: int, u2: int, out: int, dt: float, ch: float) -> None:
        """out = ssprk3 step of u (pointers); u1, u2 receive the stage states."""
        c, m, be = self.c, self.method, self.be
        be.stage(m, c["beta11"], dt, ch, u, u1)
        be.stage(m, c["beta22"], dt, ch, u1, u2, base=u, comb_a=c["alfa21"], comb_b=c["alfa22"])
        be.stage(m, c["beta33"], dt, ch, u2, out, base=u, comb_a=c["alfa31"], comb_b=c["alfa33"])

    def _psi_decay(self, lam: torch.Tensor, beta: float, dt: float, ch: float) -> None:
        """With GLM: lam's psi row times the stage's psi decay, in place."""
        if self.glm:
            lam[:, self.psi].mul_(math.exp(-beta * dt * ch * self.par.mhd_glm_alpha / self.mindx))

    def _tape(self, like: torch.Tensor) -> torch.Tensor:
        """Stage states of one segment: tape[j] = (u0, u1, u2) of its step j
        (every+1 entries, the last holding the segment's end state)."""
        if getattr(self, "_tape_buf", None) is None or self._tape_buf.device != like.device:
            self._tape_buf = torch.empty((self.every + 1, 3) + tuple(like.shape), dtype=torch.float64,
                                         device=like.device)
        return self._tape_buf

    def _on_own_stream(self, fn, arg: torch.Tensor, *rest):
        """fn(arg, *rest) on the run's own stream (default priority), ordered
        after the caller's stream and before its later work.  The library
        and the torch operations between its calls (the psi decay, copies)
        then share one stream: on the legacy default stream those
        operations would serialise with every other stream, the
        recompute's included."""
        if not self.device_backend:
            return fn(arg, *rest)
        caller = torch.cuda.current_stream()
        if getattr(self, "_own", None) is None:
            self._own = torch.cuda.Stream()
        own = self._own
        own.wait_stream(caller)
        arg.record_stream(own)
        with torch.cuda.stream(own):
            out = fn(arg, *rest)
        caller.wait_stream(own)
        out.record_stream(caller)
        return out

    def forward(self, u0: torch.Tensor, dts: list[float] | None = None,
                chs: list[float] | None = None) -> torch.Tensor:
        """Final state; keeps checkpoints for backward.  With ``dts`` the
        given time steps are replayed instead of the CFL ones, and with
        them the GLM speeds ``chs`` (by default those of the previous
        forward pass)."""
        return self._on_own_stream(self._forward, u0, dts, chs)

    def _forward(self, u0: torch.Tensor, dts: list[float] | None = None,
                 chs: list[float] | None = None) -> torch.Tensor:
        """Final state; keeps checkpoints for backward.  With ``dts`` the
        given time steps are replayed instead of the CFL ones, and with
        them the GLM speeds ``chs`` (by default those of the previous
        forward pass).  The last two
        segments run on the two stage tapes the backward pass double-buffers,
        so it starts without recomputing them."""
        if u0.shape != self.shape or u0.dtype != torch.float64 or u0.is_cuda != self.device_backend:
            raise ValueError(f"initial state must be a float64 {'CUDA ' if self.device_backend else ''}tensor "
                             f"of shape {self.shape}")
        if dts is not None and len(dts) != self.nsteps:
            raise ValueError("dts must hold one time step per step")
        if dts is not None and chs is None:
            chs = self.chs if self.chs is not None else [0.0] * self.nsteps
        if chs is not None and len(chs) != self.nsteps:
            raise ValueError("chs must hold one GLM speed per step")
        self._bind_stream()
        u = u0.detach().contiguous().clone()
        u1 = torch.empty_like(u)
        u2 = torch.empty_like(u)
        self.checkpoints = {0: u.clone()}
        n_last = ((self.nsteps - 1) // self.every) * self.every
        # first step on a tape: the segment before the last one (second
        # tape), if there is one
        n_tape = n_last - self.every if n_last >= self.every else n_last
        tapes = {n_last: self._tape(u)}
        if n_tape < n_last:
            tapes[n_tape] = self._tape2(u)
        out = torch.empty_like(u)
        used, used_ch = [], []
        for n in range(self.nsteps):
            if n < n_tape:
                cur = u.data_ptr()
            else:
                s0 = n_last if n >= n_last else n_tape
                tape, j = tapes[s0], n - s0
                if n == n_tape:
                    tape[0, 0].copy_(u)
                cur = tape[j, 0].data_ptr()
            if dts is not None:
                dt, ch = dts[n], chs[n]
            else:
                dt, ch = self.be.getdt_state(cur, self.par.courantpar)
            used.append(dt)
            used_ch.append(ch)
            if n < n_tape:
                # the third stage overwrites u in place (its RK base, read per cell)
                self._step(cur, u1.data_ptr(), u2.data_ptr(), cur, dt, ch)
                if (n + 1) % self.every == 0:
                    self.checkpoints[n + 1] = u.clone()
            else:
                if n + 1 == self.nsteps:
                    nxt = out.data_ptr()
                elif n + 1 == n_last:  # into the last segment's tape
                    nxt = tapes[n_last][0, 0].data_ptr()
                else:
                    nxt = tape[j + 1, 0].data_ptr()
                self._step(cur, tape[j, 1].data_ptr(), tape[j, 2].data_ptr(), nxt, dt, ch)
                if n + 1 == n_last:
                    self.checkpoints[n_last] = tapes[n_last][0, 0].clone()
        self.dts = used
        self.chs = used_ch
        self._tape_segment = n_last
        self._tape2_segment = n_tape if n_tape < n_last else -1
        return out

    def _recompute(self, be, tape: torch.Tensor, n0: int, n1: int) -> None:
        """Stage states of steps n0..n1-1 into tape, from the checkpoint."""
        c, m = self.c, self.method
        tape[0, 0].copy_(self.checkpoints[n0])
        for j in range(n1 - n0):
            dt, ch = self.dts[n0 + j], self.chs[n0 + j]
            u0, u1, u2 = (tape[j, q].data_ptr() for q in range(3))
            be.stage(m, c["beta11"], dt, ch, u0, u1)
            be.stage(m, c["beta22"], dt, ch, u1, u2, base=u0, comb_a=c["alfa21"], comb_b=c["alfa22"])
            if j + 1 < n1 - n0:
                be.stage(m, c["beta33"], dt, ch, u2, tape[j + 1, 0].data_ptr(), base=u0,
                         comb_a=c["alfa31"], comb_b=c["alfa33"])

    def _reverse(self, tape: torch.Tensor, n0: int, n1: int, lam, l1, l2) -> None:
        """The adjoint stages of steps n1-1..n0 from the stage tape."""
        c, m, be = self.c, self.method, self.be
        for j in range(n1 - n0 - 1, -1, -1):
            dt, ch = self.dts[n0 + j], self.chs[n0 + j]
            u0, u1, u2 = (tape[j, q].data_ptr() for q in range(3))
            # u3 = a31 u0 + a33 u2 + L3(u2) (psi row: times the decay D3)
            self._psi_decay(lam, c["beta33"], dt, ch)
            be.adjoint1(m, c["beta33"], dt, u2, lam.data_ptr(), c["alfa33"], l2.data_ptr(), ch=ch)
            # u2 = a21 u0 + a22 u1 + L2(u1)
            self._psi_decay(l2, c["beta22"], dt, ch)
            be.adjoint1(m, c["beta22"], dt, u1, l2.data_ptr(), c["alfa22"], l1.data_ptr(), ch=ch)
            # u1 = u0 + L1(u0); lam_u0 = l1 + L1^T l1 + a31 lam + a21 l2 (in
            # place; lam and l2 already carry their stages' decay)
            self._psi_decay(l1, c["beta11"], dt, ch)
            be.adjoint1(m, c["beta11"], dt, u0, l1.data_ptr(), 1.0, lam.data_ptr(),
                        add_a=lam.data_ptr(), ca=c["alfa31"], add_b=l2.data_ptr(), cb=c["alfa21"], ch=ch)

    def backward(self, lam: torch.Tensor) -> torch.Tensor:
        """Adjoint of the whole run: d(final)/d(initial)^T lam."""
        return self._on_own_stream(self._backward, lam)

    def _backward(self, lam: torch.Tensor) -> torch.Tensor:
        """Adjoint of the whole run: d(final)/d(initial)^T lam.

        On the device the recompute of the next segment (on a second stream
        and context, into a second tape) overlaps the reverse sweep of the
        current one; events order each tape's reuse."""
        self._bind_stream()
        lam = lam.detach().to(torch.float64).contiguous().clone()
        l1 = torch.empty_like(lam)
        l2 = torch.empty_like(lam)
        starts = sorted(self.checkpoints)
        segs = [(n0, starts[i + 1] if i + 1 < len(starts) else self.nsteps) for i, n0 in enumerate(starts)][::-1]
        have_last = segs[0][0] == self._tape_segment
        have_prev = len(segs) > 1 and segs[1][0] == getattr(self, "_tape2_segment", -1)
        self._tape_segment = self._tape2_segment = -1  # the tapes are reused from here on
        if not self.device_backend:
            tape = self._tape(lam)
            for i, (n0, n1) in enumerate(segs):
                if i > 0 or not have_last:
                    self._recompute(self.be, tape, n0, n1)
                self._reverse(tape, n0, n1, lam, l1, l2)
            return lam
        tapes = [self._tape(lam), self._tape2(lam)]
        main = torch.cuda.current_stream()
        side = self._side_stream()
        side.wait_stream(main)
        ready: list = [None] * len(segs)
        free: list = [None, None]

        def launch_recompute(i: int) -> None:
            k = i % 2
            with torch.cuda.stream(side):
                if free[k] is not None:
                    side.wait_event(free[k])
                self._recompute(self._side_be, tapes[k], *segs[i])
                ev = torch.cuda.Event()
                ev.record(side)
                ready[i] = ev

        if not have_last:
            launch_recompute(0)
        for i, (n0, n1) in enumerate(segs):
            if i + 1 < len(segs) and not (i + 1 == 1 and have_prev):
                launch_recompute(i + 1)
            if ready[i] is not None:
                main.wait_event(ready[i])
            self._reverse(tapes[i % 2], n0, n1, lam, l1, l2)
            ev = torch.cuda.Event()
            ev.record(main)
            free[i % 2] = ev
        main.wait_stream(side)
        return lam

    def _tape2(self, like: torch.Tensor) -> torch.Tensor:
        if getattr(self, "_tape_buf2", None) is None or self._tape_buf2.device != like.device:
            self._tape_buf2 = torch.empty_like(self._tape(like))
        return self._tape_buf2

    def _side_stream(self):
        """A second stream (default priority) and a second context on it for
        the overlapped recompute."""
        if getattr(self, "_side", None) is None:
            self._side = torch.cuda.Stream()
            self._side_be = DeviceFV(self.par, self.mesh, precision=self.precision)
            self._side_be.set_stream(self._side.cuda_stream)
        return self._side


class _RunFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, u0, run: SSPRK3Run, dts, chs):
        ctx.run = run
        return run.forward(u0, dts, chs)

    @staticmethod
    def backward(ctx, grad):
        return ctx.run.backward(grad), None, None, None


def run_differentiable(u0: torch.Tensor, run: SSPRK3Run, dts: list[float] | None = None,
                       chs: list[float] | None = None) -> torch.Tensor:
    """Final state of ``run`` from ``u0`` (block layout [nb, nw, (z,) y, x]),
    differentiable w.r.t. u0 through the adjoint kernels; ``dts`` (and the
    GLM speeds ``chs``) replay a time-step sequence."""
    return _RunFunction.apply(u0, run, dts, chs)


# ---------------------------------------------------------------------------
# global fields <-> block layout
# ---------------------------------------------------------------------------
def block_index(mesh: UniformMesh) -> np.ndarray:
    """Flat global cell index (row-major over (z,) y, x) of every interior
    cell in the block layout [nb, (z,) y, x]."""
    nd = mesh.ndim
    shp = tuple(mesh.block_nx[::-1])
    dom = tuple(mesh.domain_nx[::-1])
    out = np.zeros((mesh.nblocks,) + shp, dtype=np.int64)
    for b in range(mesh.nblocks):
        coords = []
        for a in range(nd):
            d = nd - 1 - a
            g = (int(mesh.ig[b, d]) - 1) * mesh.block_nx[d] + np.arange(mesh.block_nx[d])
            coords.append(g)
        grids = np.meshgrid(*coords, indexing="ij")
        out[b] = np.ravel_multi_index(tuple(grids), dom)
    return out


def to_blocks(u: torch.Tensor, index: torch.Tensor) -> torch.Tensor:
    """Global fields [nw, (nz,) ny, nx] -> block layout [nb, nw, (z,) y, x]
    (a differentiable gather)."""
    nw = u.shape[0]
    g = u.reshape(nw, -1)[:, index.reshape(-1)].reshape((nw,) + tuple(index.shape))
    return g.movedim(0, 1).contiguous()


def from_blocks(blocks: torch.Tensor, index: torch.Tensor, domain_shape) -> torch.Tensor:
    """Block layout [nb, nw, (z,) y, x] -> global fields [nw, (nz,) ny, nx]
    (a differentiable scatter)."""
    nw = blocks.shape[1]
    src = blocks.movedim(1, 0).reshape(nw, -1)
    out = torch.zeros((nw, int(np.prod(domain_shape))), dtype=blocks.dtype, device=blocks.device)
    out = out.index_copy(1, index.reshape(-1), src)
    return out.reshape((nw,) + tuple(domain_shape))


def hd_conserved(rho: torch.Tensor, v: list[torch.Tensor], p: torch.Tensor, gamma: float) -> torch.Tensor:
    """hd_to_conserved on fields: [nw, ...]."""
    vsq = v[0] * v[0]
    for c in v[1:]:
        vsq = vsq + c * c
    e = p * (1.0 / (gamma - 1.0)) + 0.5 * vsq * rho
    return torch.stack([rho] + [rho * c for c in v] + [e])


def kh_fields(mesh: UniformMesh, seed: int, amp: float = 1.0e-3, device=None, rho_amp: float = 0.0):
    """Kelvin-Helmholtz primitive fields (rho, vx, vy, p) on the global grid:
    rho=2, vx=+0.5 for |y-0.5|<0.25, else rho=1, vx=-0.5; p=2.5;
    vy = 0.01 sin(4 pi x + phase) with a seeded phase; seeded uniform noise
    of amplitude ``amp`` on vx and vy and of ``rho_amp`` on rho (hash of the
    global cell index).  With rho_amp=0 every density slope inside the bands
    is exactly zero, the kink of the limiters, where the run has one-sided
    derivatives only."""
    from .problems import hashed_uniform
    ny, nx = mesh.domain_nx[1], mesh.domain_nx[0]
    dx = [(mesh.xprobmax[d] - mesh.xprobmin[d]) / mesh.domain_nx[d] for d in range(2)]
    x = mesh.xprobmin[0] + (np.arange(nx) + 0.5) * dx[0]
    y = mesh.xprobmin[1] + (np.arange(ny) + 0.5) * dx[1]
    Y, X = np.meshgrid(y, x, indexing="ij")
    gi = np.stack(np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")[::-1])  # [2, ny, nx] (x, y)
    phase = 2.0 * math.pi * float(np.random.default_rng(seed).random())
    inner = np.abs(Y - 0.5) < 0.25
    rho = np.where(inner, 2.0, 1.0)
    vx = np.where(inner, 0.5, -0.5) + amp * hashed_uniform(seed, 1, gi, 1)
    vy = 0.01 * np.sin(4.0 * math.pi * X + phase) + amp * hashed_uniform(seed, 1, gi, 2)
    if rho_amp:
        rho = rho + rho_amp * hashed_uniform(seed, 1, gi, 0)
    p = np.full_like(rho, 2.5)
    t = lambda a: torch.as_tensor(a, dtype=torch.float64, device=device)
    return t(rho), t(vx), t(vy), t(p)


def mhd_conserved(rho: torch.Tensor, v: list[torch.Tensor], p: torch.Tensor, b: list[torch.Tensor],
                  gamma: float, psi: torch.Tensor | None = None) -> torch.Tensor:
    """mhd_to_conserved on fields (no B0): [rho, m_1..3, e, b_1..3] (and
    psi with GLM)."""
    vsq = v[0] * v[0]
    for c in v[1:]:
        vsq = vsq + c * c
    bsq = b[0] * b[0]
    for c in b[1:]:
        bsq = bsq + c * c
    e = p * (1.0 / (gamma - 1.0)) + 0.5 * (vsq * rho + bsq)
    return torch.stack([rho] + [rho * c for c in v] + [e] + list(b) + ([psi] if psi is not None else []))


def orszag_tang_fields(mesh: UniformMesh, seed: int, rho_amp: float = 1.0e-2, vz_amp: float = 1.0e-3, device=None,
                       field_amp: float = 0.0):
    """3-D Orszag-Tang fields on the global grid [nz, ny, nx]: rho =
    25/(36 pi) (1 + rho_amp r), p = 5/(12 pi), v = (-sin 2 pi y, sin 2 pi x,
    vz_amp r'), and the unit-amplitude field shape (-sin 2 pi y, sin 4 pi x,
    0)/sqrt(4 pi) that the scalar B amplitude multiplies; r, r' seeded
    uniform noise in [-1, 1) (hash of the global cell index).  field_amp > 0
    adds seeded noise of that amplitude to vx, vy, p (relative) and to the
    field shape (in units of 1/sqrt(4 pi)): the plain fields are constant
    along some axes, where every limiter slope sits on its kink."""
    from .problems import hashed_uniform
    nx, ny, nz = mesh.domain_nx
    dx = [(mesh.xprobmax[d] - mesh.xprobmin[d]) / mesh.domain_nx[d] for d in range(3)]
    x = mesh.xprobmin[0] + (np.arange(nx) + 0.5) * dx[0]
    y = mesh.xprobmin[1] + (np.arange(ny) + 0.5) * dx[1]
    Z, Y, X = np.meshgrid(np.arange(nz), y, x, indexing="ij")
    gz, gy, gx = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    gi = np.stack([gx, gy, gz])
    rho = 25.0 / (36.0 * math.pi) * (1.0 + rho_amp * hashed_uniform(seed, 1, gi, 0))
    vx = -np.sin(2.0 * math.pi * Y)
    vy = np.sin(2.0 * math.pi * X)
    vz = vz_amp * hashed_uniform(seed, 1, gi, 3)
    p = np.full_like(rho, 5.0 / (12.0 * math.pi))
    s4 = 1.0 / math.sqrt(4.0 * math.pi)
    bx = -s4 * np.sin(2.0 * math.pi * Y)
    by = s4 * np.sin(4.0 * math.pi * X)
    bz = np.zeros_like(rho)
    if field_amp:
        vx = vx + field_amp * hashed_uniform(seed, 1, gi, 1)
        vy = vy + field_amp * hashed_uniform(seed, 1, gi, 2)
        p = p * (1.0 + field_amp * hashed_uniform(seed, 1, gi, 4))
        bx = bx + s4 * field_amp * hashed_uniform(seed, 1, gi, 5)
        by = by + s4 * field_amp * hashed_uniform(seed, 1, gi, 6)
        bz = bz + s4 * field_amp * hashed_uniform(seed, 1, gi, 7)
    t = lambda a: torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device=device)
    return t(rho), [t(vx), t(vy), t(vz)], t(p), [t(bx), t(by), t(bz)]

