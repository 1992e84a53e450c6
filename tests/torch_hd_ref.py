"""FP64 PyTorch restatement of the 2-D hydrodynamic scheme on a periodic
uniform grid (test infrastructure): threestep (SSPRK3), TVD reconstruction of
the primitive variables with vanleer or minmod, Einfeldt bounds and the HLLC
or HLL Riemann flux, as in src/mod_finite_volume.t, src/limiter/mod_limiter.t,
src/hd/mod_hd_phys.t and src/hd/mod_hd_hllc.t.  Written with torch tensor
operations on global fields so that torch.autograd differentiates it: the
independent reverse-mode reference for the hand-written adjoint kernels.

Fields are [nw, ny, nx] (rho, m_x, m_y, e); dimension d is tensor axis
2 - d (x along the last axis)."""
from __future__ import annotations

import torch


def to_prim(u, gamma):
    rho = u[0]
    inv = 1.0 / rho
    vx = u[1] * inv
    vy = u[2] * inv
    vsq = vx * vx + vy * vy
    p = (gamma - 1.0) * (u[3] - 0.5 * rho * vsq)
    return torch.stack([rho, vx, vy, p])


def to_cons(w, gamma):
    vsq = w[1] * w[1] + w[2] * w[2]
    e = w[3] * (1.0 / (gamma - 1.0)) + 0.5 * vsq * w[0]
    return torch.stack([w[0], w[0] * w[1], w[0] * w[2], e])


def limiter(a, b, kind):
    if kind == "vanleer":
        return 2.0 * torch.clamp(a * b, min=0.0) / (b + a + 1.0e-12)
    s = torch.sign(b)
    s = torch.where(s == 0, torch.ones_like(s), s)
    return s * torch.clamp(torch.minimum(torch.abs(b), s * a), min=0.0)


def flux(wc, w, idim):
    vn = w[1 + idim]
    f = [vn * w[0], vn * wc[1], vn * wc[2], vn * (wc[3] + w[3])]
    f[1 + idim] = f[1 + idim] + w[3]
    return torch.stack(f)


def face_flux(w, idim, gamma, lim, riemann):
    """Flux at the face between cell i and i+1 along idim, indexed by i."""
    ax = 2 - idim
    sh = lambda k: torch.roll(w, shifts=-k, dims=ax)  # value at i+k
    wm, w0, wp, wpp = sh(-1), w, sh(1), sh(2)
    wL = w0 + 0.5 * limiter(w0 - wm, wp - w0, lim)
    wR = wp - 0.5 * limiter(wp - w0, wpp - wp, lim)
    cL, cR = to_cons(wL, gamma), to_cons(wR, gamma)
    fL, fR = flux(cL, wL, idim), flux(cR, wR, idim)
    mn = 1 + idim
    t1, t2 = torch.sqrt(wL[0]), torch.sqrt(wR[0])
    t3 = 1.0 / (t1 + t2)
    um = (wL[mn] * t1 + wR[mn] * t2) * t3
    csL = gamma * wL[3] / wL[0]
    csR = gamma * wR[3] / wR[0]
    dv = wR[mn] - wL[mn]
    dmean = torch.sqrt((t1 * csL + t2 * csR) * t3 + 0.5 * t1 * t2 * (t3 * t3) * (dv * dv))
    if riemann == "tvdlf":  # tvdlfeps = 1, src/mod_finite_volume.t:348-370
        cm = torch.abs(um) + dmean
        return 0.5 * (fL + fR) - 0.5 * cm * (cR - cL)
    cmin, cmax = um - dmean, um + dmean
    mid = (cmin < 0) & (cmax > 0)
    D = torch.where(mid, cmax - cmin, torch.ones_like(cmax))
    fhll = (cmax * fL - cmin * fR + cmin * cmax * (cR - cL)) / D
    hll = torch.where(cmin >= 0, fL, torch.where(cmax <= 0, fR, fhll))
    if riemann == "hll":
        return hll
    whr = (cmax * cR[0] - cmin * cL[0] + fL[0] - fR[0]) / D
    whm = (cmax * cR[mn] - cmin * cL[mn] + fL[mn] - fR[mn]) / D
    whr = torch.where(mid, whr, torch.ones_like(whr))
    lam = whm / whr
    # patch: -1 left star state, 1 right star state, 3 HLL
    left = (cmin < 0) & (lam > 0) & (lam < cmax)
    right = ~left & (cmax > 0) & (lam < 0) & (lam > cmin)
    degen = ~left & ~right & ((lam >= cmax) | (lam <= cmin))
    zero = ~left & ~right & ~degen & (lam == 0)
    right = right | (zero & (-cmin >= cmax))
    left = left | (zero & ~(-cmin >= cmax))
    use_right = right

    def star(cs, wS, fS):
        vS = wS[mn] / wS[0]
        den = torch.where(mid & ~degen, cs - lam, torch.ones_like(cs))
        csmls = 1.0 / den
        wcd = [None] * 4
        wcd[0] = wS[0] * (cs - vS) * csmls
        t = 1 - idim  # transverse momentum index - 1
        wcd[1 + t] = (cs * wS[1 + t] - fS[1 + t]) * csmls
        wcd[mn] = wcd[0] * lam
        pcd = wS[0] * (cs - vS) * (lam - vS) + fS[mn] - wS[mn] * vS
        wcd[3] = (cs * wS[3] - fS[3] + lam * pcd) * csmls
        wcd = torch.stack(wcd)
        return fS + cs * (wcd - wS)

    fcd = torch.where(use_right, star(cmax, cR, fR), star(cmin, cL, fL))
    inner = torch.where(left | right, fcd, fhll)
    return torch.where(cmin >= 0, fL, torch.where(cmax <= 0, fR, inner))


def stage(u_in, init, qdt, dx, gamma, lim, riemann):
    """init + sum_d (-qdt/dx_d)(fC(i) - fC(i-1)), dimension by dimension."""
    w = to_prim(u_in, gamma)
    acc = init
    for idim in range(2):
        ax = 2 - idim
        dq = -qdt / dx[idim]
        f = dq * face_flux(w, idim, gamma, lim, riemann)
        acc = acc + f - torch.roll(f, shifts=1, dims=ax)
    return acc


def ssprk3_step(u, dt, dx, gamma, lim="vanleer", riemann="hllc"):
    a21, a31 = 3.0 / 4.0, 1.0 / 3.0
    a22, a33 = 1.0 - a21, 1.0 - a31
    u1 = stage(u, u, 1.0 * dt, dx, gamma, lim, riemann)
    u2 = stage(u1, a21 * u + a22 * u1, 0.25 * dt, dx, gamma, lim, riemann)
    return stage(u2, a31 * u + a33 * u2, (2.0 / 3.0) * dt, dx, gamma, lim, riemann)


def run(u0, dts, dx, gamma, lim="vanleer", riemann="hllc"):
    u = u0
    for dt in dts:
        u = ssprk3_step(u, dt, dx, gamma, lim, riemann)
    return u
