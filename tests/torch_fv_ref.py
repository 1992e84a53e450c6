"""FP64 PyTorch restatement of the finite-volume stage on a periodic uniform
grid (test infrastructure), for 2-D / 3-D hydrodynamics and ideal MHD
without divergence cleaning or with GLM: threestep (SSPRK3), TVD
reconstruction of the primitive variables with vanleer, minmod or cada3,
Einfeldt bounds, and the TVDLF, HLL or (hydrodynamics) HLLC Riemann flux, as
in src/mod_finite_volume.t (:116-339, :348-504), src/limiter/mod_limiter.t
(limiter_cada3 :253-290), src/hd/mod_hd_phys.t, src/hd/mod_hd_hllc.t and
src/mhd/mod_mhd_phys.t (mhd_get_flux :3810-3889, mhd_get_cbounds
:3037-3127, mhd_get_csound_prim :3329-3379, flux_type :580-588; GLM:
mhd_modify_wLR :6326-6377, add_source_glm :5472-5531).  Written with torch tensor operations on global fields so that
torch.autograd differentiates it: the independent reverse-mode reference
for the hand-written adjoint kernels.

Fields are [nw, (nz,) ny, nx]: rho, m_1..ndir, e (, b_1..ndir); dimension d
(0 = x) is tensor axis ndim - d.
"""
from __future__ import annotations

import torch


class Scheme:
    def __init__(self, ndim, physics="hd", gamma=1.4, lim="vanleer", riemann="hllc", ndir=None, glm=False,
                 glm_alpha=0.5, glm_extended=True, cada3_radius=0.1):
        self.ndim = ndim
        self.mhd = physics == "mhd"
        self.ndir = ndir or (3 if self.mhd else ndim)
        self.gamma = gamma
        self.lim = lim
        self.riemann = riemann
        self.glm = self.mhd and glm
        self.glm_alpha = glm_alpha
        self.glm_extended = glm_extended
        self.cada3_radius = cada3_radius
        self.E = self.ndir + 1
        self.MAG = self.ndir + 2
        self.PSI = 2 * self.ndir + 2
        self.nw = (2 * self.ndir + 2 + (1 if self.glm else 0)) if self.mhd else self.ndir + 2

    # -- conversions -----------------------------------------------------------
    def to_prim(self, u):
        nd, E = self.ndir, self.E
        rho = u[0]
        inv = 1.0 / rho
        v = [u[1 + k] * inv for k in range(nd)]
        vsq = v[0] * v[0]
        for k in range(1, nd):
            vsq = vsq + v[k] * v[k]
        if self.mhd:
            b = [u[self.MAG + k] for k in range(nd)]
            bsq = b[0] * b[0]
            for k in range(1, nd):
                bsq = bsq + b[k] * b[k]
            p = (self.gamma - 1.0) * (u[E] - 0.5 * (rho * vsq + bsq))
            return torch.stack([rho] + v + [p] + b + ([u[self.PSI]] if self.glm else []))
        p = (self.gamma - 1.0) * (u[E] - 0.5 * rho * vsq)
        return torch.stack([rho] + v + [p])

    def to_cons(self, w):
        nd, E = self.ndir, self.E
        vsq = w[1] * w[1]
        for k in range(1, nd):
            vsq = vsq + w[1 + k] * w[1 + k]
        m = [w[0] * w[1 + k] for k in range(nd)]
        if self.mhd:
            bsq = w[self.MAG] * w[self.MAG]
            for k in range(1, nd):
                bsq = bsq + w[self.MAG + k] * w[self.MAG + k]
            e = w[E] * (1.0 / (self.gamma - 1.0)) + 0.5 * (vsq * w[0] + bsq)
            return torch.stack([w[0]] + m + [e] + [w[self.MAG + k] for k in range(nd)]
                               + ([w[self.PSI]] if self.glm else []))
        e = w[E] * (1.0 / (self.gamma - 1.0)) + 0.5 * vsq * w[0]
        return torch.stack([w[0]] + m + [e])

    # -- fluxes and speeds -----------------------------------------------------
    def flux(self, wc, w, idim, ch=0.0):
        nd, E, mn = self.ndir, self.E, 1 + idim
        vn = w[mn]
        if not self.mhd:
            f = [vn * w[0]] + [vn * wc[1 + k] for k in range(nd)] + [vn * (wc[E] + w[E])]
            f[mn] = f[mn] + w[E]
            return torch.stack(f)
        M = self.MAG
        bn = w[M + idim]
        bsq = w[M] * w[M]
        bv = w[M] * w[1]
        for k in range(1, nd):
            bsq = bsq + w[M + k] * w[M + k]
            bv = bv + w[M + k] * w[1 + k]
        ptot = w[E] + 0.5 * bsq
        f = [vn * w[0]] + [wc[mn] * w[1 + k] - bn * w[M + k] for k in range(nd)]
        f[mn] = f[mn] + ptot
        f.append(vn * (wc[E] + ptot) - bn * bv)
        f += [vn * w[M + k] - bn * w[1 + k] for k in range(nd)]
        if self.glm:  # the normal field carries psi, psi carries ch^2 B_n
            f[M + idim] = w[self.PSI]
            f.append(ch * ch * bn)
        return torch.stack(f)

    def csq(self, w, idim):
        """Squared sound speed (hydrodynamics) or fast speed (MHD)."""
        E = self.E
        inv = 1.0 / w[0]
        cs = self.gamma * w[E] * inv
        if not self.mhd:
            return cs
        M = self.MAG
        b2 = w[M] * w[M]
        for k in range(1, self.ndir):
            b2 = b2 + w[M + k] * w[M + k]
        cf2 = b2 * inv + cs
        bn = w[M + idim]
        av = torch.clamp(cf2 * cf2 - 4.0 * cs * (bn * bn) * inv, min=0.0)
        c = torch.sqrt(0.5 * (cf2 + torch.sqrt(av)))
        return c * c

    def limiter(self, a, b):
        if self.lim == "vanleer":
            return 2.0 * torch.clamp(a * b, min=0.0) / (b + a + 1.0e-12)
        s = torch.sign(b)
        s = torch.where(s == 0, torch.ones_like(s), s)
        return s * torch.clamp(torch.minimum(torch.abs(b), s * a), min=0.0)

    @staticmethod
    def cada3_side(x, d0, rdel):
        """One face side of limiter_cada3 as face_limit forms it:
        phi(x / (d0 + sign(eps, d0)), (d0^2 + x^2) rdel) * d0."""
        eps = 1.4901161193847656e-08
        th = x / (d0 + torch.copysign(torch.full_like(d0, eps), d0))
        eta = (d0 * d0 + x * x) * rdel
        A = (2.0 + th) * (1.0 / 3.0)
        lim = torch.clamp(torch.minimum(A, torch.maximum(-0.5 * th, torch.clamp(torch.minimum(2.0 * th, A),
                                                                                max=1.6))), min=0.0)
        t2 = (eta - 1.0) * 1.0e14
        blend = 0.5 * ((1.0 - t2) * A + (1.0 + t2) * lim)
        phi = torch.where(eta <= 1.0 - 1.0e-14, A, torch.where(eta >= 1.0 + 1.0e-14, lim, blend))
        return phi * d0

    def face_flux(self, w, idim, dx=None, ch=0.0):
        """Flux at the face between cell i and i+1 along idim, indexed by i."""
        ax = self.ndim - idim
        sh = lambda k: torch.roll(w, shifts=-k, dims=ax)  # value at i+k
        wm, w0, wp, wpp = sh(-1), w, sh(1), sh(2)
        if self.lim == "cada3":
            rdel = 1.0 / (self.cada3_radius * dx[idim]) ** 2
            wL = w0 + 0.5 * self.cada3_side(w0 - wm, wp - w0, rdel)
            wR = wp - 0.5 * self.cada3_side(wpp - wp, wp - w0, rdel)
        else:
            wL = w0 + 0.5 * self.limiter(w0 - wm, wp - w0)
            wR = wp - 0.5 * self.limiter(wp - w0, wpp - wp)
        if self.glm:  # shared interface state of the normal field and psi
            bn, S = self.MAG + idim, self.PSI
            bs = 0.5 * (wR[bn] + wL[bn] - (wR[S] - wL[S]) / ch)
            ps = 0.5 * (wR[S] + wL[S] - (wR[bn] - wL[bn]) * ch)
            rows = [wL[i] for i in range(self.nw)]
            rows[bn], rows[S] = bs, ps
            wL = torch.stack(rows)
            rows = [wR[i] for i in range(self.nw)]
            rows[bn], rows[S] = bs, ps
            wR = torch.stack(rows)
        cL, cR = self.to_cons(wL), self.to_cons(wR)
        fL, fR = self.flux(cL, wL, idim, ch), self.flux(cR, wR, idim, ch)
        mn = 1 + idim
        t1, t2 = torch.sqrt(wL[0]), torch.sqrt(wR[0])
        t3 = 1.0 / (t1 + t2)
        um = (wL[mn] * t1 + wR[mn] * t2) * t3
        dv = wR[mn] - wL[mn]
        dmean = torch.sqrt((t1 * self.csq(wL, idim) + t2 * self.csq(wR, idim)) * t3
                           + 0.5 * t1 * t2 * (t3 * t3) * (dv * dv))
        if self.riemann == "tvdlf":  # tvdlfeps = 1, src/mod_finite_volume.t:348-370
            cm = torch.abs(um) + dmean
            return 0.5 * (fL + fR) - 0.5 * cm * (cR - cL)
        cmin, cmax = um - dmean, um + dmean
        mid = (cmin < 0) & (cmax > 0)
        D = torch.where(mid, cmax - cmin, torch.ones_like(cmax))
        fhll = (cmax * fL - cmin * fR + cmin * cmax * (cR - cL)) / D
        hll = torch.where(cmin >= 0, fL, torch.where(cmax <= 0, fR, fhll))
        if self.riemann == "hll":
            if self.mhd and self.ndim > 1 and not self.glm:
                # normal field: Lax-Friedrichs dissipation (flux_type)
                bnv = self.MAG + idim
                M = torch.maximum(cmax, torch.abs(cmin))
                rows = [hll[i] for i in range(self.nw)]
                rows[bnv] = -0.5 * M * (cR[bnv] - cL[bnv])
                hll = torch.stack(rows)
            return hll
        return self._hllc(cL, cR, fL, fR, fhll, cmin, cmax, mid, D, idim)

    def _hllc(self, cL, cR, fL, fR, fhll, cmin, cmax, mid, D, idim):
        E, mn = self.E, 1 + idim
        whr = (cmax * cR[0] - cmin * cL[0] + fL[0] - fR[0]) / D
        whm = (cmax * cR[mn] - cmin * cL[mn] + fL[mn] - fR[mn]) / D
        whr = torch.where(mid, whr, torch.ones_like(whr))
        lam = whm / whr
        left = (cmin < 0) & (lam > 0) & (lam < cmax)
        right = ~left & (cmax > 0) & (lam < 0) & (lam > cmin)
        degen = ~left & ~right & ((lam >= cmax) | (lam <= cmin))
        zero = ~left & ~right & ~degen & (lam == 0)
        right = right | (zero & (-cmin >= cmax))
        left = left | (zero & ~(-cmin >= cmax))

        def star(cs, wS, fS):
            vS = wS[mn] / wS[0]
            den = torch.where(mid & ~degen, cs - lam, torch.ones_like(cs))
            csmls = 1.0 / den
            wcd = [None] * self.nw
            wcd[0] = wS[0] * (cs - vS) * csmls
            for k in range(self.ndir):
                if k != idim:
                    wcd[1 + k] = (cs * wS[1 + k] - fS[1 + k]) * csmls
            wcd[mn] = wcd[0] * lam
            pcd = wS[0] * (cs - vS) * (lam - vS) + fS[mn] - wS[mn] * vS
            wcd[E] = (cs * wS[E] - fS[E] + lam * pcd) * csmls
            wcd = torch.stack(wcd)
            return fS + cs * (wcd - wS)

        fcd = torch.where(right, star(cmax, cR, fR), star(cmin, cL, fL))
        inner = torch.where(left | right, fcd, fhll)
        return torch.where(cmin >= 0, fL, torch.where(cmax <= 0, fR, inner))

    # -- time stepping -----------------------------------------------------------
    def stage(self, u_in, init, qdt, dx, ch=0.0):
        """init + sum_d (-qdt/dx_d)(fC(i) - fC(i-1)), dimension by dimension;
        with GLM the psi decay and the extended sources follow."""
        w = self.to_prim(u_in)
        acc = init
        for idim in range(self.ndim):
            ax = self.ndim - idim
            f = (-qdt / dx[idim]) * self.face_flux(w, idim, dx, ch)
            acc = acc + f - torch.roll(f, shifts=1, dims=ax)
        if not self.glm:
            return acc
        import math
        rows = [acc[i] for i in range(self.nw)]
        rows[self.PSI] = math.exp(-qdt * ch * self.glm_alpha / min(dx[:self.ndim])) * rows[self.PSI]
        if self.glm_extended:
            M, S = self.MAG, self.PSI
            nb = lambda q, d, k: torch.roll(q, shifts=-k, dims=self.ndim - 1 - d)  # field q at i + k e_d
            gp = [((nb(w[S], d, 1) - nb(w[S], d, -1)) * 0.5) / dx[d] for d in range(self.ndim)]
            divb = 0.5 * (nb(w[M], 0, 1) - nb(w[M], 0, -1)) * (1.0 / dx[0])
            for d in range(1, self.ndim):
                divb = divb + 0.5 * (nb(w[M + d], d, 1) - nb(w[M + d], d, -1)) * (1.0 / dx[d])
            for d in range(self.ndim):
                rows[self.E] = rows[self.E] - qdt * w[M + d] * gp[d]
            for k in range(self.ndir):
                rows[1 + k] = rows[1 + k] - qdt * w[M + k] * divb
        return torch.stack(rows)

    def ssprk3_step(self, u, dt, dx, ch=0.0):
        a21, a31 = 3.0 / 4.0, 1.0 / 3.0
        a22, a33 = 1.0 - a21, 1.0 - a31
        u1 = self.stage(u, u, 1.0 * dt, dx, ch)
        u2 = self.stage(u1, a21 * u + a22 * u1, 0.25 * dt, dx, ch)
        return self.stage(u2, a31 * u + a33 * u2, (2.0 / 3.0) * dt, dx, ch)

    def run(self, u0, dts, dx, chs=None):
        u = u0
        for n, dt in enumerate(dts):
            u = self.ssprk3_step(u, dt, dx, chs[n] if chs is not None else 0.0)
        return u


def run(u0, dts, dx, gamma, lim="vanleer", riemann="hllc"):
    """2-D hydrodynamics (the Kelvin-Helmholtz setups)."""
    return Scheme(2, "hd", gamma, lim, riemann).run(u0, dts, dx)
