"""Exact solution of the Riemann problem for the 1-D Euler equations of an
ideal gas (E. F. Toro, Riemann Solvers and Numerical Methods for Fluid
Dynamics, 3rd ed., ch. 4: pressure function (4.5)-(4.8), Newton iteration
for p*, sampling (4.52)-(4.66)).  Used as a closed-form check of the
finite-volume schemes."""
from __future__ import annotations

import math

import numpy as np


def _f(p, rho, pk, ck, g):
    if p > pk:  # shock
        A = 2.0 / ((g + 1.0) * rho)
        B = (g - 1.0) / (g + 1.0) * pk
        q = math.sqrt(A / (p + B))
        return (p - pk) * q, q * (1.0 - 0.5 * (p - pk) / (B + p))
    # rarefaction
    r = (p / pk) ** ((g - 1.0) / (2.0 * g))
    f = 2.0 * ck / (g - 1.0) * (r - 1.0)
    df = 1.0 / (rho * ck) * (p / pk) ** (-(g + 1.0) / (2.0 * g))
    return f, df


def star_state(rl, ul, pl, rr, ur, pr, g):
    cl = math.sqrt(g * pl / rl)
    cr = math.sqrt(g * pr / rr)
    p = max(1e-12, 0.5 * (pl + pr) - 0.125 * (ur - ul) * (rl + rr) * (cl + cr))
    for _ in range(100):
        fl, dfl = _f(p, rl, pl, cl, g)
        fr, dfr = _f(p, rr, pr, cr, g)
        pn = p - (fl + fr + ur - ul) / (dfl + dfr)
        pn = max(pn, 1e-14)
        if abs(pn - p) / (0.5 * (pn + p)) < 1e-14:
            p = pn
            break
        p = pn
    fl, _ = _f(p, rl, pl, cl, g)
    fr, _ = _f(p, rr, pr, cr, g)
    u = 0.5 * (ul + ur) + 0.5 * (fr - fl)
    return p, u


def sample(x_over_t, rl, ul, pl, rr, ur, pr, g):
    """Primitive (rho, u, p) at similarity coordinate S = x/t."""
    ps, us = star_state(rl, ul, pl, rr, ur, pr, g)
    cl = math.sqrt(g * pl / rl)
    cr = math.sqrt(g * pr / rr)
    S = x_over_t
    gm = (g - 1.0) / (g + 1.0)
    if S <= us:  # left of contact
        if ps > pl:
            sl = ul - cl * math.sqrt((g + 1.0) / (2.0 * g) * ps / pl + (g - 1.0) / (2.0 * g))
            if S <= sl:
                return rl, ul, pl
            return rl * (ps / pl + gm) / (gm * ps / pl + 1.0), us, ps
        shl = ul - cl
        if S <= shl:
            return rl, ul, pl
        cml = cl * (ps / pl) ** ((g - 1.0) / (2.0 * g))
        stl = us - cml
        if S > stl:
            return rl * (ps / pl) ** (1.0 / g), us, ps
        c = 2.0 / (g + 1.0) * (cl + (g - 1.0) / 2.0 * (ul - S))
        u = 2.0 / (g + 1.0) * (cl + (g - 1.0) / 2.0 * ul + S)
        rho = rl * (c / cl) ** (2.0 / (g - 1.0))
        return rho, u, pl * (c / cl) ** (2.0 * g / (g - 1.0))
    if ps > pr:
        sr = ur + cr * math.sqrt((g + 1.0) / (2.0 * g) * ps / pr + (g - 1.0) / (2.0 * g))
        if S >= sr:
            return rr, ur, pr
        return rr * (ps / pr + gm) / (gm * ps / pr + 1.0), us, ps
    shr = ur + cr
    if S >= shr:
        return rr, ur, pr
    cmr = cr * (ps / pr) ** ((g - 1.0) / (2.0 * g))
    str_ = us + cmr
    if S <= str_:
        return rr * (ps / pr) ** (1.0 / g), us, ps
    c = 2.0 / (g + 1.0) * (cr - (g - 1.0) / 2.0 * (ur - S))
    u = 2.0 / (g + 1.0) * (-cr + (g - 1.0) / 2.0 * ur + S)
    rho = rr * (c / cr) ** (2.0 / (g - 1.0))
    return rho, u, pr * (c / cr) ** (2.0 * g / (g - 1.0))


def solution(x, t, x0, left, right, g):
    """Arrays (rho, u, p) at positions x and time t for a discontinuity at x0."""
    out = np.array([sample((xi - x0) / t, *left, *right, g) for xi in np.asarray(x).ravel()])
    return out[:, 0], out[:, 1], out[:, 2]
