"""Closed-form and invariant checks of the CPU oracle, for the parts of the
scheme no reproducible golden log covers (HLLC, GLM, Einfeldt MHD bounds,
WENO5, the limiters one by one).

Every check is a function ``check_*(lib_path)`` returning a message (empty
when it passes), so that tests/test_oracle_mutants.py can run the same
checks against builds of the oracle with planted mistakes and show that each
mistake is caught.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from amrvac_amd import problems
from amrvac_amd.host import FS, Par, TimeIntegrator
from amrvac_amd.mesh import UniformMesh
from oracle.oracle import OracleFV, weno5_face

import exact_riemann


def _sim(par, mesh=None, lib_path=None):
    mesh = mesh or UniformMesh(par.ndim, (16,) * par.ndim, (16,) * par.ndim, (0.0,) * par.ndim,
                               (1.0,) * par.ndim, ("periodic",) * (2 * par.ndim), nghost=par.nghost)
    return OracleFV(par, mesh, lib_path=lib_path)


def _prims_const(left, right, n=4):
    """n cells, the left half = left, the right half = right."""
    return np.array([left] * (n // 2) + [right] * (n // 2), dtype=np.float64)


# --------------------------------------------------------------------------
# Riemann problem of the Euler equations (Toro ch. 4)
# --------------------------------------------------------------------------
def shock_tube(par: Par, nx: int, t_end: float, left, right, lib_path=None):
    mesh = UniformMesh(2, (nx, 16), (16, 16), (0.0, 0.0), (1.0, 16.0 / nx), ("cont", "cont", "periodic", "periodic"),
                       nghost=par.nghost)
    x = mesh.cell_centers()
    L = np.where(x[:, 0] < 0.5, 1.0, 0.0)
    rho = L * left[0] + (1 - L) * right[0]
    u = L * left[1] + (1 - L) * right[1]
    p = L * left[2] + (1 - L) * right[2]
    w = problems.hd_to_conserved(np.stack([rho, u, 0 * u, p], axis=1), par.gamma, 2)
    be = OracleFV(par, mesh, lib_path=lib_path)
    be.set_state(0, w)
    be.getbc(0)
    import dataclasses
    par2 = dataclasses.replace(par, time_max=t_end)
    TimeIntegrator(par2, mesh, be).run()
    g = mesh.to_global(be.get_state(0))
    xc = (np.arange(nx) + 0.5) / nx
    rho_num = g[0][0]
    rho_ex, _, _ = exact_riemann.solution(xc, t_end, 0.5, left, right, par.gamma)
    return float(np.abs(rho_num - rho_ex).mean()), rho_num


SOD = ((1.0, 0.0, 1.0), (0.125, 0.0, 0.1))

SCHEMES = {
    "3step_hllc_vanleer": dict(time_stepper="threestep", flux_scheme="hllc", limiter="vanleer"),
    "3step_hll_cada3": dict(time_stepper="threestep", flux_scheme="hll", limiter="cada3"),
    "2step_tvdlf_minmod": dict(time_stepper="twostep", flux_scheme="tvdlf", limiter="minmod"),
    "3step_tvdlf_weno5": dict(time_stepper="threestep", flux_scheme="tvdlf", limiter="weno5"),
    "3step_hll_koren_einfeldt": dict(time_stepper="threestep", flux_scheme="hll", limiter="koren"),
}


def check_sod(scheme: str, lib_path=None) -> str:
    par = Par(physics="hd", ndim=2, ndir=2, gamma=1.4, courantpar=0.4, **SCHEMES[scheme])
    e1, _ = shock_tube(par, 128, 0.2, *SOD, lib_path=lib_path)
    e2, _ = shock_tube(par, 256, 0.2, *SOD, lib_path=lib_path)
    if not (e2 < 0.006 and e1 / e2 > 1.6):
        return f"{scheme}: L1(rho) {e1:.4g} (128), {e2:.4g} (256)"
    return ""


@pytest.mark.parametrize("scheme", sorted(SCHEMES))
def test_sod_shock_tube_converges_to_exact_solution(scheme):
    assert check_sod(scheme) == ""


# --------------------------------------------------------------------------
# Smooth advection: order of accuracy
# --------------------------------------------------------------------------
def density_wave_error(par: Par, nx: int, lib_path=None, t_end: float = 1.0) -> float:
    """L1 error of a density wave 1 + 0.2 sin(2 pi x) carried once around a
    periodic unit domain by u = 1 at uniform pressure (an exact solution of
    the Euler equations: pure advection of the contact)."""
    import dataclasses
    mesh = UniformMesh(2, (nx, 16), (16, 16), (0.0, 0.0), (1.0, 16.0 / nx), ("periodic",) * 4, nghost=par.nghost)
    x = mesh.cell_centers()[:, 0]
    dx = 1.0 / nx
    # exact cell averages
    rho = 1.0 + 0.2 * (np.cos(2 * np.pi * (x - dx / 2)) - np.cos(2 * np.pi * (x + dx / 2))) / (2 * np.pi * dx)
    w = problems.hd_to_conserved(np.stack([rho, 0 * rho + 1.0, 0 * rho, 0 * rho + 1.0], axis=1), par.gamma, 2)
    be = OracleFV(par, mesh, lib_path=lib_path)
    be.set_state(0, w)
    be.getbc(0)
    TimeIntegrator(dataclasses.replace(par, time_max=t_end), mesh, be).run()
    return float(np.abs(mesh.to_global(be.get_state(0))[0] - mesh.to_global(w)[0]).mean())


# scheme: (settings, minimum error ratio 64 -> 128 cells, maximum error at 128)
SMOOTH = {
    "twostep_tvdlf_vanleer": (dict(time_stepper="twostep", flux_scheme="tvdlf", limiter="vanleer"), 3.3, 8e-4),
    "threestep_hllc_vanleer": (dict(time_stepper="threestep", flux_scheme="hllc", limiter="vanleer"), 3.3, 6e-4),
    "threestep_hll_cada3": (dict(time_stepper="threestep", flux_scheme="hll", limiter="cada3"), 6.0, 2e-5),
    "threestep_tvdlf_weno5": (dict(time_stepper="threestep", flux_scheme="tvdlf", limiter="weno5"), 16.0, 1.5e-7),
}


def check_smooth_order(scheme: str, lib_path=None) -> str:
    settings, min_ratio, max_err = SMOOTH[scheme]
    par = Par(physics="hd", ndim=2, ndir=2, gamma=1.4, courantpar=0.6, **settings)
    e1 = density_wave_error(par, 64, lib_path)
    e2 = density_wave_error(par, 128, lib_path)
    if not (e2 < max_err and e1 / e2 > min_ratio):
        return f"{scheme}: L1(rho) {e1:.4g} (64), {e2:.4g} (128), ratio {e1 / e2:.3g}"
    return ""


@pytest.mark.parametrize("scheme", sorted(SMOOTH))
def test_smooth_density_wave_order_of_accuracy(scheme):
    """Second order for the TVD limiters (Hancock predictor-corrector and
    SSPRK3), third order for cada3, fifth for WENO5 in space."""
    assert check_smooth_order(scheme) == ""


def check_hllc_stationary_contact(lib_path=None) -> str:
    """HLLC keeps an isolated stationary contact exactly; HLL smears it."""
    out = {}
    for fs in ("hllc", "hll"):
        par = Par(physics="hd", ndim=2, ndir=2, gamma=1.4, time_stepper="threestep", flux_scheme=fs,
                  limiter="minmod", courantpar=0.5)
        mesh = UniformMesh(2, (64, 16), (16, 16), (0.0, 0.0), (1.0, 0.25), ("cont", "cont", "periodic", "periodic"))
        x = mesh.cell_centers()
        rho = np.where(x[:, 0] < 0.5, 1.0, 0.1)
        w = problems.hd_to_conserved(np.stack([rho, 0 * rho, 0 * rho, 0 * rho + 1.0], axis=1), 1.4, 2)
        be = OracleFV(par, mesh, lib_path=lib_path)
        be.set_state(0, w)
        be.getbc(0)
        TimeIntegrator(par, mesh, be).run(nsteps=20)
        out[fs] = float(np.abs(mesh.interior(be.get_state(0))[:, 0] - mesh.interior(w)[:, 0]).max())
    if not (out["hllc"] < 1e-13 and out["hll"] > 1e-3):
        return f"contact: hllc {out['hllc']:.3g}, hll {out['hll']:.3g}"
    return ""


def test_hllc_keeps_a_stationary_contact():
    assert check_hllc_stationary_contact() == ""


# --------------------------------------------------------------------------
# Single-face identities
# --------------------------------------------------------------------------
def euler_flux(prim, gamma, idim, ndir):
    """Textbook flux of the Euler equations in direction idim."""
    rho, v, p = prim[0], prim[1:1 + ndir], prim[1 + ndir]
    e = p / (gamma - 1.0) + 0.5 * rho * np.dot(v, v)
    f = np.zeros(ndir + 2)
    f[0] = rho * v[idim]
    f[1:1 + ndir] = rho * v[idim] * v
    f[1 + idim] += p
    f[1 + ndir] = (e + p) * v[idim]
    return f


def mhd_flux(prim, gamma, idim, ndir=3):
    """Textbook flux of ideal MHD (no GLM) in direction idim."""
    rho, v, p, B = prim[0], prim[1:4], prim[4], prim[5:8]
    e = p / (gamma - 1.0) + 0.5 * rho * np.dot(v, v) + 0.5 * np.dot(B, B)
    pt = p + 0.5 * np.dot(B, B)
    f = np.zeros(8)
    f[0] = rho * v[idim]
    f[1:4] = rho * v[idim] * v - B[idim] * B
    f[1 + idim] += pt
    f[4] = (e + pt) * v[idim] - B[idim] * np.dot(B, v)
    f[5:8] = v[idim] * B - B[idim] * v
    return f


def check_flux_consistency(lib_path=None) -> str:
    """F_Riemann(u, u) = f(u) for every solver, HD and MHD, all directions."""
    rng = np.random.default_rng(7)
    msgs = []
    hd = _sim(Par(physics="hd", ndim=3, ndir=3, gamma=1.4, limiter="minmod"), lib_path=lib_path)
    mhd = _sim(Par(physics="mhd", ndim=3, ndir=3, gamma=5 / 3, typedivbfix="none", limiter="minmod"),
               lib_path=lib_path)
    for _ in range(20):
        pr = np.concatenate([[0.5 + rng.random()], rng.normal(size=3), [0.2 + rng.random()]])
        for fs in ("tvdlf", "hll", "hllc"):
            for d in range(3):
                f, _, _ = hd.face_flux(FS[fs], d, _prims_const(pr, pr))
                ref = euler_flux(pr, 1.4, d, 3)
                if not np.allclose(f, ref, rtol=1e-12, atol=1e-12):
                    msgs.append(f"hd {fs} dim {d}: {f} vs {ref}")
        pm = np.concatenate([[0.5 + rng.random()], rng.normal(size=3), [0.2 + rng.random()], rng.normal(size=3)])
        for fs in ("tvdlf", "hll", "hllc", "hlld"):
            for d in range(3):
                f, _, _ = mhd.face_flux(FS[fs], d, _prims_const(pm, pm))
                ref = mhd_flux(pm, 5 / 3, d)
                if not np.allclose(f, ref, rtol=1e-12, atol=1e-12):
                    msgs.append(f"mhd {fs} dim {d}: {f} vs {ref}")
    return "; ".join(msgs[:3])


def test_riemann_fluxes_are_consistent():
    assert check_flux_consistency() == ""


def check_signal_speeds(lib_path=None) -> str:
    """Einfeldt bounds of identical states are u -+ c (Roe average of equal
    states); the MHD fast speed is max(c_s, c_A) for B along the normal and
    sqrt(c_s^2 + c_A^2) for B across it."""
    msgs = []
    g = 1.4
    hd = _sim(Par(physics="hd", ndim=2, ndir=2, gamma=g, limiter="minmod", typeboundspeed="Einfeldt"),
              lib_path=lib_path)
    pr = np.array([0.7, 0.3, -0.2, 1.3])
    _, cmax, cmin = hd.face_flux(FS["hll"], 0, _prims_const(pr, pr))
    c = math.sqrt(g * 1.3 / 0.7)
    if not (abs(cmax - (0.3 + c)) < 1e-14 and abs(cmin - (0.3 - c)) < 1e-14):
        msgs.append(f"hd Einfeldt {cmin},{cmax} vs {0.3 - c},{0.3 + c}")
    gm = 5 / 3
    mhd = _sim(Par(physics="mhd", ndim=3, ndir=3, gamma=gm, typedivbfix="none", limiter="minmod",
                   typeboundspeed="Einfeldt"), lib_path=lib_path)
    rho, p = 0.9, 0.4
    cs = math.sqrt(gm * p / rho)
    for B, expect in (([1.2, 0, 0], max(cs, 1.2 / math.sqrt(rho))),
                      ([0, 0.8, 0.6], math.sqrt(cs ** 2 + 1.0 / rho))):
        pm = np.array([rho, 0.0, 0.0, 0.0, p] + B)
        _, cmax, cmin = mhd.face_flux(FS["hll"], 0, _prims_const(pm, pm))
        if not (abs(cmax - expect) < 1e-12 and abs(cmin + expect) < 1e-12):
            msgs.append(f"mhd fast speed B={B}: {cmin},{cmax} vs +-{expect}")
        cons = problems.mhd_to_conserved(pm[None, :], gm, 3)[0]
        cm = mhd.cell_cmax(cons, 0)
        if abs(cm - expect) > 1e-12:
            msgs.append(f"mhd cell cmax B={B}: {cm} vs {expect}")
    return "; ".join(msgs)


def test_signal_speeds_special_cases():
    assert check_signal_speeds() == ""


def check_glm_interface_state(lib_path=None) -> str:
    """mhd_modify_wLR solves the linear (B_n, psi) Riemann problem: the face
    values carry the characteristic variables psi + c_h B_n from the left and
    psi - c_h B_n from the right (Dedner et al. 2002, eq. 42).  The face
    fluxes of B_n and psi are psi* and c_h^2 B_n*."""
    gm = 5 / 3
    ch = 2.3
    sim = _sim(Par(physics="mhd", ndim=3, ndir=3, gamma=gm, typedivbfix="glm", limiter="minmod"),
               lib_path=lib_path)
    msgs = []
    rng = np.random.default_rng(3)
    for d in range(3):
        pl = np.concatenate([[1.0, 0.1, -0.2, 0.05, 0.6], rng.normal(size=3), [rng.normal()]])
        pr = np.concatenate([[0.8, 0.0, 0.1, -0.1, 0.5], rng.normal(size=3), [rng.normal()]])
        f, _, _ = sim.face_flux(FS["hll"], d, _prims_const(pl, pr), ch=ch)
        psi_s, bn_s = f[5 + d], f[8] / ch ** 2
        if abs((psi_s + ch * bn_s) - (pl[8] + ch * pl[5 + d])) > 1e-12:
            msgs.append(f"dim {d}: right-going invariant")
        if abs((psi_s - ch * bn_s) - (pr[8] - ch * pr[5 + d])) > 1e-12:
            msgs.append(f"dim {d}: left-going invariant")
    return "; ".join(msgs)


def test_glm_interface_riemann_solution():
    assert check_glm_interface_state() == ""


def check_limiters(lib_path=None) -> str:
    """dwlimiter2 identities: minmod/vanleer/koren/cada3 vanish for opposite
    signs and are symmetric in scaling; minmod picks the smaller slope;
    vanleer is the harmonic mean; cada3 in smooth regions (eta < 1) is the
    third-order slope (2+theta)/3 dwC(i)."""
    msgs = []
    base = dict(physics="hd", ndim=2, ndir=2, gamma=1.4)
    sims = {lim: _sim(Par(limiter=lim, **base), lib_path=lib_path)
            for lim in ("minmod", "vanleer", "koren", "cada", "cada3", "woodward", "superbee", "mcbeta")}
    rng = np.random.default_rng(11)
    for _ in range(200):
        a, b = rng.normal(size=2)
        for lim, s in sims.items():
            l, r = s.limiter(a, b)
            if a * b < 0 and lim not in ("cada", "cada3") and (l != 0.0 or r != 0.0):
                msgs.append(f"{lim} not zero for opposite signs")
        l, r = sims["minmod"].limiter(a, b)
        if a * b > 0 and l != math.copysign(min(abs(a), abs(b)), b):
            msgs.append("minmod value")
        l, _ = sims["vanleer"].limiter(a, b)
        if a * b > 0 and abs(l - 2 * a * b / (a + b)) > 1e-9 * abs(l):
            msgs.append("vanleer value")
        # TVD bounds: 0 <= phi <= 2 min(1, theta)
        for lim in ("minmod", "vanleer", "koren", "superbee", "mcbeta", "woodward"):
            l, _ = sims[lim].limiter(a, b)
            if a * b > 0 and not (0 <= l / b <= 2 * min(1.0, a / b) + 1e-12):
                msgs.append(f"{lim} outside TVD region")
    # cada3 (Cada & Torrilhon 2009): where the local variation is small
    # against cada3_radius*dx (eta < 1) the slope is the third-order one,
    # (2 dwC(i+1) + dwC(i))/3 to the left face and (2 dwC(i) + dwC(i+1))/3 to
    # the right (up to the sqrt(eps) guard in theta); with a=b large it is b.
    s = sims["cada3"]
    for a, b in ((1e-3, 2e-3), (-3e-3, 1e-3), (4e-3, -1e-3), (2e-3, 2e-3)):
        l, r = s.limiter(a, b)
        if abs(l - (2 * b + a) / 3) > 1e-4 * abs(b) or abs(r - (2 * a + b) / 3) > 1e-4 * abs(a):
            msgs.append(f"cada3 smooth {a},{b}: {l},{r}")
    l, r = s.limiter(0.5, 0.5)
    if abs(l - 0.5) > 1e-7 or abs(r - 0.5) > 1e-7:
        msgs.append(f"cada3 equal slopes {l},{r}")
    l, r = s.limiter(0.5, 0.05)
    # theta=10 in the limited region: min(A, max(.., min(2 theta, A, 1.6))) = 1.6
    if abs(l - 1.6 * 0.05) > 1e-9:
        msgs.append(f"cada3 limited {l}")
    # cada is cada3's limited branch everywhere: the same value as cada3 where
    # eta > 1, and its slope ratio stays within [0, 1.6] (cadgamma)
    c, c3 = sims["cada"], sims["cada3"]
    for _ in range(200):
        a, b = rng.normal(size=2)
        l, r = c.limiter(a, b)
        l3, r3 = c3.limiter(a, b)
        if (a * a + b * b) > 1.0 and (l != l3 or r != r3):  # eta >> 1 at dx = 1/16
            msgs.append(f"cada vs cada3 limited {a},{b}")
        if a * b > 0 and not (0.0 <= l / b <= 1.6 * (1 + 1e-12) and 0.0 <= r / a <= 1.6 * (1 + 1e-12)):
            msgs.append(f"cada slope ratio outside [0, 1.6] at {a},{b}")
    l, r = c.limiter(0.5, 0.5)
    if abs(l - 0.5) > 1e-7 or abs(r - 0.5) > 1e-7:
        msgs.append(f"cada equal slopes {l},{r}")
    return "; ".join(msgs[:4])


def test_limiter_identities():
    assert check_limiters() == ""


def check_weno5_polynomials(lib_path=None) -> str:
    """WENO5 face values are exact for cell averages of polynomials of degree
    <= 2, whatever the nonlinear weights."""
    rng = np.random.default_rng(5)
    msgs = []
    for _ in range(50):
        c = rng.normal(size=3)
        h = 0.1
        # cell averages over [x_j - h/2, x_j + h/2] of c0 + c1 x + c2 x^2, x_j = (j-2) h
        xs = (np.arange(6) - 2) * h
        avg = c[0] + c[1] * xs + c[2] * (xs ** 2 + h * h / 12.0)
        xf = 0.5 * h
        exact = c[0] + c[1] * xf + c[2] * xf * xf
        wl, wr = weno5_face(avg, lib_path=lib_path)
        if abs(wl - exact) > 1e-12 or abs(wr - exact) > 1e-12:
            msgs.append(f"weno5 {wl} {wr} vs {exact}")
    return "; ".join(msgs[:2])


def test_weno5_exact_for_quadratics():
    assert check_weno5_polynomials() == ""


def check_fix_limiter(lib_path=None) -> str:
    """fix_limiter (src/limiter/mod_weno.t:100-133): where the WENO5 left or
    right face state fails the small-value check, both fall back to the
    first-order states, the two cells next to the face.  With rho=1 and v=0
    the TVDLF momentum flux is (p_L + p_R)/2 exactly, so the face states can
    be read off the flux."""
    msgs = []
    gamma = 5.0 / 3.0

    def mom_flux(p_line, small_pressure, v=0.0, rho_line=None):
        par = Par(physics="hd", ndim=2, ndir=2, gamma=gamma, limiter="weno5", small_pressure=small_pressure)
        sim = _sim(par, lib_path=lib_path)
        n = len(p_line)
        rho = np.ones(n) if rho_line is None else np.asarray(rho_line)
        prims = np.stack([rho, v * np.ones(n), np.zeros(n), np.asarray(p_line)], axis=1)
        f, _, _ = sim.face_flux(FS["tvdlf"], 0, prims)
        return f
    # a pressure minimum on the face: cell averages of p0 + (x - x_f)^2
    # (h = 1), whose WENO5 face value is p0 = 0.1 exactly; every cell holds
    # at least 0.1 + 1/4 + 1/12
    xs = np.arange(6) - 2.5
    p_avg = 0.1 + xs ** 2 + 1.0 / 12.0
    f_weno = mom_flux(p_avg, 0.05)[1]
    f_fixed = mom_flux(p_avg, 0.2)[1]
    if abs(f_weno - 0.1) > 1e-12:
        msgs.append(f"unflagged WENO5 face pressure {f_weno} != 0.1")
    if abs(f_fixed - p_avg[2]) > 1e-12:
        msgs.append(f"flagged face: momentum flux {f_fixed} != cell pressure {p_avg[2]}")
    # the same for density: rho minimum on the face, v = 1 (mass flux rho_face)
    g_weno = mom_flux(np.ones(6), 0.0, v=1.0, rho_line=p_avg)[0]
    par = Par(physics="hd", ndim=2, ndir=2, gamma=gamma, limiter="weno5", small_density=0.2)
    sim = _sim(par, lib_path=lib_path)
    prims = np.stack([p_avg, np.ones(6), np.zeros(6), np.ones(6)], axis=1)
    g_fixed = sim.face_flux(FS["tvdlf"], 0, prims)[0][0]
    if abs(g_weno - 0.1) > 1e-12 or abs(g_fixed - p_avg[2]) > 1e-12:
        msgs.append(f"density fallback: mass flux {g_weno} / {g_fixed}")
    # only the right state fails (its WENO5 value is 0.297, the left one
    # 1.777, every cell >= 0.6): both still fall back
    p_line = [2.4, 0.75, 2.2, 0.6, 1.25, 1.35]
    wl, wr = weno5_face(p_line, lib_path=lib_path)
    f_one = mom_flux(p_line, 0.5)[1]
    if not (wr < 0.5 < wl and min(p_line) > 0.5):
        msgs.append(f"right-only case lost: {wl} {wr}")
    if abs(f_one - 0.5 * (2.2 + 0.6)) > 1e-12:
        msgs.append(f"right-only failure: momentum flux {f_one} != {0.5 * (2.2 + 0.6)}")
    return "; ".join(msgs)


def test_weno5_fix_limiter_falls_back_to_first_order():
    assert check_fix_limiter() == ""


# --------------------------------------------------------------------------
# Whole-step invariants
# --------------------------------------------------------------------------
def check_glm_decay(lib_path=None) -> str:
    """A uniform MHD state with uniform psi: fluxes cancel and the only change
    is the implicit psi decay exp(-qdt c_h alpha / dx) of add_source_glm,
    combined by ssprk3."""
    par = Par(physics="mhd", ndim=3, ndir=3, gamma=5 / 3, typedivbfix="glm", time_stepper="threestep",
              flux_scheme="hll", limiter="cada3", courantpar=0.6)
    mesh = UniformMesh(3, (16, 16, 16), (16, 16, 16), (0.0,) * 3, (1.0,) * 3, ("periodic",) * 6)
    prim = np.array([1.0, 0.3, -0.1, 0.2, 0.5, 0.4, 0.2, -0.3, 0.01])
    w = problems.mhd_to_conserved(np.broadcast_to(prim[None, :, None, None, None],
                                                  (1, 9) + mesh.nxg[::-1]).copy(), 5 / 3, 3)
    be = OracleFV(par, mesh, lib_path=lib_path)
    be.set_state(0, w)
    be.getbc(0)
    ti = TimeIntegrator(par, mesh, be)
    ti.step()
    dt, ch = ti.state.dt, ti.state.cmax_global
    alpha, dx = 0.5, 1.0 / 16
    d = [math.exp(-b * dt * ch * alpha / dx) for b in (1.0, 0.25, 2.0 / 3.0)]
    psi1 = d[0] * 0.01
    psi2 = d[1] * (0.75 * 0.01 + (1.0 - 0.75) * psi1)
    psi3 = d[2] * (1.0 / 3.0 * 0.01 + (1.0 - 1.0 / 3.0) * psi2)
    got = mesh.interior(be.get_state(0))
    msgs = []
    if abs(got[0, 8].max() - psi3) > 1e-15 or abs(got[0, 8].min() - psi3) > 1e-15:
        msgs.append(f"psi {got[0, 8].max()} vs {psi3}")
    if np.abs(got[0, :8] - mesh.interior(w)[0, :8]).max() > 1e-14:
        msgs.append("non-psi variables changed")
    return "; ".join(msgs)


def test_glm_psi_decay_closed_form():
    assert check_glm_decay() == ""


def check_conservation_and_uniform_states(lib_path=None) -> str:
    msgs = []
    # periodic KH: mass, momentum and energy totals to round-off
    par = Par(physics="hd", ndim=2, ndir=2, gamma=1.4, time_stepper="threestep", flux_scheme="hllc",
              limiter="vanleer", courantpar=0.8)
    mesh = UniformMesh(2, (64, 64), (16, 16), (0.0, 0.0), (1.0, 1.0), ("periodic",) * 4)
    w0 = problems.kelvin_helmholtz_2d(mesh)
    be = OracleFV(par, mesh, lib_path=lib_path)
    be.set_state(0, w0)
    be.getbc(0)
    TimeIntegrator(par, mesh, be).run(nsteps=10)
    t0 = mesh.interior(w0).sum(axis=(0, 2, 3))
    t1 = mesh.interior(be.get_state(0)).sum(axis=(0, 2, 3))
    if np.abs(t1 - t0).max() > 1e-10 * np.abs(t0).max():
        msgs.append(f"HD totals drift {t1 - t0}")
    # periodic OT + GLM: mass to round-off
    par = Par(physics="mhd", ndim=3, ndir=3, typedivbfix="glm", time_stepper="threestep", flux_scheme="hll",
              limiter="cada3", courantpar=0.6)
    mesh = UniformMesh(3, (32, 32, 32), (16, 16, 16), (0.0,) * 3, (1.0,) * 3, ("periodic",) * 6)
    w0 = problems.orszag_tang_3d(mesh)
    be = OracleFV(par, mesh, lib_path=lib_path)
    be.set_state(0, w0)
    be.getbc(0)
    TimeIntegrator(par, mesh, be).run(nsteps=2)
    m0 = mesh.interior(w0)[:, 0].sum()
    m1 = mesh.interior(be.get_state(0))[:, 0].sum()
    if abs(m1 - m0) > 1e-12 * m0:
        msgs.append(f"MHD mass drift {m1 - m0}")
    # uniform moving states stay uniform
    for fs, lim, phys in (("hllc", "vanleer", "hd"), ("hll", "cada3", "mhd"), ("tvdlf", "weno5", "hd")):
        if phys == "hd":
            par = Par(physics="hd", ndim=2, ndir=2, gamma=1.4, time_stepper="threestep", flux_scheme=fs,
                      limiter=lim, courantpar=0.5)
            mesh = UniformMesh(2, (32, 32), (16, 16), (0.0, 0.0), (1.0, 1.0), ("periodic",) * 4, nghost=par.nghost)
            prim = np.array([1.3, 0.4, -0.7, 2.0])
            w0 = problems.hd_to_conserved(np.broadcast_to(prim[None, :, None, None], (mesh.nblocks, 4) + mesh.nxg[::-1]).copy(), 1.4, 2)
        else:
            par = Par(physics="mhd", ndim=2, ndir=3, typedivbfix="linde", time_stepper="threestep", flux_scheme=fs,
                      limiter=lim, courantpar=0.5)
            mesh = UniformMesh(2, (32, 32), (16, 16), (0.0, 0.0), (1.0, 1.0), ("periodic",) * 4)
            prim = np.array([1.3, 0.4, -0.7, 0.1, 2.0, 0.3, 0.5, -0.2])
            w0 = problems.mhd_to_conserved(np.broadcast_to(prim[None, :, None, None], (mesh.nblocks, 8) + mesh.nxg[::-1]).copy(), 5 / 3, 3)
        be = OracleFV(par, mesh, lib_path=lib_path)
        be.set_state(0, w0)
        be.getbc(0)
        TimeIntegrator(par, mesh, be).run(nsteps=3)
        dev = np.abs(mesh.interior(be.get_state(0)) - mesh.interior(w0)).max()
        if dev > 1e-13:
            msgs.append(f"uniform {phys} {fs} {lim} changed by {dev:.3g}")
    return "; ".join(msgs)


def test_conservation_and_uniform_states():
    assert check_conservation_and_uniform_states() == ""


def check_hlld_discontinuities(lib_path=None) -> str:
    """HLLD (Miyoshi & Kusano 2005) resolves isolated discontinuities
    exactly: across a stationary contact (only rho jumps, B_n != 0) and a
    stationary rotational discontinuity (v_n = -s B_n/sqrt(rho), B_t
    rotated at constant |B_t|, [v_t] = B_n [B_t]/(rho v_n)) the face flux
    is f(u_L) = f(u_R); HLL is not.  With B = 0 it is the hydrodynamic HLLC
    flux (same contact speed and star states) for subsonic states."""
    g = 5.0 / 3.0
    rng = np.random.default_rng(11)
    mhd = _sim(Par(physics="mhd", ndim=3, ndir=3, gamma=g, typedivbfix="none", limiter="minmod"), lib_path=lib_path)
    hd = _sim(Par(physics="hd", ndim=3, ndir=3, gamma=g, limiter="minmod"), lib_path=lib_path)
    msgs = []

    def rel(a, b):
        return float(np.abs(a - b).max() / np.abs(b).max())

    for _ in range(10):
        for d in range(3):
            v = np.zeros(3)
            v[[(d + 1) % 3, (d + 2) % 3]] = rng.normal(size=2)
            B = rng.normal(size=3)
            B[d] = 0.5 + rng.random()
            pl = np.concatenate([[0.5 + rng.random()], v, [0.2 + rng.random()], B])
            pr = pl.copy()
            pr[0] = 0.5 + 2.0 * rng.random()
            f, _, _ = mhd.face_flux(FS["hlld"], d, _prims_const(pl, pr))
            fh, _, _ = mhd.face_flux(FS["hll"], d, _prims_const(pl, pr))
            ref = mhd_flux(pl, g, d)
            if rel(f, ref) > 1e-13 or rel(fh, ref) < 1e-3:
                msgs.append(f"contact dim {d}: hlld {rel(f, ref):.2e} hll {rel(fh, ref):.2e}")
            rho, p, bn, sgn = 0.5 + rng.random(), 0.2 + rng.random(), 0.5 + rng.random(), rng.choice([-1.0, 1.0])
            vn = -sgn * bn / math.sqrt(rho)
            c = bn / (rho * vn)
            bt, th = 0.3 + rng.random(), rng.random(2) * 2.0 * math.pi
            w0 = rng.normal(size=2)
            prims = []
            for t in th:
                btv = bt * np.array([math.cos(t), math.sin(t)])
                vv, BB = np.zeros(3), np.zeros(3)
                vv[d], BB[d] = vn, bn
                vv[[(d + 1) % 3, (d + 2) % 3]] = c * btv + w0
                BB[[(d + 1) % 3, (d + 2) % 3]] = btv
                prims.append(np.concatenate([[rho], vv, [p], BB]))
            f, _, _ = mhd.face_flux(FS["hlld"], d, _prims_const(prims[0], prims[1]))
            fh, _, _ = mhd.face_flux(FS["hll"], d, _prims_const(prims[0], prims[1]))
            ref = mhd_flux(prims[0], g, d)
            if rel(f, ref) > 1e-13 or rel(fh, ref) < 1e-3:
                msgs.append(f"rotational dim {d}: hlld {rel(f, ref):.2e} hll {rel(fh, ref):.2e}")
            pl = np.concatenate([[0.5 + rng.random()], 0.3 * rng.normal(size=3), [0.5 + rng.random()]])
            pr = np.concatenate([[0.5 + rng.random()], 0.3 * rng.normal(size=3), [0.5 + rng.random()]])
            fc, _, _ = hd.face_flux(FS["hllc"], d, _prims_const(pl, pr))
            fm, _, _ = mhd.face_flux(FS["hlld"], d, _prims_const(np.concatenate([pl, np.zeros(3)]),
                                                                 np.concatenate([pr, np.zeros(3)])))
            if rel(fm[:5], fc) > 1e-13 or np.abs(fm[5:]).max() != 0.0:
                msgs.append(f"B=0 dim {d}: hlld vs hllc {rel(fm[:5], fc):.2e}")
    return "; ".join(msgs[:3])


def mk_hlld_flux(pl, pr, sL, sR, gamma, d):
    """HLLD flux of Miyoshi & Kusano (2005, J. Comput. Phys. 208, 315),
    written from the paper in its own variables for the face normal d:
    S_M by eq. (38), the total pressure of the star states by eq. (41)
    (the paper's closed form, not the p_T* = rho(S-u)(S_M-u) + p_T
    shortcut), rho*, v*, B*, e* by eqs. (43)-(48), the double-star states
    by eqs. (59)-(63) and the fluxes by eq. (66).  pl, pr: primitive
    (rho, v1..3, p, B1..3); sL, sR: the outer wave speeds."""
    t = [(d + 1) % 3, (d + 2) % 3]

    def unpack(pw):
        rho, v, p, B = pw[0], pw[1:4].copy(), pw[4], pw[5:8].copy()
        e = p / (gamma - 1.0) + 0.5 * rho * (v @ v) + 0.5 * (B @ B)
        return rho, v, p, B, e, p + 0.5 * (B @ B)

    rL, vL, _, BL, eL, ptL = unpack(pl)
    rR, vR, _, BR, eR, ptR = unpack(pr)
    Bx = (sR * BR[d] - sL * BL[d]) / (sR - sL)
    uL, uR = vL[d], vR[d]
    sM = ((sR - uR) * rR * uR - (sL - uL) * rL * uL - ptR + ptL) / ((sR - uR) * rR - (sL - uL) * rL)
    pts = ((sR - uR) * rR * ptL - (sL - uL) * rL * ptR + rL * rR * (sR - uR) * (sL - uL) * (uR - uL)) / (
        (sR - uR) * rR - (sL - uL) * rL)

    def star(r, v, B, e, pt, S, u):
        rs = r * (S - u) / (S - sM)
        den = r * (S - u) * (S - sM) - Bx * Bx
        vs, Bs = v.copy(), B.copy()
        vs[d], Bs[d] = sM, Bx
        if den != 0.0:
            vs[t] = v[t] - Bx * B[t] * (sM - u) / den
            Bs[t] = B[t] * (r * (S - u) ** 2 - Bx * Bx) / den
        else:
            vs[t], Bs[t] = v[t], B[t]
        es = ((S - u) * e - pt * u + pts * sM + Bx * (v @ B - vs @ Bs)) / (S - sM)
        return rs, vs, Bs, es

    rsL, vsL, BsL, esL = star(rL, vL, BL, eL, ptL, sL, uL)
    rsR, vsR, BsR, esR = star(rR, vR, BR, eR, ptR, sR, uR)
    qL, qR = math.sqrt(rsL), math.sqrt(rsR)
    sgn = math.copysign(1.0, Bx)
    vss, Bss = vsL.copy(), BsL.copy()
    vss[t] = (qL * vsL[t] + qR * vsR[t] + (BsR[t] - BsL[t]) * sgn) / (qL + qR)
    Bss[t] = (qL * BsR[t] + qR * BsL[t] + qL * qR * (vsR[t] - vsL[t]) * sgn) / (qL + qR)
    essL = esL - qL * (vsL @ BsL - vss @ Bss) * sgn
    essR = esR + qR * (vsR @ BsR - vss @ Bss) * sgn
    s1L, s1R = sM - abs(Bx) / qL, sM + abs(Bx) / qR

    def cons(r, v, e, B):
        return np.concatenate([[r], r * v, [e], B])

    UL, UR = cons(rL, vL, eL, BL), cons(rR, vR, eR, BR)
    FL, FR = mhd_flux(pl, gamma, d), mhd_flux(pr, gamma, d)
    U1L, U1R = cons(rsL, vsL, esL, BsL), cons(rsR, vsR, esR, BsR)
    U2L, U2R = cons(rsL, vss, essL, Bss), cons(rsR, vss, essR, Bss)
    if sL > 0.0:
        return FL
    if s1L >= 0.0:
        return FL + sL * (U1L - UL)
    if sM >= 0.0:
        return FL + sL * (U1L - UL) + s1L * (U2L - U1L)
    if s1R >= 0.0:
        return FR + sR * (U1R - UR) + s1R * (U2R - U1R)
    if sR >= 0.0:
        return FR + sR * (U1R - UR)
    return FR


def check_hlld_against_paper(lib_path=None) -> str:
    """The oracle's HLLD flux (the reference's transcription, with Guo's
    total-pressure shortcut) against the paper's formulation above, for
    random states in every fan region (normal B of both signs)."""
    g = 5.0 / 3.0
    rng = np.random.default_rng(5)
    mhd = _sim(Par(physics="mhd", ndim=3, ndir=3, gamma=g, typedivbfix="none", limiter="minmod"), lib_path=lib_path)
    worst = 0.0
    for _ in range(200):
        pl = np.concatenate([[0.3 + rng.random()], rng.normal(size=3), [0.1 + rng.random()], rng.normal(size=3)])
        pr = np.concatenate([[0.3 + rng.random()], rng.normal(size=3), [0.1 + rng.random()], rng.normal(size=3)])
        d = int(rng.integers(3))
        f, cmax, cmin = mhd.face_flux(FS["hlld"], d, _prims_const(pl, pr))
        ref = mk_hlld_flux(pl, pr, cmin, cmax, g, d)
        worst = max(worst, float(np.abs(f - ref).max() / np.abs(ref).max()))
    if worst > 1e-11:
        return f"hlld vs the paper's formulation: {worst:.2e}"
    return ""


def test_hlld_matches_the_published_formulation():
    assert check_hlld_against_paper() == ""


def check_mhd_hllc_contact(lib_path=None) -> str:
    """MHD HLLC (Li 2005, mhd_get_wCD) keeps an isolated stationary contact
    exactly, with and without a normal field (HLL does not), and reduces to
    the hydrodynamic HLLC flux at B = 0 for subsonic states."""
    g = 5.0 / 3.0
    rng = np.random.default_rng(13)
    mhd = _sim(Par(physics="mhd", ndim=3, ndir=3, gamma=g, typedivbfix="none", limiter="minmod"), lib_path=lib_path)
    hd = _sim(Par(physics="hd", ndim=3, ndir=3, gamma=g, limiter="minmod"), lib_path=lib_path)
    msgs = []
    for _ in range(10):
        for d in range(3):
            for bn in (0.0, 0.5 + rng.random()):
                v = np.zeros(3)
                v[[(d + 1) % 3, (d + 2) % 3]] = rng.normal(size=2)
                B = rng.normal(size=3)
                B[d] = bn
                pl = np.concatenate([[0.5 + rng.random()], v, [0.2 + rng.random()], B])
                pr = pl.copy()
                pr[0] = 0.5 + 2.0 * rng.random()
                f, _, _ = mhd.face_flux(FS["hllc"], d, _prims_const(pl, pr))
                ref = mhd_flux(pl, g, d)
                if np.abs(f - ref).max() > 1e-13 * np.abs(ref).max():
                    msgs.append(f"contact dim {d} bn {bn:.2f}: {np.abs(f - ref).max():.2e}")
            pl = np.concatenate([[0.5 + rng.random()], 0.3 * rng.normal(size=3), [0.5 + rng.random()]])
            pr = np.concatenate([[0.5 + rng.random()], 0.3 * rng.normal(size=3), [0.5 + rng.random()]])
            fc, _, _ = hd.face_flux(FS["hllc"], d, _prims_const(pl, pr))
            fm, _, _ = mhd.face_flux(FS["hllc"], d, _prims_const(np.concatenate([pl, np.zeros(3)]),
                                                                 np.concatenate([pr, np.zeros(3)])))
            if np.abs(fm[:5] - fc).max() > 1e-13 * np.abs(fc).max():
                msgs.append(f"B=0 dim {d}: {np.abs(fm[:5] - fc).max():.2e}")
    return "; ".join(msgs[:3])


def test_mhd_hllc_keeps_contacts():
    assert check_mhd_hllc_contact() == ""


def test_hlld_resolves_isolated_discontinuities():
    assert check_hlld_discontinuities() == ""


# --------------------------------------------------------------------------
# constrained transport (staggered face field)
# --------------------------------------------------------------------------

CT_SCHEMES = ("3step_hll_cada_ct", "3step_hll_cada_ct_average", "3step_hll_cada_ct_hll", "3step_hlld_cada_ct")


def _ct_field_loop(scheme: str, lib_path=None, rows: int = 50):
    from test_oracle_golden import FieldLoopIntegrator, field_loop_setup
    par, mesh = field_loop_setup(scheme)
    par.time_max = 0.01 * rows
    be = OracleFV(par, mesh, lib_path=lib_path)
    be.set_state(0, problems.field_loop_ct(mesh))
    be.getbc(0)
    st = FieldLoopIntegrator(par, mesh, be).run()
    return par, mesh, be, np.array(st.log)


def check_ct_golden(scheme: str, lib_path=None) -> str:
    """The staggered field loop (tests/mhd/field_loop_2.5D with the scheme
    file) reproduces the reference's regression log, all 50 rows."""
    import os
    from amrvac_amd.host import logs_match, read_log
    _, _, _, log = _ct_field_loop(scheme, lib_path)
    gold = read_log(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", f"fl_2d_{scheme}.log"))
    if log.shape != gold.shape or not logs_match(log, gold):
        return f"{scheme}: log differs from the reference's"
    return ""


def check_ct_divb(lib_path=None) -> str:
    """Constrained transport keeps the face-field divergence at round-off
    for every electric-field average (the defining property: the face
    update is the circulation of one edge field, so the discrete
    divergence of every cell telescopes to zero)."""
    msgs = []
    for scheme in CT_SCHEMES[:3]:
        par, mesh, be, _ = _ct_field_loop(scheme, lib_path, rows=10)
        w = be.get_state(0)
        ex = problems.field_loop_output_vars_ct(mesh, w)
        divb = np.abs(mesh.interior(ex)[:, 0]).max()
        bscale = np.abs(mesh.interior(w)[:, 5:7]).max() / min(mesh.dx)
        if not divb <= 1e-12 * bscale:
            msgs.append(f"{scheme}: max|div B| {divb:.3e} vs |B|/dx {bscale:.3e}")
    return "; ".join(msgs)


def test_ct_keeps_div_b_zero():
    assert check_ct_divb() == ""


def check_hancock_ct(lib_path=None) -> str:
    """The Hancock predictor with constrained transport (twostep with
    TVDLF, typepred1 = 'hancock', src/io/mod_input_output.t:788-810).
    hancock (src/mod_finite_volume.t:19-113) neither reads nor updates the
    face field, and advect copies ps into ps1 before the predictor
    (src/mod_advance.t:69-74).  So after the predictor ps1's faces equal
    ps's bit for bit, and ps1's cells equal those of the same predictor run
    without a face field (typedivbfix 'none') on the same cells, bit for
    bit.  The corrector then keeps div B at round-off."""
    import dataclasses
    from amrvac_amd.host import FS, TimeIntegrator
    from test_oracle_golden import field_loop_setup
    par, mesh = field_loop_setup("3step_hll_cada_ct")
    par = dataclasses.replace(par, time_stepper="twostep", flux_scheme="tvdlf", limiter="minmod")
    if par.typepred1 != "hancock":
        return "twostep with tvdlf does not select the Hancock predictor"
    w0 = problems.field_loop_ct(mesh)
    be = OracleFV(par, mesh, lib_path=lib_path)
    be.set_state(0, w0)
    be.getbc(0)
    ti = TimeIntegrator(par, mesh, be)
    ti.setdt()
    dt = ti.state.dt
    be.advect1(FS["hancock"], 0.5, dt, 0.0, 0, 1, ("copy", 0))
    w1 = be.get_state(1)
    plain = dataclasses.replace(par, typedivbfix="none")
    bp = OracleFV(plain, mesh, lib_path=lib_path)
    bp.set_state(0, np.ascontiguousarray(w0[:, :plain.nw]))
    bp.getbc(0)
    bp.advect1(FS["hancock"], 0.5, dt, 0.0, 0, 1, ("copy", 0))
    wp = bp.get_state(1)
    msgs = []
    if not np.array_equal(mesh.interior(w1)[:, par.nw:], mesh.interior(be.get_state(0))[:, par.nw:]):
        msgs.append("the predictor changed the face field")
    if not np.array_equal(mesh.interior(w1)[:, :par.nw], mesh.interior(wp)):
        msgs.append("the predictor's cells differ from the run without a face field")
    be.advect1(FS["tvdlf"], 1.0, dt, 0.0, 1, 0, None)
    w = be.get_state(0)
    ex = problems.field_loop_output_vars_ct(mesh, w)
    divb = np.abs(mesh.interior(ex)[:, 0]).max()
    if not divb <= 1e-12 * np.abs(mesh.interior(w)[:, 5:7]).max() / min(mesh.dx):
        msgs.append(f"div B {divb:.3e} after the corrector")
    if np.array_equal(mesh.interior(w)[:, par.nw:], mesh.interior(w0)[:, par.nw:]):
        msgs.append("the corrector left the face field unchanged")
    return "; ".join(msgs)


def test_hancock_predictor_with_constrained_transport():
    assert check_hancock_ct() == ""


def check_small_values_crash(lib_path=None) -> str:
    """fix_small_values with the default small_values_method 'error'
    (src/mod_finite_volume.t:325-333, src/mod_small_values.t:25-87): a
    run stops ("small value encountered", src/amrvac.t:311-318) right after
    the first step whose updated state holds a cell below small_pressure.
    Two smoothly separating streams open a rarefaction whose central
    pressure falls; with onestep (the stage output is the step's state) and minmod
    (face states bounded by their cells) the crash step is exactly the
    first step whose minimum pressure is below the floor.  Without
    fix_small_values the same run continues (MHD) or stops only where the
    cmaxmean bound speed evaluates a small pressure (HD, hd_get_pthermal
    src/hd/mod_hd_phys.t:1046-1062)."""
    import dataclasses
    msgs = []
    gamma = 5.0 / 3.0
    mesh = UniformMesh(2, (64, 16), (16, 16), (0.0, 0.0), (1.0, 0.25), ("cont", "cont", "periodic", "periodic"))
    x = mesh.cell_centers()
    vx = 0.6 * np.tanh((x[:, 0] - 0.5) / 0.05)
    one = np.ones_like(vx)

    def run(par, w0, nmax=60):
        be = OracleFV(par, mesh, lib_path=lib_path)
        be.set_state(0, w0)
        be.getbc(0)
        ti = TimeIntegrator(par, mesh, be)
        pmin = []
        crashed = False
        try:
            for _ in range(nmax):
                ti.run(nsteps=1)
                w = mesh.interior(be.get_state(0))
                pmin.append(_pmin(w, par))
                ti.state.final_dt_exit = False
        except RuntimeError as e:
            if "small value" not in str(e):
                raise
            crashed = True
            pmin.append(_pmin(mesh.interior(be.get_state(0)), par))
        return crashed, ti.state.it, pmin

    def _pmin(w, par):
        if par.physics == "hd":
            return float(((gamma - 1) * (w[:, 3] - 0.5 * (w[:, 1] ** 2 + w[:, 2] ** 2) / w[:, 0])).min())
        ke = 0.5 * (w[:, 1] ** 2 + w[:, 2] ** 2 + w[:, 3] ** 2) / w[:, 0]
        return float(((gamma - 1) * (w[:, 4] - ke - 0.5 * (w[:, 5] ** 2 + w[:, 6] ** 2 + w[:, 7] ** 2))).min())
    floor = 0.02
    hd = Par(physics="hd", ndim=2, ndir=2, gamma=gamma, time_stepper="onestep", flux_scheme="tvdlf",
             limiter="minmod", courantpar=0.4, small_pressure=floor, fix_small_values=True)
    w_hd = problems.hd_to_conserved(np.stack([one, vx, 0 * one, 0.1 * one], axis=1), gamma, 2)
    crashed, it, pmin = run(hd, w_hd)
    first_low = next((k for k, p in enumerate(pmin) if p < floor), None)
    if not crashed or first_low is None or it != first_low:
        msgs.append(f"hd fix_small_values: crashed={crashed} at it={it}, first low step {first_low}")
    # without fix_small_values the run goes on with Einfeldt bounds, and
    # stops with cmaxmean bounds once a face mean falls below the floor
    crashed0, _, _ = run(dataclasses.replace(hd, fix_small_values=False), w_hd)
    crashed1, it1, _ = run(dataclasses.replace(hd, fix_small_values=False, typeboundspeed="cmaxmean"), w_hd)
    if crashed0 or not crashed1 or first_low is None or not first_low - 3 <= it1 <= first_low + 3:
        msgs.append(f"hd check_small_values only: Einfeldt crashed={crashed0}, cmaxmean crashed={crashed1} at {it1}")
    mhd = Par(physics="mhd", ndim=2, ndir=3, gamma=gamma, typedivbfix="none", time_stepper="onestep",
              flux_scheme="hll", limiter="minmod", courantpar=0.4, small_pressure=floor, fix_small_values=True)
    w_mhd = problems.mhd_to_conserved(np.stack([one, vx, 0 * one, 0 * one, 0.1 * one, 0.1 * one, 0 * one,
                                                0 * one], axis=1), gamma, 3)
    crashed, it, pmin = run(mhd, w_mhd)
    first_low = next((k for k, p in enumerate(pmin) if p < floor), None)
    if not crashed or first_low is None or it != first_low:
        msgs.append(f"mhd fix_small_values: crashed={crashed} at it={it}, first low step {first_low}")
    crashed0, _, _ = run(dataclasses.replace(mhd, fix_small_values=False), w_mhd)
    if crashed0:
        msgs.append("mhd run with check_small_values only crashed")
    return "; ".join(msgs)


def test_small_values_stop_the_run():
    assert check_small_values_crash() == ""


def check_linde_level_jump(lib_path=None) -> str:
    """Linde's source retreats one cell layer from faces with a coarser or
    finer neighbour (src/mhd/mod_mhd_phys.t:5616-5641): the maintained 3-D
    MHD blast log (levels 2 and 3 from the start) is reproduced to 1e-8
    with the rule and leaves the compare_logs band without it."""
    from amrvac_amd.amr import AMRIntegrator
    from amrvac_amd.host import logs_match
    from golden_setups import golden, setup
    par, forest, ap, init = setup("bw3d", "3step_hll_cada")
    ti = AMRIntegrator(par, forest, ap, lambda p, f: OracleFV(p, f, threads=8, lib_path=lib_path), init)
    got = np.array(ti.run().log)
    ref = golden("bw3d", "3step_hll_cada")
    if got.shape != ref.shape or not logs_match(got, ref):
        return "bw_3d_3step_hll_cada.log not reproduced"
    return ""


ALL_CHECKS = {
    "linde_level_jump": check_linde_level_jump,
    "ct_divb": check_ct_divb,
    "hancock_ct": check_hancock_ct,
    "ct_golden_contact": lambda p: check_ct_golden("3step_hll_cada_ct", p),
    "ct_golden_average": lambda p: check_ct_golden("3step_hll_cada_ct_average", p),
    "ct_golden_hll": lambda p: check_ct_golden("3step_hll_cada_ct_hll", p),
    "ct_golden_hlld": lambda p: check_ct_golden("3step_hlld_cada_ct", p),
    "sod_hllc_vanleer": lambda p: check_sod("3step_hllc_vanleer", p),
    "sod_hll_cada3": lambda p: check_sod("3step_hll_cada3", p),
    "sod_tvdlf_minmod": lambda p: check_sod("2step_tvdlf_minmod", p),
    "sod_tvdlf_weno5": lambda p: check_sod("3step_tvdlf_weno5", p),
    "smooth_twostep_vanleer": lambda p: check_smooth_order("twostep_tvdlf_vanleer", p),
    "smooth_hllc_vanleer": lambda p: check_smooth_order("threestep_hllc_vanleer", p),
    "smooth_hll_cada3": lambda p: check_smooth_order("threestep_hll_cada3", p),
    "smooth_tvdlf_weno5": lambda p: check_smooth_order("threestep_tvdlf_weno5", p),
    "hllc_contact": check_hllc_stationary_contact,
    "hlld_discontinuities": check_hlld_discontinuities,
    "hlld_paper": check_hlld_against_paper,
    "mhd_hllc_contact": check_mhd_hllc_contact,
    "flux_consistency": check_flux_consistency,
    "signal_speeds": check_signal_speeds,
    "glm_interface": check_glm_interface_state,
    "limiters": check_limiters,
    "weno5_polynomials": check_weno5_polynomials,
    "weno5_fix_limiter": check_fix_limiter,
    "small_values_crash": check_small_values_crash,
    "glm_decay": check_glm_decay,
    "conservation": check_conservation_and_uniform_states,
}
