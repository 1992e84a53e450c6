"""Initial conditions (the user-module side of the reference,
``usr_init_one_grid``) for the workloads this package runs.

Each function returns the conserved state of every block in the reference
layout ``w[ib, iw, (k,) j, i]`` with ghost cells, evaluated at the cell
centres of ``UniformMesh.cell_centers``.  Seeded noise is a counter-based
hash of (seed, level, global cell index, component), so it does not depend on
block order.
"""
from __future__ import annotations

import numpy as np

from .mesh import UniformMesh

DPI = 3.141592653589793238462643383279502884197169399375105  # src/mod_constants.t:21

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(x: np.ndarray) -> np.ndarray:
    with np.errstate(over="ignore"):
        z = (x + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def hashed_uniform(seed: int, level: int, gidx: np.ndarray, comp: int) -> np.ndarray:
    """Uniform noise in [-1, 1) for global cell indices ``gidx[nd, ...]``."""
    h = _splitmix64(np.full(gidx.shape[1:], np.uint64(seed & 0xFFFFFFFFFFFFFFFF)))
    vals = [level] + [gidx[d] for d in range(gidx.shape[0])] + [0] * (3 - gidx.shape[0]) + [comp]
    for v in vals:
        hv = np.asarray(v).astype(np.int64).astype(np.uint64)
        h = _splitmix64(h ^ hv)
    u = (h >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    return 2.0 * u - 1.0


def hd_to_conserved(prim: np.ndarray, gamma: float, ndir: int) -> np.ndarray:
    """hd_to_conserved, src/hd/mod_hd_phys.t:677-700 (w[..., iw] order)."""
    w = prim.copy()
    inv_g1 = 1.0 / (gamma - 1.0)
    vsq = w[:, 1] * w[:, 1]
    for c in range(1, ndir):
        vsq = vsq + w[:, 1 + c] * w[:, 1 + c]
    e = ndir + 1
    w[:, e] = w[:, e] * inv_g1 + 0.5 * vsq * w[:, 0]
    for c in range(ndir):
        w[:, 1 + c] = w[:, 0] * w[:, 1 + c]
    return w


def mhd_to_conserved(prim: np.ndarray, gamma: float, ndir: int) -> np.ndarray:
    """mhd_to_conserved_origin, src/mhd/mod_mhd_phys.t:1562-1579."""
    w = prim.copy()
    inv_g1 = 1.0 / (gamma - 1.0)
    vsq = w[:, 1] * w[:, 1]
    for c in range(1, ndir):
        vsq = vsq + w[:, 1 + c] * w[:, 1 + c]
    mag0 = ndir + 2
    bsq = w[:, mag0] * w[:, mag0]
    for c in range(1, ndir):
        bsq = bsq + w[:, mag0 + c] * w[:, mag0 + c]
    e = ndir + 1
    w[:, e] = w[:, e] * inv_g1 + 0.5 * (vsq * w[:, 0] + bsq)
    for c in range(ndir):
        w[:, 1 + c] = w[:, 0] * w[:, 1 + c]
    return w


def alfven_wave_2p5d(mesh: UniformMesh, gamma: float = 5.0 / 3.0) -> np.ndarray:
    """Circularly polarised Alfven wave, tests/mhd/Alfven_wave_2.5D/mod_usr.t
    (2-D, ndir=3, Linde cleaning: nw = 8)."""
    x = mesh.cell_centers()
    A = 0.1
    ca = 1.0
    k1 = 2.0 * DPI
    alpha = np.arctan(2.0)
    beta = 0.0
    k2 = k1 * np.tan(alpha)
    phi = k1 * x[:, 0] + k2 * x[:, 1]
    rho = np.ones_like(phi)
    p = 0.1 * np.ones_like(phi)
    v = [np.zeros_like(phi), A * np.sin(phi), A * np.cos(phi)]
    b = [np.sqrt(rho) * ca, -np.sqrt(rho) * ca * A * np.sin(phi),
         -np.sqrt(rho) * ca * A * np.cos(phi)]

    def rotate(u):
        g = np.arctan(np.cos(alpha) * np.tan(beta))
        r0 = np.cos(g) * np.cos(alpha) * u[0] - np.sin(alpha) * u[1] - np.sin(g) * np.cos(alpha) * u[2]
        r1 = np.cos(g) * np.sin(alpha) * u[0] + np.cos(alpha) * u[1] - np.sin(g) * np.sin(alpha) * u[2]
        r2 = np.sin(g) * u[0] + np.cos(g) * u[2]
        return [r0, r1, r2]

    vr = rotate(v)
    br = rotate(b)
    prim = np.stack([rho, vr[0], vr[1], vr[2], p, br[0], br[1], br[2]], axis=1)
    return mhd_to_conserved(prim, gamma, 3)


def kelvin_helmholtz_2d(mesh: UniformMesh, gamma: float = 1.4, seed: int = 20261202,
                        amp: float = 1.0e-3) -> np.ndarray:
    """2-D Euler Kelvin-Helmholtz: rho=2, vx=+0.5 for |y-0.5|<0.25, else
    rho=1, vx=-0.5; p=2.5; vy=0.01 sin(4 pi x); seeded uniform noise of
    amplitude ``amp`` on vx and vy."""
    x = mesh.cell_centers()
    g = mesh.global_index()
    inside = np.abs(x[:, 1] - 0.5) < 0.25
    rho = np.where(inside, 2.0, 1.0)
    vx = np.where(inside, 0.5, -0.5) + amp * hashed_uniform(seed, 1, g.transpose(1, 0, *range(2, g.ndim)), 0)
    vy = 0.01 * np.sin(4.0 * DPI * x[:, 0]) + amp * hashed_uniform(seed, 1, g.transpose(1, 0, *range(2, g.ndim)), 1)
    p = 2.5 * np.ones_like(rho)
    prim = np.stack([rho, vx, vy, p], axis=1)
    return hd_to_conserved(prim, gamma, 2)


def orszag_tang_3d(mesh: UniformMesh, gamma: float = 5.0 / 3.0, seed: int = 20261202,
                   amp: float = 1.0e-3, b_amp: float = 1.0, glm: bool = True,
                   rho_field: np.ndarray | None = None) -> np.ndarray:
    """3-D Orszag-Tang vortex: rho=25/(36 pi), p=5/(12 pi),
    v=(-sin 2 pi y, sin 2 pi x, vz), B=b_amp/sqrt(4 pi) (-sin 2 pi y,
    sin 4 pi x, 0), vz = ``amp`` seeded noise; psi=0 when GLM is on."""
    x = mesh.cell_centers()
    g = mesh.global_index()
    gi = g.transpose(1, 0, *range(2, g.ndim))
    rho0 = 25.0 / 36.0 / DPI
    p0 = 5.0 / 12.0 / DPI
    b0 = b_amp / np.sqrt(4.0 * DPI)
    rho = rho0 * np.ones_like(x[:, 0])
    if rho_field is not None:
        rho = rho * (1.0 + rho_field)
    vx = -np.sin(2.0 * DPI * x[:, 1])
    vy = np.sin(2.0 * DPI * x[:, 0])
    vz = amp * hashed_uniform(seed, 1, gi, 2)
    p = p0 * np.ones_like(rho)
    bx = -b0 * np.sin(2.0 * DPI * x[:, 1])
    by = b0 * np.sin(4.0 * DPI * x[:, 0])
    bz = np.zeros_like(rho)
    comps = [rho, vx, vy, vz, p, bx, by, bz]
    if glm:
        comps.append(np.zeros_like(rho))
    prim = np.stack(comps, axis=1)
    return mhd_to_conserved(prim, gamma, 3)


def _gradient_c(q: np.ndarray, axis: int, dx: float) -> np.ndarray:
    """Central gradient of src/mod_geometry.t:329-403 (nth=1) on a
    [..., (k,) j, i] array; returns values for interior+ghost cells with the
    outermost layer left at zero."""
    out = np.zeros_like(q)
    n = q.shape[axis]
    sl_c = [slice(None)] * q.ndim
    sl_p = [slice(None)] * q.ndim
    sl_m = [slice(None)] * q.ndim
    sl_c[axis] = slice(1, n - 1)
    sl_p[axis] = slice(2, n)
    sl_m[axis] = slice(0, n - 2)
    out[tuple(sl_c)] = (q[tuple(sl_p)] - q[tuple(sl_m)]) * 0.5
    return out / dx


def field_loop_az(mesh: UniformMesh, x: np.ndarray, v0, t: float,
                  A0: float = 1.0e-3, R0: float = 0.3) -> np.ndarray:
    """vectorpot_z of tests/mhd/field_loop_2.5D/mod_usr.t (periodic shift)."""
    lo, hi = mesh.xprobmin, mesh.xprobmax
    x1 = x[:, 0] - v0[0] * t - lo[0]
    x1 = lo[0] + (x1 - np.floor(x1 / (hi[0] - lo[0])) * (hi[0] - lo[0]))
    x2 = x[:, 1] - v0[1] * t - lo[1]
    x2 = lo[1] + (x2 - np.floor(x2 / (hi[1] - lo[1])) * (hi[1] - lo[1]))
    r2 = x1 ** 2 + x2 ** 2
    return np.where(r2 < R0 ** 2, A0 * (R0 - np.sqrt(r2)), 0.0)


def field_loop_bfield(mesh: UniformMesh, x: np.ndarray, v0, t: float):
    """bfield_solution (init_from_vectorpot branch): B = curl(Az z) with the
    central gradient, evaluated on the interior."""
    az = field_loop_az(mesh, x, v0, t)
    b1 = _gradient_c(az, az.ndim - 2, mesh.dx[1])   # d/dy
    b2 = -_gradient_c(az, az.ndim - 1, mesh.dx[0])  # -d/dx
    return b1, b2


def field_loop_2p5d(mesh: UniformMesh, gamma: float = 5.0 / 3.0) -> np.ndarray:
    """Advected field loop, tests/mhd/field_loop_2.5D/mod_usr.t (iprob=1):
    rho=1, v=(2,1,0), p=1, B from the vector potential; the 8 MHD
    variables only (the three diagnostic variables of that module are
    computed at output time)."""
    x = mesh.cell_centers()
    v0 = (2.0, 1.0)
    b1, b2 = field_loop_bfield(mesh, x, v0, 0.0)
    one = np.ones_like(b1)
    prim = np.stack([one, 2.0 * one, 1.0 * one, 0.0 * one, 1.0 * one, b1, b2, 0.0 * one], axis=1)
    return mhd_to_conserved(prim, gamma, 3)


def field_loop_output_vars(mesh: UniformMesh, w: np.ndarray, t: float) -> np.ndarray:
    """set_output_vars of the field-loop module: div B (2nd and 4th order,
    src/mod_geometry.t:582-655) and |B - B_exact|; returns [nb, 3, ...]."""
    ng = mesh.nghost
    dx, dy = mesh.dx
    bx, by = w[:, 5], w[:, 6]
    ny, nx = w.shape[-2:]
    d2 = np.zeros_like(bx)
    d4 = np.zeros_like(bx)
    I = slice(ng, nx - ng)
    J = slice(ng, ny - ng)
    div = np.zeros_like(bx[:, J, I])
    div = div + 0.5 * (bx[:, J, ng + 1:nx - ng + 1] - bx[:, J, ng - 1:nx - ng - 1]) * (1.0 / dx)
    div = div + 0.5 * (by[:, ng + 1:ny - ng + 1, I] - by[:, ng - 1:ny - ng - 1, I]) * (1.0 / dy)
    d2[:, J, I] = div
    div4 = np.zeros_like(bx[:, J, I])
    div4 = div4 + (-bx[:, J, ng + 2:nx - ng + 2] + 8.0 * bx[:, J, ng + 1:nx - ng + 1]
                   - 8.0 * bx[:, J, ng - 1:nx - ng - 1] + bx[:, J, ng - 2:nx - ng - 2]) / (12.0 * dx)
    div4 = div4 + (-by[:, ng + 2:ny - ng + 2, I] + 8.0 * by[:, ng + 1:ny - ng + 1, I]
                   - 8.0 * by[:, ng - 1:ny - ng - 1, I] + by[:, ng - 2:ny - ng - 2, I]) / (12.0 * dy)
    d4[:, J, I] = div4
    x = mesh.cell_centers()
    e1, e2 = field_loop_bfield(mesh, x, (2.0, 1.0), t)
    err = np.sqrt((bx - e1) ** 2 + (by - e2) ** 2)
    berr = np.zeros_like(bx)
    berr[:, J, I] = err[:, J, I]
    return np.stack([d2, d4, berr], axis=1)


def field_loop_faces(mesh: UniformMesh) -> np.ndarray:
    """The face field of the field loop from its vector potential
    (b_from_vector_potentialA src/mod_constrained_transport.t:103-173 with
    initvecpot_usr): A_z = A0 (R0 - r) inside r <= R0 at the cell corners
    x + dx/2, y + dy/2, then B = curl A as circulation / face area.
    Returns ws [nb, 2, ny, nx]: component d at the index of the cell whose
    upper d face it sits on, set on the faces of the interior cells and
    their lower neighbours (ghost faces are left zero for getbc)."""
    a0, r0 = 1.0e-3, 0.3
    x = mesh.cell_centers()
    dx, dy = mesh.dx
    xc = x[:, 0] + 0.5 * dx
    yc = x[:, 1] + 0.5 * dy
    rx = np.sqrt(xc ** 2 + yc ** 2)
    az = np.where(rx <= r0, a0 * (r0 - rx), 0.0) * 1.0  # times dsC = 1 (ignorable z)
    ng = mesh.nghost
    ny, nx = az.shape[-2:]
    ws = np.zeros((az.shape[0], 2, ny, nx))
    J, I = slice(ng, ny - ng), slice(ng, nx - ng)
    I1, J1 = slice(ng - 1, nx - ng), slice(ng - 1, ny - ng)
    # x faces: lvc(1,2,3) = +1, area dy
    ws[:, 0, J, I1] = (0.0 + (az[:, J, I1] - az[:, ng - 1:ny - ng - 1, I1])) / dy
    # y faces: lvc(2,1,3) = -1, area dx
    ws[:, 1, J1, I] = (0.0 + (-1.0) * (az[:, J1, I] - az[:, J1, ng - 1:nx - ng - 1])) / dx
    return ws


def field_loop_ct(mesh: UniformMesh, gamma: float = 5.0 / 3.0) -> np.ndarray:
    """The field loop on a staggered mesh (stagger_grid branch of
    initonegrid_usr): face field from the vector potential, cell field by
    mhd_face_to_center (src/mhd/mod_mhd_phys.t:7586-7648).  Returns the
    backend state [nb, 8 + 2, ny, nx]: conserved cell variables, then the
    face field."""
    ws = field_loop_faces(mesh)
    dx, dy = mesh.dx
    ng = mesh.nghost
    ny, nx = ws.shape[-2:]
    J, I = slice(ng, ny - ng), slice(ng, nx - ng)
    b1 = np.zeros_like(ws[:, 0])
    b2 = np.zeros_like(ws[:, 0])
    b1[:, J, I] = 0.5 / dy * (ws[:, 0, J, I] * dy + ws[:, 0, J, ng - 1:nx - ng - 1] * dy)
    b2[:, J, I] = 0.5 / dx * (ws[:, 1, J, I] * dx + ws[:, 1, ng - 1:ny - ng - 1, I] * dx)
    one = np.ones_like(b1)
    prim = np.stack([one, 2.0 * one, 1.0 * one, 0.0 * one, 1.0 * one, b1, b2, 0.0 * one], axis=1)
    return np.concatenate([mhd_to_conserved(prim, gamma, 3), ws], axis=1)


def field_loop_3d_ct(mesh: UniformMesh, gamma: float = 5.0 / 3.0, a0: float = 1.0e-3,
                     r0: float = 0.3) -> np.ndarray:
    """The 3-D advected field loop of Gardiner & Stone (2008, J. Comput.
    Phys. 227, 4123, sec. 5.4) on a staggered mesh over the periodic box
    [-0.5, 0.5] x [-0.5, 0.5] x [-1, 1]: a cylinder of field lines along the
    axis e3 = (-1, 0, 2)/sqrt(5), vector potential A = A0 (R0 - r) e3 inside
    r <= R0, where r is the distance from that axis (periodic images by
    wrapping x1 = (2x + z)/sqrt(5) and y), advected with v = (1, 1, 2) at
    rho = p = 1.  The face field is built as b_from_vector_potentialA does
    (src/mod_constrained_transport.t:103-173): each potential component at
    the centre of its edges times the edge length, then per face the
    circulation in lvc order over the face area; the cell field is the face
    average (mhd_face_to_center src/mhd/mod_mhd_phys.t:7586-7648).  Returns
    the backend state [nb, 8 + 3, nz, ny, nx] (ghost cells and ghost faces
    zero, for getbc)."""
    assert mesh.ndim == 3
    x = mesh.cell_centers()
    dx = mesh.dx
    s5 = np.sqrt(5.0)
    p1 = 2.0 / s5  # period of x1 over the box

    def a3(xe, ye, ze):
        x1 = (2.0 * xe + ze) / s5
        x1 = x1 - p1 * np.floor(x1 / p1 + 0.5)
        x2 = ye - np.floor(ye + 0.5)
        r = np.sqrt(x1 * x1 + x2 * x2)
        return np.where(r <= r0, a0 * (r0 - r), 0.0)

    xc, yc, zc = x[:, 0], x[:, 1], x[:, 2]
    h = 0.5 * np.asarray(dx)
    # edge potentials times edge length: A_x at (x, y + dy/2, z + dz/2),
    # A_y = 0, A_z at (x + dx/2, y + dy/2, z)
    edge = [-a3(xc, yc + h[1], zc + h[2]) / s5 * dx[0],
            np.zeros_like(xc),
            2.0 * a3(xc + h[0], yc + h[1], zc) / s5 * dx[2]]
    lvc = {(0, 1, 2): 1, (1, 2, 0): 1, (2, 0, 1): 1, (0, 2, 1): -1, (2, 1, 0): -1, (1, 0, 2): -1}
    area = (dx[1] * dx[2], dx[0] * dx[2], dx[0] * dx[1])
    ng = mesh.nghost
    nz, ny, nx = x.shape[-3:]
    # faces of the interior cells and their lower neighbours
    nb = x.shape[0]
    ws = np.zeros((nb, 3, nz, ny, nx))
    for d in range(3):
        lo = [ng, ng, ng]
        lo[d] = ng - 1
        sl = tuple(slice(lo[e], (nx, ny, nz)[e] - ng) for e in range(3))[::-1]
        circ = np.zeros_like(xc[(slice(None),) + sl])
        for idim2 in range(3):
            for idir in range(3):
                l = lvc.get((d, idim2, idir), 0)
                if l == 0:
                    continue
                lo2 = list(lo)
                lo2[idim2] -= 1
                sl2 = tuple(slice(lo2[e], lo2[e] + (sl[2 - e].stop - sl[2 - e].start)) for e in range(3))[::-1]
                da = edge[idir][(slice(None),) + sl] - edge[idir][(slice(None),) + sl2]
                circ = circ + da if l == 1 else circ - da
        ws[(slice(None), d) + sl] = circ / area[d]
    inner = tuple(slice(ng, n - ng) for n in (nz, ny, nx))
    b = np.zeros((nb, 3, nz, ny, nx))
    for d in range(3):
        lower = list(inner)
        lower[2 - d] = slice(ng - 1, (nz, ny, nx)[2 - d] - ng - 1)
        b[(slice(None), d) + inner] = 0.5 / area[d] * (ws[(slice(None), d) + inner] * area[d] +
                                                       ws[(slice(None), d) + tuple(lower)] * area[d])
    one = np.ones_like(xc)
    prim = np.stack([one, 1.0 * one, 1.0 * one, 2.0 * one, 1.0 * one, b[:, 0], b[:, 1], b[:, 2]], axis=1)
    return np.concatenate([mhd_to_conserved(prim, gamma, 3), ws], axis=1)


def field_loop_output_vars_ct(mesh: UniformMesh, w: np.ndarray) -> np.ndarray:
    """set_output_vars of the field-loop module on a staggered mesh: div B
    of the face field (get_divb src/physics/mod_functions_bfield.t:30-37,
    both orders) and |B - B(t=0)| against the face-averaged initial field;
    w is the backend state with the face field after the 8 cell variables.
    Returns [nb, 3, ny, nx]."""
    ng = mesh.nghost
    dx, dy = mesh.dx
    bx, by = w[:, 5], w[:, 6]
    ws = w[:, 8:10]
    ny, nx = bx.shape[-2:]
    J, I = slice(ng, ny - ng), slice(ng, nx - ng)
    div = np.zeros_like(bx[:, J, I])
    div = div + ws[:, 0, J, I] * dy - ws[:, 0, J, ng - 1:nx - ng - 1] * dy
    div = div + ws[:, 1, J, I] * dx - ws[:, 1, ng - 1:ny - ng - 1, I] * dx
    div = div / (dx * dy)
    d2 = np.zeros_like(bx)
    d2[:, J, I] = div
    w0 = field_loop_faces(mesh)
    e1 = 0.5 * (w0[:, 0, J, I] + w0[:, 0, J, ng - 1:nx - ng - 1])
    e2 = 0.5 * (w0[:, 1, J, I] + w0[:, 1, ng - 1:ny - ng - 1, I])
    berr = np.zeros_like(bx)
    berr[:, J, I] = np.sqrt((bx[:, J, I] - e1) ** 2 + (by[:, J, I] - e2) ** 2)
    return np.stack([d2, d2.copy(), berr], axis=1)


def blast_wave(mesh: UniformMesh, gamma: float = 5.0 / 3.0, seed: int = 20261202,
               rho0: float = 1.0, p_out: float = 1.0e-2, p_in: float = 10.0,
               radius: float = 0.05, noise: float = 0.01) -> np.ndarray:
    """Hydrodynamic blast wave: rho=1 with seeded relative density noise,
    v=0, p=p_in inside r<radius around the domain centre, p_out outside."""
    x = mesh.cell_centers()
    g = mesh.global_index()
    gi = g.transpose(1, 0, *range(2, g.ndim))
    nd = mesh.ndim
    r2 = np.zeros_like(x[:, 0])
    for d in range(nd):
        c = 0.5 * (mesh.xprobmin[d] + mesh.xprobmax[d])
        r2 = r2 + (x[:, d] - c) ** 2
    rho = rho0 * (1.0 + noise * hashed_uniform(seed, 1, gi, 0))
    p = np.where(r2 < radius ** 2, p_in, p_out)
    comps = [rho] + [np.zeros_like(rho)] * nd + [p]
    return hd_to_conserved(np.stack(comps, axis=1), gamma, nd)


def orszag_tang_2p5d(mesh: UniformMesh, gamma: float = 5.0 / 3.0, seed: int = 20261202,
                     amp: float = 1.0e-3, glm: bool = True) -> np.ndarray:
    """2-D Orszag-Tang vortex with three vector components (as
    tests/mhd/Orszag_Tang_2D/mod_usr.t, plus a seeded vz perturbation)."""
    x = mesh.cell_centers()
    g = mesh.global_index()
    gi = g.transpose(1, 0, *range(2, g.ndim))
    rho0 = 25.0 / 36.0 / DPI
    p0 = 5.0 / 12.0 / DPI
    b0 = 1.0 / np.sqrt(4.0 * DPI)
    one = np.ones_like(x[:, 0])
    comps = [rho0 * one, -np.sin(2.0 * DPI * x[:, 1]), np.sin(2.0 * DPI * x[:, 0]),
             amp * hashed_uniform(seed, 1, gi, 2), p0 * one, -b0 * np.sin(2.0 * DPI * x[:, 1]),
             b0 * np.sin(4.0 * DPI * x[:, 0]), 0.0 * one]
    if glm:
        comps.append(0.0 * one)
    return mhd_to_conserved(np.stack(comps, axis=1), gamma, 3)


def blast_wave_blocks(forest, blocks, gamma: float = 5.0 / 3.0, rho0: float = 1.0, p_out: float = 1.0,
                      p_in: float = 100.0, radius: float = 0.2, noise: float = 0.0,
                      seed: int = 20261202) -> np.ndarray:
    """Blast wave on the given leaves of an AMR forest (ghosted conserved
    state [len(blocks), nw, ...]): rho=rho0 (times 1 + seeded relative noise
    hashed on the level and the level's global cell index), v=0, p=p_in for
    r < radius around the domain centre, p_out outside.  With the defaults
    it is the initial condition of tests/hd/blast_wave_Cartesian_2D/3D
    (mod_usr.t: p=100 inside r<0.2, 1 outside)."""
    x = forest.cell_centers(blocks)
    nd = forest.ndim
    r2 = np.zeros_like(x[:, 0])
    for d in range(nd):
        c = 0.5 * (forest.xprobmin[d] + forest.xprobmax[d])
        r2 = r2 + (x[:, d] - c) ** 2
    rho = rho0 * np.ones_like(r2)
    if noise:
        for k, b in enumerate(blocks):
            gi = forest.global_index(b)
            rho[k] = rho0 * (1.0 + noise * hashed_uniform(seed, int(forest.level[b]), gi, 0))
    p = np.where(r2 < radius ** 2, p_in, p_out)
    comps = [rho] + [np.zeros_like(rho)] * nd + [p]
    return hd_to_conserved(np.stack(comps, axis=1), gamma, nd)


def kelvin_helmholtz_2d_blocks(forest, blocks, gamma: float = 5.0 / 3.0, vextra: float = 0.0) -> np.ndarray:
    """Kelvin-Helmholtz shear layers of tests/hd/Kelvin_Helmholtz_2D/mod_usr.t
    (iprob=1): rho=10 for 0.25<=y<=0.75 else 1, p=2.5, vx=+-0.5 with linear
    ramps of half-width 0.025 across y=0.25 and y=0.75, vy=0.01 sin(4 pi x)
    times two Gaussians of variance 0.00125 centred on the layers."""
    x = forest.cell_centers(blocks)
    y = x[:, 1]
    rhodens, rholight, kx, pint, dlin, sigma = 10.0, 1.0, 4.0 * DPI, 2.5, 0.025, 0.00125
    rho = np.where((y < 0.25) | (y > 0.75), rholight, rhodens)
    p = pint * np.ones_like(y)
    vx = np.zeros_like(y)
    vx = np.where((y <= 0.25 - dlin) | (y >= 0.75 + dlin), vextra - 0.5, vx)
    vx = np.where((y >= 0.25 + dlin) & (y <= 0.75 - dlin), vextra + 0.5, vx)
    ramp1 = (y - (0.25 - dlin)) * ((vextra + 0.5) - (vextra - 0.5)) / (2.0 * dlin) + (vextra - 0.5)
    vx = np.where((y > 0.25 - dlin) & (y < 0.25 + dlin), ramp1, vx)
    ramp2 = (y - (0.75 - dlin)) * ((vextra - 0.5) - (vextra + 0.5)) / (2.0 * dlin) + (vextra + 0.5)
    vx = np.where((y > 0.75 - dlin) & (y < 0.75 + dlin), ramp2, vx)
    vy = 0.01 * np.sin(kx * x[:, 0]) * (np.exp(-0.5 * (y - 0.25) ** 2 / sigma)
                                        + np.exp(-0.5 * (y - 0.75) ** 2 / sigma))
    return hd_to_conserved(np.stack([rho, vx, vy, p], axis=1), gamma, 2)


def mhd_kelvin_helmholtz_3d_blocks(forest, blocks, gamma: float = 5.0 / 3.0, glm: bool = False) -> np.ndarray:
    """MHD Kelvin-Helmholtz jet of tests/mhd/Kelvin_Helmholtz_3D/mod_usr.t
    (Keppens & Toth): rho=1, p=1, B=(0.129, 0, 0), vx=0.645 tanh((r-0.5)/0.05)
    with r the distance from the x axis, and a 0.01 perturbation of (vy, vz)
    localised on the jet surface; nw=8, plus psi=0 with GLM cleaning."""
    x = forest.cell_centers(blocks)
    r = np.sqrt(x[:, 1] ** 2 + x[:, 2] ** 2)
    phi = np.arctan2(x[:, 2], x[:, 1])
    mpol, k1 = 1.0, 2.0 * DPI / (forest.xprobmax[0] - forest.xprobmin[0])
    rjet, qv, b0, dv = 0.5, 0.645, 0.129, 0.01
    width = 0.1 * rjet
    one = np.ones_like(r)
    vx = qv * np.tanh((r - rjet) / width)
    env = dv * np.exp(-((r - rjet) / (4.0 * width)) ** 2) * np.cos(mpol * phi) * np.sin(k1 * x[:, 0])
    vy = env * np.cos(phi)
    vz = env * np.sin(phi)
    comps = [one, vx, vy, vz, one, b0 * one, 0.0 * one, 0.0 * one] + ([0.0 * one] if glm else [])
    return mhd_to_conserved(np.stack(comps, axis=1), gamma, 3)


def mhd_blast_wave_3d_blocks(forest, blocks, gamma: float = 5.0 / 3.0, busr: float = 1.0,
                             glm: bool = False) -> np.ndarray:
    """MHD blast wave of tests/mhd/blast_wave_Cartesian_3D/mod_usr.t: rho=1,
    v=0, p=100 inside r<0.2 around the domain centre and 1 outside, uniform
    B=(Busr, Busr, Busr); nw=8, plus psi=0 with GLM cleaning."""
    x = forest.cell_centers(blocks)
    r2 = np.zeros_like(x[:, 0])
    for d in range(3):
        c = (forest.xprobmin[d] + forest.xprobmax[d]) * 0.5
        r2 = r2 + (x[:, d] - c) ** 2
    one = np.ones_like(r2)
    p = np.where(r2 < 0.2 ** 2, 100.0, 1.0)
    comps = [one, 0.0 * one, 0.0 * one, 0.0 * one, p, busr * one, busr * one, busr * one]
    comps += [0.0 * one] if glm else []
    return mhd_to_conserved(np.stack(comps, axis=1), gamma, 3)


def orszag_tang_2d_blocks(forest, blocks, gamma: float = 5.0 / 3.0, glm: bool = False) -> np.ndarray:
    """Orszag-Tang vortex of tests/mhd/Orszag_Tang_2D/mod_usr.t with two
    vector components (ndir=2: nw=6, plus psi=0 with GLM cleaning)."""
    x = forest.cell_centers(blocks)
    rho0 = 25.0 / 36.0 / DPI
    p0 = 5.0 / 12.0 / DPI
    b0 = 1.0 / np.sqrt(4.0 * DPI)
    one = np.ones_like(x[:, 0])
    comps = [rho0 * one, -np.sin(2.0 * DPI * x[:, 1]), np.sin(2.0 * DPI * x[:, 0]), p0 * one,
             -b0 * np.sin(2.0 * DPI * x[:, 1]), b0 * np.sin(4.0 * DPI * x[:, 0])] + ([0.0 * one] if glm else [])
    return mhd_to_conserved(np.stack(comps, axis=1), gamma, 2)
