"""Block-structured AMR host: the octree forest, its connectivity, the Löhner
error estimator and the regrid decisions of MPI-AMRVAC, on one rank.

The forest mirrors src/amr/mod_forest.t: level-1 roots cover the domain in
blocks of ``block_nx`` cells, every refinement splits a block into 2^ndim
children with half the cell size (src/mod_initialize.t:101-113), and the
leaves are ordered along the Morton curve (src/amr/mod_space_filling_curve.t).
A leaf's block index is its position in that order; backends store one state
block per leaf in that order.

Neighbour types per direction (i in {-1,0,1}^ndim, src/amr/mod_amr_neighbors.t
find_neighbor): 1 boundary, 2 coarse, 3 sibling, 4 fine.
"""
from __future__ import annotations

import itertools
import math
from dataclasses import dataclass, field

import numpy as np

from . import amr_native

NB_BOUNDARY, NB_COARSE, NB_SIBLING, NB_FINE = 1, 2, 3, 4


def morton_encode(ig0: tuple[int, ...]) -> int:
    """Bit interleave of 0-based indices, dimension 1 in the lowest bit
    (mortonEncode, src/amr/mod_space_filling_curve.t)."""
    nd = len(ig0)
    key = 0
    for bit in range(64 // nd + 1):
        for d, v in enumerate(ig0):
            key |= ((v >> bit) & 1) << (nd * bit + d)
    return key


class Node:
    __slots__ = ("level", "ig", "parent", "children", "block")

    def __init__(self, level: int, ig: tuple[int, ...], parent: "Node | None"):
        self.level = level
        self.ig = ig
        self.parent = parent
        self.children: list[Node] | None = None  # 2^ndim, index sum((ic_d-1) << d)
        self.block = -1

    @property
    def leaf(self) -> bool:
        return self.children is None

    def ic(self) -> tuple[int, ...]:
        """Position within the parent, 1 or 2 per dimension."""
        return tuple(1 + (g - 1) % 2 for g in self.ig)


def child_index(ic: tuple[int, ...]) -> int:
    return sum((c - 1) << d for d, c in enumerate(ic))


def dir_index(i: tuple[int, ...]) -> int:
    """Index of direction i in the 3^ndim neighbour table (dimension 1
    fastest), as the uniform mesh's neighbour table."""
    return sum((v + 1) * 3 ** d for d, v in enumerate(i))


@dataclass
class AMRForest:
    ndim: int
    domain_nx: tuple[int, ...]
    block_nx: tuple[int, ...]
    xprobmin: tuple[float, ...]
    xprobmax: tuple[float, ...]
    bc: tuple[str, ...]        # 2*ndim entries: min1, max1, min2, max2, ...
    max_level: int = 1
    nghost: int = 2
    nodes: dict = field(init=False, default_factory=dict)
    roots: list = field(init=False, default_factory=list)

    def __post_init__(self) -> None:
        nd = self.ndim
        self.periodic = tuple(self.bc[2 * d] == "periodic" for d in range(nd))
        self.ng1 = tuple(self.domain_nx[d] // self.block_nx[d] for d in range(nd))
        # dx(level) = dx(level-1)/2, dg(level) = dg(level-1)/2
        # (src/mod_initialize.t:95-113)
        self.dx = {1: tuple((self.xprobmax[d] - self.xprobmin[d]) / self.domain_nx[d] for d in range(nd))}
        self.dg = {1: tuple(self.dx[1][d] * float(self.block_nx[d]) for d in range(nd))}
        for lev in range(2, max(self.max_level, 1) + 2):
            self.dx[lev] = tuple(v * 0.5 for v in self.dx[lev - 1])
            self.dg[lev] = tuple(0.5 * v for v in self.dg[lev - 1])
        # level-1 roots in Morton order over the enclosing power-of-two square
        # (level1_Morton_order)
        ngsq = max(2 ** math.ceil(math.log(float(n)) / math.log(2.0)) if n > 1 else 1 for n in self.ng1)
        order = []
        for ig0 in itertools.product(*[range(ngsq)] * nd):
            if all(ig0[d] < self.ng1[d] for d in range(nd)):
                order.append((morton_encode(ig0), tuple(v + 1 for v in ig0)))
        order.sort()
        for _, ig in order:
            n = Node(1, ig, None)
            self.nodes[(1, ig)] = n
            self.roots.append(n)
        self.update()

    # -- structure -------------------------------------------------------------
    def ng(self, level: int) -> tuple[int, ...]:
        return tuple(n * 2 ** (level - 1) for n in self.ng1)

    def refine_node(self, node: Node) -> list[Node]:
        assert node.leaf
        kids = [None] * (2 ** self.ndim)
        for ic in itertools.product(*[(1, 2)] * self.ndim):
            ig = tuple(2 * (node.ig[d] - 1) + ic[d] for d in range(self.ndim))
            c = Node(node.level + 1, ig, node)
            self.nodes[(c.level, ig)] = c
            kids[child_index(ic)] = c
        node.children = kids
        return kids

    def coarsen_node(self, parent: Node) -> None:
        for c in parent.children:
            assert c.leaf
            del self.nodes[(c.level, c.ig)]
        parent.children = None

    def update(self) -> None:
        """Morton order of the leaves (amr_Morton_order) and connectivity."""
        leaves: list[Node] = []

        def walk(n: Node) -> None:
            if n.leaf:
                leaves.append(n)
            else:
                for c in n.children:  # ic(1) fastest
                    walk(c)

        for r in self.roots:
            walk(r)
        for b, n in enumerate(leaves):
            n.block = b
        self.leaves = leaves
        # the leaves in the order the forest created them (the level lists
        # proper_nesting walks)
        self.leaf_order = np.fromiter((n.block for n in self.nodes.values() if n.leaf), dtype=np.int32,
                                      count=len(leaves))
        self._connectivity()

    @property
    def nblocks(self) -> int:
        return len(self.leaves)

    @property
    def levmin(self) -> int:
        return int(self.level.min())

    @property
    def levmax(self) -> int:
        return int(self.level.max())

    def _connectivity(self) -> None:
        """Neighbour tables of the leaves (build_connectivity,
        src/mod_functions_connectivity.t:55-431) from the native tree
        routines: per leaf and direction the type, the same-level or coarser
        leaf, and the children of a finer neighbour that touch the leaf."""
        nd = self.ndim
        nb = self.nblocks
        self.level = np.array([n.level for n in self.leaves], dtype=np.int32)
        self.ig = np.array([n.ig for n in self.leaves], dtype=np.int32).reshape(nb, nd)
        self.icpos = np.array([n.ic() for n in self.leaves], dtype=np.int32).reshape(nb, nd)
        self.ntype, self.nbr, self.nchild = amr_native.connectivity(nd, self.ng1, self.periodic, self.level, self.ig)
        # block extents (alloc_node, src/amr/mod_amr_solution_node.t:150-160),
        # elementwise in float64 as the scalar formulas round
        levs = sorted(self.dg)
        dg = np.array([self.dg[lv] for lv in levs], dtype=np.float64)[self.level - levs[0]]
        dxl = np.array([self.dx[lv] for lv in levs], dtype=np.float64)[self.level - levs[0]]
        xmin = np.array(self.xprobmin, dtype=np.float64)
        xmax = np.array(self.xprobmax, dtype=np.float64)
        self.xlohi = np.empty((nb, nd, 2), dtype=np.float64)
        self.xlohi[:, :, 0] = xmin + (self.ig - 1).astype(np.float64) * dg
        self.xlohi[:, :, 1] = np.minimum(xmin + self.ig.astype(np.float64) * dg, xmax)
        self.dxb = np.ascontiguousarray(dxl)

    # -- geometry of the leaves (the UniformMesh interface) ----------------
    @property
    def nxg(self) -> tuple[int, ...]:
        return tuple(self.block_nx[d] + 2 * self.nghost for d in range(self.ndim))

    @property
    def bc_codes(self) -> tuple[int, ...]:
        from .mesh import bc_code
        return tuple(bc_code(b) for b in self.bc)

    def cell_centers(self, blocks=None) -> np.ndarray:
        """x[ib, d, (k,) j, i] of the given leaves (all by default), ghost cells
        included, as alloc_node sets them."""
        nd = self.ndim
        ng = self.nghost
        blocks = range(self.nblocks) if blocks is None else blocks
        shp = self.nxg[::-1]
        out = np.zeros((len(blocks), nd) + shp, dtype=np.float64)
        for k, ib in enumerate(blocks):
            for d in range(nd):
                nxd = self.nxg[d]
                bn = self.block_nx[d]
                ix = np.arange(1, nxd + 1)
                lo, hi = self.xlohi[ib, d]
                dx = self.dxb[ib, d]
                v = np.where(ix <= bn, lo + ((ix - ng).astype(np.float64) - 0.5) * dx,
                             hi + ((ix - (bn + ng)).astype(np.float64) - 0.5) * dx)
                bshape = [1] * nd
                bshape[nd - 1 - d] = nxd
                out[k, d] = np.broadcast_to(v.reshape(bshape), shp)
        return out

    def global_index(self, b: int) -> np.ndarray:
        """0-based cell index on the leaf's level for every cell of leaf b
        (ghosts included), shape [nd, (k,) j, i]."""
        nd = self.ndim
        shp = self.nxg[::-1]
        out = np.zeros((nd,) + shp, dtype=np.int64)
        for d in range(nd):
            nxd = self.nxg[d]
            g = (int(self.ig[b, d]) - 1) * self.block_nx[d] + np.arange(nxd) - self.nghost
            bshape = [1] * nd
            bshape[nd - 1 - d] = nxd
            out[d] = np.broadcast_to(g.reshape(bshape), shp)
        return out

    def interior(self, w: np.ndarray) -> np.ndarray:
        ng = self.nghost
        sl = tuple(slice(ng, ng + self.block_nx[d]) for d in range(self.ndim))[::-1]
        return w[(slice(None), slice(None)) + sl]

    def cell_volume(self) -> np.ndarray:
        """Cell volume per leaf (uniform within a block)."""
        return np.prod(self.dxb, axis=1)


# ---------------------------------------------------------------------------
# Löhner error estimator, lohner_grid src/amr/mod_errest.t:63-199 (Cartesian,
# no B0, no user variable); w is one ghosted block [nw, (k,) j, i]
# ---------------------------------------------------------------------------
def lohner_error(w: np.ndarray, ndim: int, nghost: int, weights: dict[int, float],
                 wavefilter: float, logflag: dict[int, bool] | None = None) -> np.ndarray:
    """Error estimate per interior cell (shape (k,) j, i)."""
    logflag = logflag or {}
    epsilon = 1.0e-6
    shp = w.shape[1:]
    ng = nghost
    # axis of dimension d in the numpy array
    ax = [ndim - 1 - d for d in range(ndim)]

    def sl(lo_off, hi_off, shift_d=None, shift=0):
        """slice tuple of the range [ng+lo_off, n-ng+hi_off) shifted by
        `shift` along dimension shift_d"""
        out = []
        for a in range(ndim):
            d = ndim - 1 - a
            n = shp[a]
            s = shift if d == shift_d else 0
            out.append(slice(ng + lo_off + s, n - ng + hi_off + s))
        return tuple(out)

    error = np.zeros(tuple(n - 2 * ng for n in shp))
    for iflag, weight in sorted(weights.items()):
        if weight == 0.0:
            continue
        q = w[iflag]
        lg = logflag.get(iflag, False)
        numerator = np.zeros_like(error)
        for d in range(ndim):
            # tmp(ix) = w(jx) - w(hx) on ix = ixM extended by one cell
            if lg:
                tmp = np.log10(q[sl(-1, 1, d, 1)]) - np.log10(q[sl(-1, 1, d, -1)])
            else:
                tmp = q[sl(-1, 1, d, 1)] - q[sl(-1, 1, d, -1)]
            for d2 in range(ndim):
                # tmp over ixM shifted by +-1 in d2, tmp indices start at ixM-1
                jj = tuple(slice(1 + (1 if dd == d2 else 0), tmp.shape[a] - 1 + (1 if dd == d2 else 0))
                           for a, dd in ((a, ndim - 1 - a) for a in range(ndim)))
                hh = tuple(slice(1 - (1 if dd == d2 else 0), tmp.shape[a] - 1 - (1 if dd == d2 else 0))
                           for a, dd in ((a, ndim - 1 - a) for a in range(ndim)))
                diff = tmp[jj] - tmp[hh]
                numerator = numerator + diff * diff
        denominator = np.zeros_like(error)
        for d in range(ndim):
            absq = np.abs(np.log10(q)) if lg else np.abs(q)
            tmp2 = absq[sl(-1, 1, d, 1)] + absq[sl(-1, 1, d, -1)]
            if lg:
                t = (np.abs(np.log10(q[sl(0, 0, d, 2)]) - np.log10(q[sl(0, 0)]))
                     + np.abs(np.log10(q[sl(0, 0)]) - np.log10(q[sl(0, 0, d, -2)])))
            else:
                t = np.abs(q[sl(0, 0, d, 2)] - q[sl(0, 0)]) + np.abs(q[sl(0, 0)] - q[sl(0, 0, d, -2)])
            for d2 in range(ndim):
                jj = tuple(slice(1 + (1 if dd == d2 else 0), tmp2.shape[a] - 1 + (1 if dd == d2 else 0))
                           for a, dd in ((a, ndim - 1 - a) for a in range(ndim)))
                hh = tuple(slice(1 - (1 if dd == d2 else 0), tmp2.shape[a] - 1 - (1 if dd == d2 else 0))
                           for a, dd in ((a, ndim - 1 - a) for a in range(ndim)))
                v = t + wavefilter * (tmp2[jj] + tmp2[hh])
                denominator = denominator + v * v
        error = error + weight * np.sqrt(numerator / np.maximum(denominator, epsilon))
    return error


# ---------------------------------------------------------------------------
# regrid decisions: errest (src/amr/mod_errest.t:10-61), proper_nesting and
# the coarsening checks (src/amr/mod_coarsen_refine.t:175-320)
# ---------------------------------------------------------------------------
@dataclass
class AMRPar:
    """The meshlist parameters of the regrid (defaults of
    src/io/mod_input_output.t:370-390)."""
    refine_max_level: int = 1
    refine_threshold: float = 0.1
    derefine_ratio: float = 1.0 / 8.0
    amr_wavefilter: float = 1.0e-2
    w_refine_weight: dict = field(default_factory=lambda: {0: 1.0})
    ditregrid: int = 1


def errest_flags(forest: AMRForest, w: np.ndarray, ap: AMRPar) -> tuple[np.ndarray, np.ndarray]:
    """Refine / coarsen flags per leaf from the Löhner estimate of the
    ghosted state w[nb, nw, ...]."""
    nb = forest.nblocks
    refine = np.zeros(nb, dtype=bool)
    coarsen = np.zeros(nb, dtype=bool)
    for n in forest.leaves:
        err = lohner_error(w[n.block], forest.ndim, forest.nghost, ap.w_refine_weight, ap.amr_wavefilter)
        thr = ap.refine_threshold
        rflag = err >= thr
        cflag = (~rflag) & (err <= ap.derefine_ratio * thr)
        if rflag.any() and n.level < ap.refine_max_level:
            refine[n.block] = True
        if cflag.all() and n.level > 1:
            coarsen[n.block] = True
    return refine, coarsen


def proper_nesting(forest: AMRForest, ap: AMRPar, refine: np.ndarray, coarsen: np.ndarray) -> None:
    """2:1 balance (proper_nesting, src/amr/mod_coarsen_refine.t:186-319,
    native): refining a block refines its parent's outward leaf neighbours;
    a coarsening needs all siblings flagged and no neighbour refining or
    staying finer (flags updated in place).  The coarsening checks visit the
    leaves of a level in the order the forest created them."""
    amr_native.proper_nesting(forest.ndim, forest.ng1, forest.periodic, forest.level, forest.ig, forest.leaf_order,
                              ap.refine_max_level, refine, coarsen)


def apply_regrid(forest: AMRForest, refine: np.ndarray, coarsen: np.ndarray) -> list[tuple]:
    """Change the forest (amr_coarsen_refine: coarsen first, then refine) and
    return, per new leaf in Morton order, where its data comes from:
    ('copy', old) | ('coarsen', [old children in ic order]) |
    ('refine', old parent, ic)."""
    old_leaves = list(forest.leaves)
    origin: dict[int, tuple] = {}
    done_parents = set()
    for n in old_leaves:
        if coarsen[n.block] and id(n.parent) not in done_parents:
            p = n.parent
            done_parents.add(id(p))
            kids = [c.block for c in p.children]
            forest.coarsen_node(p)
            origin[id(p)] = ("coarsen", kids)
    for n in old_leaves:
        if refine[n.block]:
            kids = forest.refine_node(n)
            for c in kids:
                origin[id(c)] = ("refine", n.block, c.ic())
    for n in old_leaves:
        if not coarsen[n.block] and not refine[n.block]:
            origin[id(n)] = ("copy", n.block)
    forest.update()
    return [origin[id(n)] for n in forest.leaves]


# ---------------------------------------------------------------------------
# regrid data: coarsen_grid into the parent, prolong_2nd into the children
# (src/amr/mod_coarsen.t:12-79, src/amr/mod_refine.t:43-167), on ghosted
# blocks [nw, (k,) j, i]
# ---------------------------------------------------------------------------
def coarsen_child(wfi: np.ndarray, ic: tuple[int, ...], wco: np.ndarray, ndim: int, nghost: int) -> None:
    ng = nghost
    bnx = [wfi.shape[ndim - d] - 2 * ng for d in range(ndim)]  # x first
    sl_co = []
    for a in range(ndim):
        d = ndim - 1 - a
        lo = ng + (ic[d] - 1) * bnx[d] // 2
        sl_co.append(slice(lo, lo + bnx[d] // 2))
    fine = wfi[(slice(None),) + tuple(slice(ng, ng + bnx[ndim - 1 - a]) for a in range(ndim))]
    # the 2^ndim fine cells summed in array-element order (x fastest)
    acc = None
    for o in itertools.product(*[(0, 1)] * ndim):  # numpy axis order: x last
        part = fine[(slice(None),) + tuple(slice(o[a], None, 2) for a in range(ndim))]
        acc = part.copy() if acc is None else acc + part
    wco[(slice(None),) + tuple(sl_co)] = acc * (1.0 / float(2 ** ndim))


def prolong_2nd(wco: np.ndarray, ic: tuple[int, ...], wfi: np.ndarray, ndim: int, nghost: int) -> None:
    ng = nghost
    bnx = [wco.shape[ndim - d] - 2 * ng for d in range(ndim)]
    # coarse range of child ic with a one-cell rim for the slopes
    def co_slice(shift_d=None, shift=0):
        out = []
        for a in range(ndim):
            d = ndim - 1 - a
            lo = ng + (ic[d] - 1) * bnx[d] // 2 + (shift if d == shift_d else 0)
            out.append(slice(lo, lo + bnx[d] // 2))
        return (slice(None),) + tuple(out)
    c0 = wco[co_slice()]
    slopes = []
    for d in range(ndim):
        slope_l = c0 - wco[co_slice(d, -1)]
        slope_r = wco[co_slice(d, 1)] - c0
        slope_c = 0.5 * (slope_r + slope_l)
        sign_c = np.copysign(1.0, slope_c)
        slopes.append(sign_c * np.maximum(0.0, np.minimum(np.minimum(np.abs(slope_c), sign_c * slope_l),
                                                          sign_c * slope_r)))
    for o in itertools.product(*[(0, 1)] * ndim):  # numpy axis order
        v = c0.copy()
        for d in range(ndim):
            eta = 0.5 * (float(o[ndim - 1 - d]) - 0.5)
            v = v + slopes[d] * eta
        sl = tuple(slice(ng + o[a], ng + bnx[ndim - 1 - a], 2) for a in range(ndim))
        wfi[(slice(None),) + sl] = v


# ---------------------------------------------------------------------------
# time integration with regrid (src/amrvac.t:189-355, settree / resettree
# src/amr/mod_amr_grid.t:13-74, amr_coarsen_refine
# src/amr/mod_coarsen_refine.t:15-173)
# ---------------------------------------------------------------------------
from .host import Par, TimeIntegrator  # noqa: E402


class AMRIntegrator(TimeIntegrator):
    """The reference's time loop over an AMR forest.  ``make_backend(par,
    forest)`` returns a backend (the host protocol plus
    ``set_fix_conserve``) for the current leaves; ``init_fn(forest, blocks)``
    returns the ghosted conserved initial state of those leaves.  With
    ``restart`` the forest and init_fn come from a snapshot and the initial
    refinement is skipped (the reference's restart reads the tree instead of
    running settree unless reset_grid, src/amrvac.t:47-95)."""

    def __init__(self, par: Par, forest: AMRForest, ap: AMRPar, make_backend, init_fn, restart: bool = False):
        self.forest = forest
        self.ap = ap
        self.make_backend = make_backend
        self.init_fn = init_fn
        self.regrids = 0
        be = make_backend(par, forest)
        super().__init__(par, forest, be)
        # initlevelone + settree: refine on the initial condition
        be.set_state(0, init_fn(forest, list(range(forest.nblocks))))
        be.getbc(0)
        if restart:
            return
        for levnew in range(2, ap.refine_max_level + 1):
            self.regrid(time_advance=False)
            if forest.levmax != levnew:
                break

    @property
    def mesh(self):
        return self.forest

    @mesh.setter
    def mesh(self, value):
        pass

    def regrid(self, time_advance: bool = True, flags=None) -> None:
        """errest + amr_coarsen_refine (+ getbc) on ps.  A backend with a
        device regrid (amr_device_regrid) estimates and moves the data
        without a round trip through the host; the new leaves of the
        initial tree (time_advance False) take the initial condition.
        flags: the (refine, coarsen) estimate of ps if already made."""
        forest, par = self.forest, self.par
        device = getattr(self.be, "amr_device_regrid", False)
        if flags is not None:
            refine, coarsen = flags
        elif device:
            refine, coarsen = self.be.errest_flags(forest, self.ap)
        else:
            w = self.be.get_state(0)
            refine, coarsen = errest_flags(forest, w, self.ap)
        if refine.any() or coarsen.any():
            proper_nesting(forest, self.ap, refine, coarsen)
        if not refine.any() and not coarsen.any():
            return
        if device and time_advance:
            plan = apply_regrid(forest, refine, coarsen)
            self.be.regrid(plan, forest)
            self.regrids += 1
            return
        if device:
            w = self.be.get_state(0)
        plan = apply_regrid(forest, refine, coarsen)
        nd, ng = forest.ndim, forest.nghost
        new = np.zeros((forest.nblocks,) + w.shape[1:])
        init_blocks = []
        for b, org in enumerate(plan):
            if org[0] == "copy":
                new[b] = w[org[1]]
            elif org[0] == "coarsen":
                for j, ob in enumerate(org[1]):
                    ic = tuple(1 + ((j >> d) & 1) for d in range(nd))
                    coarsen_child(w[ob], ic, new[b], nd, ng)
            elif time_advance:
                prolong_2nd(w[org[1]], org[2], new[b], nd, ng)
            else:
                init_blocks.append(b)
        if init_blocks:
            new[init_blocks] = self.init_fn(forest, init_blocks)
        self.be = self.make_backend(par, forest)
        self.be.set_state(0, new)
        self.be.getbc(0)
        self.regrids += 1

    def advect(self) -> None:
        fix = self.forest.levmax > self.forest.levmin
        par, st = self.par, self.state
        if par.integrator == "Predictor_Corrector":
            from .host import FS
            self.be.set_fix_conserve(False)
            self.be.advect1(FS[par.typepred1], 0.5, st.dt, st.cmax_global, 0, 1, ("copy", 0))
            self.be.set_fix_conserve(fix)
            self.be.advect1(FS[par.flux_scheme], 1.0, st.dt, st.cmax_global, 1, 0, None)
        else:
            self.be.set_fix_conserve(fix)
            super().advect()

    def step(self) -> None:
        super().step()
        if self.ap.refine_max_level > 1 and self.state.it % self.ap.ditregrid == 0:
            self.regrid()

    def run(self, nsteps: int | None = None):
        par, st = self.par, self.state
        steps = 0
        while True:
            self.setdt()
            if par.dtsave_log < 1e99 and self._log_due():
                self.log_line()
            if (st.it >= par.it_max or st.global_time >= par.time_max
                    or st.final_dt_exit or (nsteps is not None and steps >= nsteps)):
                break
            self.advect()
            st.dts.append(st.dt)
            # the device estimate's flag copy also brings the small-value
            # flag, so the crash check after it waits on nothing more
            regrid_due = self.ap.refine_max_level > 1 and (st.it + 1) % self.ap.ditregrid == 0
            flags = self.be.errest_flags(self.forest, self.ap) \
                if regrid_due and getattr(self.be, "amr_device_regrid", False) else None
            if st.crash or self.be.crash():
                st.crash = True
                raise RuntimeError("Error: small value encountered, run crash.")
            if regrid_due:
                self.regrid(flags=flags)
            st.it += 1
            st.global_time = st.global_time + st.dt
            steps += 1
        return st
