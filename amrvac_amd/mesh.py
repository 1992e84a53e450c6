"""Single-level block mesh: block list, neighbour table and block extents.

This is the host-side connectivity the reference keeps in its forest
(src/amr/mod_forest.t, src/mod_connectivity.t): for a uniform grid every
block sits on level 1, blocks are numbered along a Morton curve
(src/amr/mod_space_filling_curve.t) and a periodic side wraps the neighbour
search (src/amr/mod_amr_neighbors.t:24-27), so periodic faces are plain
same-level copies.

The neighbour table has 3**ndim entries per block, indexed
``n = (i1+1) + 3*(i2+1) + 9*(i3+1)`` for the direction ``(i1, i2, i3)`` in
{-1,0,1}**ndim; a negative entry marks a physical boundary.
"""
from __future__ import annotations

import itertools
from dataclasses import dataclass, field

import numpy as np

# src/mod_global_parameters.t:541-550
BC_CONT = 2
BC_SYMM = 3
BC_ASYMM = 4
BC_PERIODIC = 5

_BC_CODES = {"cont": BC_CONT, "symm": BC_SYMM, "asymm": BC_ASYMM,
             "periodic": BC_PERIODIC}


def bc_code(name: str) -> int:
    try:
        return _BC_CODES[name]
    except KeyError as exc:
        raise ValueError(f"unsupported boundary type {name!r}") from exc


def _morton_key(ig: tuple[int, ...]) -> int:
    key = 0
    nbits = 21
    for bit in range(nbits):
        for d, v in enumerate(ig):
            key |= ((v >> bit) & 1) << (bit * len(ig) + d)
    return key


@dataclass
class UniformMesh:
    """A uniform (refine_max_level=1) block mesh.

    Mirrors the level-1 layout of the reference: ``dx = (xmax-xmin)/nx``
    (src/io/mod_input_output.t:1610-1625), ``dg = dx*block_nx`` and
    ``rpxmin = xprobmin + (ig-1)*dg`` (src/amr/mod_amr_solution_node.t:
    164-167, src/mod_initialize.t:104-110).
    """

    ndim: int
    domain_nx: tuple[int, ...]
    block_nx: tuple[int, ...]
    xprobmin: tuple[float, ...]
    xprobmax: tuple[float, ...]
    bc: tuple[str, ...]  # 2*ndim entries: min1, max1, min2, max2, ...
    nghost: int = 2
    ig: np.ndarray = field(init=False)
    neighbor: np.ndarray = field(init=False)
    xlohi: np.ndarray = field(init=False)
    dx: np.ndarray = field(init=False)

    def __post_init__(self) -> None:
        nd = self.ndim
        if not (len(self.domain_nx) == len(self.block_nx) == len(self.xprobmin)
                == len(self.xprobmax) == nd and len(self.bc) == 2 * nd):
            raise ValueError("mesh arguments must have ndim entries (bc: 2*ndim)")
        for d in range(nd):
            if self.domain_nx[d] % self.block_nx[d] != 0:
                raise ValueError("Grid (domain_nx) and block (block_nx) must be consistent")
            lo_p = self.bc[2 * d] == "periodic"
            hi_p = self.bc[2 * d + 1] == "periodic"
            if lo_p != hi_p:
                raise ValueError("periodic boundaries must be paired")
        self.dx = np.array([(self.xprobmax[d] - self.xprobmin[d]) / self.domain_nx[d]
                            for d in range(nd)], dtype=np.float64)
        ng = [self.domain_nx[d] // self.block_nx[d] for d in range(nd)]
        self.ng_blocks = tuple(ng)
        all_ig = list(itertools.product(*[range(1, n + 1) for n in ng]))
        all_ig.sort(key=lambda t: _morton_key(tuple(v - 1 for v in t)))
        self.ig = np.array(all_ig, dtype=np.int32).reshape(-1, nd)
        index = {tuple(int(v) for v in row): i for i, row in enumerate(self.ig)}
        nb = len(all_ig)
        nnb = 3 ** nd
        nbr = np.full((nb, nnb), -1, dtype=np.int32)
        periodic = [self.bc[2 * d] == "periodic" for d in range(nd)]
        for ib, row in enumerate(self.ig):
            for n, dirs in enumerate(itertools.product(*([(-1, 0, 1)] * nd))):
                # itertools.product varies the last axis fastest; the table
                # index has axis 0 fastest
                dirs = dirs[::-1]
                idx = sum((dirs[d] + 1) * 3 ** d for d in range(nd))
                tgt = []
                ok = True
                for d in range(nd):
                    v = int(row[d]) + dirs[d]
                    if v < 1 or v > ng[d]:
                        if periodic[d]:
                            v = (v - 1) % ng[d] + 1
                        else:
                            ok = False
                            break
                    tgt.append(v)
                nbr[ib, idx] = index[tuple(tgt)] if ok else -1
        self.neighbor = nbr
        dg = [self.dx[d] * float(self.block_nx[d]) for d in range(nd)]
        xl = np.zeros((nb, nd, 2), dtype=np.float64)
        for ib, row in enumerate(self.ig):
            for d in range(nd):
                lo = self.xprobmin[d] + float(row[d] - 1) * dg[d]
                hi = self.xprobmin[d] + float(row[d]) * dg[d]
                if hi > self.xprobmax[d]:
                    hi = self.xprobmax[d]
                xl[ib, d, 0] = lo
                xl[ib, d, 1] = hi
        self.xlohi = xl

    @property
    def nblocks(self) -> int:
        return int(self.ig.shape[0])

    @property
    def bc_codes(self) -> tuple[int, ...]:
        return tuple(bc_code(b) for b in self.bc)

    @property
    def nxg(self) -> tuple[int, ...]:
        """Block shape including ghost cells (ixG), x first."""
        return tuple(self.block_nx[d] + 2 * self.nghost for d in range(self.ndim))

    def cell_centers(self) -> np.ndarray:
        """x[ib, d, (k,) j, i] cell centres in the reference layout
        (src/amr/mod_amr_solution_node.t:176-186), ghost cells included."""
        nd = self.ndim
        ng = self.nghost
        shp = self.nxg[::-1]  # numpy order: slowest first
        out = np.zeros((self.nblocks, nd) + shp, dtype=np.float64)
        for d in range(nd):
            nxd = self.nxg[d]
            bn = self.block_nx[d]
            ix = np.arange(1, nxd + 1)
            for ib in range(self.nblocks):
                lo, hi = self.xlohi[ib, d]
                v = np.where(ix <= bn,
                             lo + ((ix - ng).astype(np.float64) - 0.5) * self.dx[d],
                             hi + ((ix - (bn + ng)).astype(np.float64) - 0.5) * self.dx[d])
                bshape = [1] * nd
                bshape[nd - 1 - d] = nxd
                out[ib, d] = np.broadcast_to(v.reshape(bshape), shp)
        return out

    def global_index(self) -> np.ndarray:
        """Global 0-based cell index per axis for every cell (ghosts may fall
        outside [0, domain_nx)); shape [nb, nd, (k,) j, i]."""
        nd = self.ndim
        shp = self.nxg[::-1]
        out = np.zeros((self.nblocks, nd) + shp, dtype=np.int64)
        for d in range(nd):
            nxd = self.nxg[d]
            loc = np.arange(nxd) - self.nghost
            for ib in range(self.nblocks):
                g = (int(self.ig[ib, d]) - 1) * self.block_nx[d] + loc
                bshape = [1] * nd
                bshape[nd - 1 - d] = nxd
                out[ib, d] = np.broadcast_to(g.reshape(bshape), shp)
        return out

    def interior(self, w: np.ndarray) -> np.ndarray:
        """View of the interior cells of a [nb, nw, (k,) j, i] array."""
        ng = self.nghost
        sl = tuple(slice(ng, ng + self.block_nx[d]) for d in range(self.ndim))[::-1]
        return w[(slice(None), slice(None)) + sl]

    def to_global(self, w: np.ndarray) -> np.ndarray:
        """Assemble interiors into one [nw, (nz,) ny, nx] array."""
        nd = self.ndim
        nw = w.shape[1]
        out = np.zeros((nw,) + tuple(self.domain_nx[::-1]), dtype=w.dtype)
        inner = self.interior(w)
        for ib in range(self.nblocks):
            sl = tuple(slice((int(self.ig[ib, d]) - 1) * self.block_nx[d],
                             int(self.ig[ib, d]) * self.block_nx[d]) for d in range(nd))[::-1]
            out[(slice(None),) + sl] = inner[ib]
        return out
