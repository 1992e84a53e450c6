"""Restatement in Python of the AMR tree routines that the package runs
natively (amrvac_amd/csrc/amr_tree.cpp): find_neighbor
(src/amr/mod_amr_neighbors.t:54-124), build_connectivity
(src/mod_functions_connectivity.t:55-431) and proper_nesting
(src/amr/mod_coarsen_refine.t:186-319), over the Node objects of an
AMRForest.  Test infrastructure: the native routines are checked against it."""
from __future__ import annotations

import itertools

import numpy as np

from amrvac_amd.amr import NB_BOUNDARY, NB_COARSE, NB_FINE, NB_SIBLING, dir_index


def node_at(forest, level: int, ig: tuple[int, ...]):
    """The node at (level, ig) with periodic wrap; 'out' outside the domain."""
    ng = forest.ng(level)
    g = list(ig)
    for d in range(forest.ndim):
        if g[d] < 1 or g[d] > ng[d]:
            if forest.periodic[d]:
                g[d] = 1 + (g[d] - 1) % ng[d]
            else:
                return "out"
    return forest.nodes.get((level, tuple(g)))


def find_neighbor(forest, node, i: tuple[int, ...]):
    """(type, node): the sibling / coarse neighbour leaf, or for a fine
    neighbour the same-level node whose children are leaves or deeper."""
    tgt = tuple(node.ig[d] + i[d] for d in range(forest.ndim))
    n = node_at(forest, node.level, tgt)
    if n == "out":
        return NB_BOUNDARY, None
    if n is None:
        # no node at this level: the parent's neighbour is a leaf
        ptgt = tuple((tgt[d] - 1) // 2 + 1 for d in range(forest.ndim))
        p = node_at(forest, node.level - 1, ptgt)
        if p == "out" or p is None:
            return NB_BOUNDARY, None
        return NB_COARSE, p
    return (NB_SIBLING if n.leaf else NB_FINE), n


def connectivity(forest):
    nd = forest.ndim
    nb = forest.nblocks
    nn = 3 ** nd
    nc = 2 ** nd
    ntype = np.zeros((nb, nn), dtype=np.int32)
    nbr = np.full((nb, nn), -1, dtype=np.int32)            # sibling / coarse leaf
    nchild = np.full((nb, nn, nc), -1, dtype=np.int32)     # fine: children of the node
    for n in forest.leaves:
        for i in itertools.product(*[(-1, 0, 1)] * nd):
            i = i[::-1]
            if all(v == 0 for v in i):
                ntype[n.block, dir_index(i)] = 0
                nbr[n.block, dir_index(i)] = n.block
                continue
            t, m = find_neighbor(forest, n, i)
            k = dir_index(i)
            ntype[n.block, k] = t
            if t in (NB_SIBLING, NB_COARSE):
                nbr[n.block, k] = m.block
            elif t == NB_FINE:
                # the children of the neighbour node that touch this block
                for j, c in enumerate(m.children):
                    icj = tuple(1 + ((j >> d) & 1) for d in range(nd))
                    if any((i[d] == -1 and icj[d] != 2) or (i[d] == 1 and icj[d] != 1) for d in range(nd)):
                        continue
                    if not c.leaf:
                        raise RuntimeError("level jump larger than one between neighbours")
                    nchild[n.block, k, j] = c.block
    return ntype, nbr, nchild


def _outward_dirs(ic: tuple[int, ...]):
    for i in itertools.product(*[range(c - 2, c) for c in ic[::-1]]):
        i = i[::-1]
        if any(v != 0 for v in i):
            yield i


def proper_nesting(forest, ap, refine: np.ndarray, coarsen: np.ndarray) -> None:
    """2:1 balance: refining a block refines its parent's outward leaf
    neighbours; a coarsening needs all siblings flagged and no neighbour
    refining or staying finer (in place)."""
    nodes_by_level: dict = {}
    for n in forest.nodes.values():
        nodes_by_level.setdefault(n.level, []).append(n)
    levmin, levmax = forest.levmin, forest.levmax
    for level in range(min(levmax, ap.refine_max_level - 1), levmin, -1):
        for n in nodes_by_level.get(level, []):
            if not n.leaf or not refine[n.block]:
                continue
            p = n.parent
            for inp in _outward_dirs(n.ic()):
                m = node_at(forest, level - 1, tuple(p.ig[d] + inp[d] for d in range(forest.ndim)))
                if m is None or m == "out":
                    continue
                if m.leaf:
                    refine[m.block] = True
    coarsen &= ~refine
    for level in range(levmax, max(2, levmin) - 1, -1):
        for n in nodes_by_level.get(level, []):
            if not n.leaf or not coarsen[n.block]:
                continue
            parent = n.parent
            sibs = parent.children
            if not all(c.leaf and coarsen[c.block] for c in sibs):
                for c in sibs:
                    if c.leaf:
                        coarsen[c.block] = False
                continue
            ok = True
            for c in sibs:
                for i in _outward_dirs(c.ic()):
                    t, m = find_neighbor(forest, c, i)
                    if t == NB_SIBLING:
                        if refine[m.block]:
                            ok = False
                    elif t == NB_FINE:
                        c1 = m.children[0]
                        if not (c1.leaf and coarsen[c1.block]):
                            ok = False
                    if not ok:
                        break
                if not ok:
                    break
            if not ok:
                for c in sibs:
                    if c.leaf:
                        coarsen[c.block] = False
