"""The native AMR tree routines (amrvac_amd/csrc/amr_tree.cpp: neighbour
tables and the 2:1 balance of the regrid flags) against their Python
restatement (tests/amr_tree_ref.py), on forests grown and pruned by random
regrids in 2-D and 3-D, periodic and bounded, and with random flags for the
balance."""
from __future__ import annotations

import numpy as np
import pytest

from amrvac_amd import amr_native
from amrvac_amd.amr import AMRForest, AMRPar, apply_regrid, proper_nesting

import amr_tree_ref as ref

CASES = [
    (2, (64, 64), (16, 16), ("periodic",) * 4, 4),
    (2, (48, 32), (16, 16), ("cont",) * 4, 4),
    (2, (64, 32), (16, 16), ("periodic", "periodic", "cont", "cont"), 5),
    (3, (32, 32, 32), (8, 8, 8), ("periodic",) * 6, 3),
    (3, (24, 16, 16), (8, 8, 8), ("cont", "cont", "periodic", "periodic", "cont", "cont"), 3),
]


def _forests(case, seed, nregrid):
    """Yield the forest after each of nregrid random regrids."""
    nd, nx, bx, bc, lmax = case
    rng = np.random.default_rng(seed)
    forest = AMRForest(nd, nx, bx, (0.0,) * nd, (1.0,) * nd, bc, max_level=lmax)
    ap = AMRPar(refine_max_level=lmax)
    for _ in range(nregrid):
        nb = forest.nblocks
        refine = (rng.random(nb) < 0.15) & (forest.level < lmax)
        coarsen = (rng.random(nb) < 0.5) & (forest.level > 1)  # as the estimator flags
        proper_nesting(forest, ap, refine, coarsen)
        if refine.any() or coarsen.any():
            apply_regrid(forest, refine, coarsen)
        yield forest, ap, rng


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}d-{c[1]}-{c[3][0]}-{c[3][-1]}")
def test_native_connectivity_matches_restatement(case):
    levels = set()
    for forest, _, _ in _forests(case, 7, 25):
        nt, nb, nc = ref.connectivity(forest)
        assert np.array_equal(forest.ntype, nt)
        assert np.array_equal(forest.nbr, nb)
        assert np.array_equal(forest.nchild, nc)
        levels.add(forest.levmax - forest.levmin)
    assert max(levels) >= 2  # the forests did span several levels


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{c[0]}d-{c[1]}-{c[3][0]}-{c[3][-1]}")
def test_native_proper_nesting_matches_restatement(case):
    changed = 0
    for forest, ap, rng in _forests(case, 11, 20):
        for _ in range(4):
            nb = forest.nblocks
            r0 = (rng.random(nb) < rng.uniform(0.05, 0.4)) & (forest.level < ap.refine_max_level)
            c0 = (rng.random(nb) < rng.uniform(0.3, 0.95)) & (forest.level > 1)
            r1, c1 = r0.copy(), c0.copy()
            r2, c2 = r0.copy(), c0.copy()
            proper_nesting(forest, ap, r1, c1)
            ref.proper_nesting(forest, ap, r2, c2)
            assert np.array_equal(r1, r2) and np.array_equal(c1, c2)
            changed += int((r1 != r0).sum() + (c1 != c0).sum())
    assert changed > 0  # the balance did change flags


def test_level_jump_of_two_is_rejected():
    """A leaf two levels finer than its neighbour is not a valid tree."""
    forest = AMRForest(2, (64, 64), (16, 16), (0.0, 0.0), (1.0, 1.0), ("periodic",) * 4, max_level=4)
    n = forest.leaves[0]
    kids = forest.refine_node(n)
    forest.refine_node(kids[0])
    with pytest.raises(RuntimeError, match="level jump"):
        forest.update()


def test_bad_leaf_sets_are_rejected():
    with pytest.raises(ValueError, match="two leaves"):
        amr_native.connectivity(2, (4, 4), (1, 1), np.array([1, 1]), np.array([[1, 1], [1, 1]]))
    with pytest.raises(ValueError, match="below another leaf"):
        amr_native.connectivity(2, (4, 4), (1, 1), np.array([1, 2]), np.array([[1, 1], [1, 1]]))
    with pytest.raises(ValueError, match="out of range"):
        amr_native.connectivity(2, (4, 4), (1, 1), np.array([1]), np.array([[5, 1]]))
