"""Pins the neighbor-list oracle (CPU)."""

import math

import numpy as np
import pytest
from scipy.spatial import cKDTree

from oracle import nblist
from torchff.workloads import make_workload


def _uniform(n, density, seed):
    L = (n / density) ** (1 / 3)
    rng = np.random.default_rng(seed)
    return rng.uniform(0, L, (n, 3)), np.eye(3) * L


def _as_set(p):
    return set(map(tuple, np.asarray(p).tolist()))


@pytest.mark.parametrize("n,rc", [(2000, 0.5), (3000, 0.8)])
def test_cell_search_equals_brute_force(n, rc):
    x, box = _uniform(n, 100.0, n)
    a = nblist.brute_force(x, box, rc)
    b = nblist.cell_search(x, box, rc)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("n,rc", [(1500, 0.6), (800, 0.45)])
def test_periodic_set_matches_kdtree(n, rc):
    # independent library routine: scipy's periodic k-d tree
    x, box = _uniform(n, 100.0, 7 * n)
    L = box[0, 0]
    tree = cKDTree(np.mod(x, L), boxsize=L)
    ref = np.array(sorted(tree.query_pairs(rc)), dtype=np.int64).reshape(-1, 2)
    got = nblist.neighbor_pairs(x, box, rc)
    assert _as_set(got) == _as_set(ref)


def test_non_periodic_matches_kdtree():
    x, _ = _uniform(600, 100.0, 3)
    ref = np.array(sorted(cKDTree(x).query_pairs(0.5)), dtype=np.int64)
    got = nblist.neighbor_pairs(x, None, 0.5)
    assert _as_set(got) == _as_set(ref)


def test_zero_and_huge_cutoff():
    x, box = _uniform(30, 100.0, 5)
    assert len(nblist.neighbor_pairs(x, box, 0.0)) == 0
    big = nblist.neighbor_pairs(x, box, box[0, 0] * 10)
    assert len(big) == 30 * 29 // 2


def test_include_self():
    x, box = _uniform(50, 100.0, 11)
    p = nblist.brute_force(x, box, 0.5, include_self=True)
    assert np.sum(p[:, 0] == p[:, 1]) == 50
    q = nblist.brute_force(x, box, 0.5, include_self=False)
    assert len(p) == len(q) + 50


def test_count_matches_density():
    n, rho, rc = 20000, 100.0, 0.8
    x, box = _uniform(n, rho, 99)
    p = nblist.cell_search(x, box, rc)
    expect = 0.5 * n * rho * 4 / 3 * math.pi * rc ** 3
    assert abs(len(p) / expect - 1) < 0.02


def test_water_exclusions_removed_and_csr():
    w = make_workload("water900")
    excl = w["exclusions"]
    rp, col = nblist.coo_to_csr(excl, w["natoms"])
    # each atom excludes exactly its own water (3 entries including itself)
    assert np.all(np.diff(rp) == 3)
    rows = np.repeat(np.arange(w["natoms"]), 3)
    assert np.all(col // 3 == rows // 3)
    p = nblist.neighbor_pairs(w["coords"], w["box"], w["cutoff"], exclusions=excl)
    assert not np.any(p[:, 0] // 3 == p[:, 1] // 3)
    allp = nblist.neighbor_pairs(w["coords"], w["box"], w["cutoff"])
    intra = allp[allp[:, 0] // 3 == allp[:, 1] // 3]
    assert len(intra) == 3 * w["n_waters"]
    assert len(allp) - len(intra) == len(p)


def test_canonicalize_drops_padding():
    raw = np.array([[5, 2], [-1, -1], [1, 3], [-1, -1]])
    np.testing.assert_array_equal(nblist.canonicalize(raw), [[1, 3], [2, 5]])
    np.testing.assert_array_equal(nblist.canonicalize(raw, npairs=1), [[2, 5]])
