"""Pins the all-images oracle of small periodic water boxes (CPU)."""

import itertools

import numpy as np
import pytest

from oracle import composite, images
from oracle.pme import grid_size_for_spacing
from torchff.workloads import box_length_for_density, ewald_alpha, water_box


def _small_box(n_waters, seed):
    L = box_length_for_density(n_waters)
    return water_box(n_waters, L, seed=seed)


def _supercell(w, m):
    """The m x m x m periodic supercell of a water-box dict (copies at (a, b, c) L)."""
    n = w["natoms"]
    L = w["box"][0, 0]
    shifts = [np.array(s, dtype=np.float64) * L for s in itertools.product(range(m), repeat=3)]
    big = dict(w)
    big["coords"] = np.concatenate([w["coords"] + s for s in shifts])
    big["box"] = w["box"] * m
    for k in ("charges", "sigma", "epsilon", "atom_types"):
        big[k] = np.tile(w[k], len(shifts))
    big["exclusions"] = np.concatenate([w["exclusions"] + c * n for c in range(len(shifts))])
    big["pairs_excl"] = np.concatenate([w["pairs_excl"] + c * n for c in range(len(shifts))])
    big["natoms"] = n * len(shifts)
    return big


def test_image_range():
    # edge >= 2 rc: the minimum image alone; edge 0.84 < rc: one image each way
    assert images.image_range([2.0, 1.8, 1.81], 0.9) == [0, 1, 0]
    assert images.image_range([0.843] * 3, 0.9) == [1, 1, 1]
    assert images.image_range([0.5] * 3, 0.9) == [2, 2, 2]


def test_matches_neighbor_list_composition_for_large_box():
    # 300 waters, edge 2.08 nm >= 2 x 0.9: every pair has one image in range
    w = water_box(300, 2.08, n_lat=7, seed=1)
    rc, K, order = 0.9, 24, 5
    alpha = ewald_alpha(rc)
    E0, F0, _, _ = composite.energy_forces(w, cutoff=rc, alpha=alpha, K=K, order=order, skin=0.0)
    E1, F1, _ = images.energy_forces(w, rc, alpha, K, order)
    assert abs(E1 - E0) < 1e-10 * abs(E0)
    assert np.max(np.abs(F1 - F0)) < 1e-9 * np.max(np.abs(F0))


@pytest.mark.parametrize("n_waters,seed", [(20, 3), (35, 4)])
def test_small_box_equals_supercell_share(n_waters, seed):
    # 20 waters: edge 0.843 nm < rc = 0.9 < 2 rc; 35 waters: edge 1.02 nm, between rc and 2 rc
    w = _small_box(n_waters, seed)
    rc, order = 0.9, 5
    alpha = ewald_alpha(rc)
    L = w["box"][0, 0]
    K = grid_size_for_spacing(L, 0.1)
    E, F, _ = images.energy_forces(w, rc, alpha, K, order)
    big = _supercell(w, 3)
    assert 3 * L >= 2 * rc
    Eb, Fb, _, _ = composite.energy_forces(big, cutoff=rc, alpha=alpha, K=3 * K, order=order, skin=0.0)
    assert abs(Eb / 27.0 - E) < 1e-10 * abs(E), (Eb / 27.0, E)
    n = w["natoms"]
    scale = np.max(np.abs(F))
    for c in range(27):
        assert np.max(np.abs(Fb[c * n:(c + 1) * n] - F)) < 1e-9 * scale


def test_extra_images_change_nothing():
    w = _small_box(20, 5)
    rc, order = 0.9, 5
    alpha = ewald_alpha(rc)
    K = grid_size_for_spacing(w["box"][0, 0], 0.1)
    E0, F0, _ = images.energy_forces(w, rc, alpha, K, order)
    E1, F1, _ = images.energy_forces(w, rc, alpha, K, order, extra_images=1)
    assert abs(E1 - E0) < 1e-12 * abs(E0)
    assert np.max(np.abs(F1 - F0)) < 1e-12 * np.max(np.abs(F0))
