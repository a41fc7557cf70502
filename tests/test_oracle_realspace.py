"""Pins the real-space oracles (CPU)."""

import math

import numpy as np
import torch

from oracle import realspace
from torchff.workloads import make_workload, water_box

K_C = 138.935456


def _pair_system(r, box_len=10.0):
    coords = torch.tensor([[1.0, 1.0, 1.0], [1.0 + r, 1.0, 1.0]], dtype=torch.float64)
    box = torch.eye(3, dtype=torch.float64) * box_len
    pairs = torch.tensor([[0, 1]])
    return coords, box, pairs


def test_lj_minimum_is_minus_epsilon():
    sig, eps = 0.31507, 0.6364
    r = 2 ** (1 / 6) * sig
    coords, box, pairs = _pair_system(r)
    s = torch.tensor([sig, sig], dtype=torch.float64)
    e = torch.tensor([eps, eps], dtype=torch.float64)
    q = torch.zeros(2, dtype=torch.float64)
    E = realspace.nonbonded_energy(coords, pairs, box, s, e, q, K_C, 1.0, do_shift=True)
    assert abs(float(E) + eps) < 1e-12
    # and it is a minimum: zero force
    c = coords.clone().requires_grad_(True)
    g, = torch.autograd.grad(realspace.nonbonded_energy(c, pairs, box, s, e, q, K_C, 1.0), c)
    assert float(g.abs().max()) < 1e-10


def test_shifted_coulomb_pair():
    r, rc = 0.37, 0.9
    coords, box, pairs = _pair_system(r)
    q = torch.tensor([0.7, -0.3], dtype=torch.float64)
    z = torch.zeros(2, dtype=torch.float64)
    E = realspace.nonbonded_energy(coords, pairs, box, z + 0.3, z, q, K_C, rc, do_shift=True)
    assert abs(float(E) - K_C * 0.7 * -0.3 * (1 / r - 1 / rc)) < 1e-11
    E2 = realspace.nonbonded_energy(coords, pairs, box, z + 0.3, z, q, K_C, rc, do_shift=False)
    assert abs(float(E2) - K_C * 0.7 * -0.3 / r) < 1e-11
    # beyond the cutoff: nothing
    coords, box, pairs = _pair_system(rc * 1.0001)
    assert float(realspace.nonbonded_energy(coords, pairs, box, z + 0.3, z, q, K_C, rc)) == 0.0


def test_exclusion_correction_is_minus_erf():
    r, a = 0.1514, 2.92
    coords, box, pairs = _pair_system(r)
    q = torch.tensor([0.417, 0.417], dtype=torch.float64)
    empty = torch.zeros((0, 2), dtype=torch.int64)
    E = realspace.ewald_real_energy(coords, box, empty, pairs, q, 0.9, a, K_C)
    assert abs(float(E) + K_C * 0.417 ** 2 * math.erf(a * r) / r) < 1e-10
    # alpha = 0 drops the correction, alpha < 0 means bare Coulomb on the pair list
    assert float(realspace.ewald_real_energy(coords, box, empty, pairs, q, 0.9, 0.0, K_C)) == 0.0
    E3 = realspace.ewald_real_energy(coords, box, pairs, None, q, 0.9, -1.0, K_C)
    assert abs(float(E3) - K_C * 0.417 ** 2 / r) < 1e-10


def test_padding_rows_skipped():
    w = make_workload("water900")
    from oracle import nblist
    p = nblist.neighbor_pairs(w["coords"], w["box"], 0.9, exclusions=w["exclusions"])
    pad = np.concatenate([p, -np.ones((100, 2), np.int64)])
    c = torch.tensor(w["coords"])
    b = torch.tensor(w["box"])
    q = torch.tensor(w["charges"])
    e1 = realspace.ewald_real_energy(c, b, torch.tensor(p), None, q, 0.9, w["alpha"], K_C)
    e2 = realspace.ewald_real_energy(c, b, torch.tensor(pad), None, q, 0.9, w["alpha"], K_C)
    assert float(e1) == float(e2)


def test_type_table_equals_per_atom_mixing():
    w = make_workload("water900")
    from oracle import nblist
    p = torch.tensor(nblist.neighbor_pairs(w["coords"], w["box"], 0.9, exclusions=w["exclusions"]))
    c = torch.tensor(w["coords"])
    b = torch.tensor(w["box"])
    e_tab = realspace.lj_energy(c, p, b, torch.tensor(w["sigma_table"]), torch.tensor(w["epsilon_table"]),
                                0.9, atom_types=torch.tensor(w["atom_types"]))
    z = torch.zeros(w["natoms"], dtype=torch.float64)
    e_atom = realspace.nonbonded_energy(c, p, b, torch.tensor(w["sigma"]), torch.tensor(w["epsilon"]), z,
                                        K_C, 0.9, do_shift=False)
    assert abs(float(e_tab) - float(e_atom)) < 1e-9 * abs(float(e_tab))


def test_forces_match_finite_differences():
    w = water_box(20, 1.5, seed=3)
    from oracle import nblist
    rc = 0.7
    p = torch.tensor(nblist.neighbor_pairs(w["coords"], w["box"], rc, exclusions=w["exclusions"]))
    b = torch.tensor(w["box"])
    q = torch.tensor(w["charges"])
    sg = torch.tensor(w["sigma"])
    ep = torch.tensor(w["epsilon"])
    excl = torch.tensor(w["pairs_excl"])

    def energy(x):
        return realspace.ewald_lj_energy(x, p, b, sg, ep, q, K_C, rc, 3.0, excl)

    x = torch.tensor(w["coords"], requires_grad=True)
    g, = torch.autograd.grad(energy(x), x)
    h = 1e-5
    rng = np.random.default_rng(0)
    for a in rng.choice(len(x), 8, replace=False):
        for d in range(3):
            xp = x.detach().clone()
            xm = x.detach().clone()
            xp[a, d] += h
            xm[a, d] -= h
            fd = (float(energy(xp)) - float(energy(xm))) / (2 * h)
            assert abs(fd - float(g[a, d])) < 1e-5 * max(1.0, abs(fd))
