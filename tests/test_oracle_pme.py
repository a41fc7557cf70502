"""Pins the PME oracle (CPU)."""

import math

import numpy as np
import pytest
import torch

from oracle import ewald, pme
from torchff.workloads import make_workload, water_box


def _t(x):
    return torch.tensor(x, dtype=torch.float64)


def test_madelung_nacl():
    # full Ewald decomposition (real over all images + PME + self) on a rock-salt supercell
    c, q, box = ewald.nacl_supercell(2)
    alpha = 4.0
    er = ewald.real_space_all_images(_t(c), _t(box), _t(q), alpha, 1.9)
    c = c + np.array([0.013, 0.029, 0.007])  # off-grid so the spline is exercised
    ep = pme.pme_energy(_t(c), _t(box), _t(q), 64, alpha, 6)
    E = float(er + ep)
    M = -E / (len(q) / 2) * 0.5  # per ion pair, nearest-neighbour distance a/2 = 0.5
    assert abs(M - ewald.MADELUNG_NACL) < 2e-6


@pytest.mark.parametrize("order,K,tol", [(6, 24, 2e-4), (5, 24, 6e-4), (6, 36, 1e-5), (8, 36, 3e-6)])
def test_pme_converges_to_direct_ewald(order, K, tol):
    w = make_workload("water900")
    c, b, q = _t(w["coords"]), _t(w["box"]), _t(w["charges"])
    alpha = w["alpha"]
    ref = float(ewald.ewald_recip_energy(c, b, q, 14, alpha))
    got = float(pme.pme_energy(c, b, q, K, alpha, order))
    # compare the reciprocal part (self term is common); scale by the k-space magnitude
    self_e = -alpha / math.sqrt(math.pi) * float((q * q).sum())
    assert abs((got - self_e) - (ref - self_e)) < tol * abs(ref - self_e)


def test_energy_identity_and_charge_gradient():
    w = water_box(40, 1.2, seed=9)
    c, b = _t(w["coords"]), _t(w["box"])
    q = _t(w["charges"]).requires_grad_(True)
    E, phi = pme.pme_energy(c, b, q, 16, 3.1, 5, return_phi=True)
    Ek = pme.pme_energy_kspace(c, b, q, 16, 3.1, 5)
    assert abs(float(E) - float(Ek)) < 1e-12 * abs(float(E))
    g, = torch.autograd.grad(E, q)
    np.testing.assert_allclose(g.numpy(), phi.detach().numpy(), rtol=1e-10, atol=1e-12)


@pytest.mark.parametrize("order", [5, 6])
def test_forces_autograd_analytic_fd(order):
    w = water_box(30, 1.1, seed=4)
    b, q = _t(w["box"]), _t(w["charges"])
    K, alpha = 15 if order == 5 else 16, 3.0
    x = _t(w["coords"]).requires_grad_(True)
    E = pme.pme_energy(x, b, q, K, alpha, order)
    g, = torch.autograd.grad(E, x)
    Phi = pme.potential_grid(pme.spread(x.detach(), b, q, K, order), b, K, alpha, order)
    phi, F = pme.pme_gather(Phi, x.detach(), b, q, K, alpha, order)
    np.testing.assert_allclose(F.numpy(), -g.numpy(), rtol=1e-9, atol=1e-10)
    h = 1e-5
    for a in (0, 7, 50):
        for d in range(3):
            xp = x.detach().clone()
            xm = x.detach().clone()
            xp[a, d] += h
            xm[a, d] -= h
            fd = (float(pme.pme_energy(xp, b, q, K, alpha, order)) - float(pme.pme_energy(xm, b, q, K, alpha, order))) / (2 * h)
            assert abs(fd - float(g[a, d])) < 1e-6 * max(1.0, abs(fd))


def test_translation_invariance_is_approximate():
    w = make_workload("water900")
    c, b, q = _t(w["coords"]), _t(w["box"]), _t(w["charges"])
    e0 = float(pme.pme_energy(c, b, q, 24, w["alpha"], 5))
    e1 = float(pme.pme_energy(c + 0.0371, b, q, 24, w["alpha"], 5))
    assert abs(e0 - e1) < 2e-4 * abs(e0)


def test_grid_size_rule():
    assert pme.grid_size_for_spacing(2.6, 0.1) == 27
    assert pme.grid_size_for_spacing(0.843, 0.1) == 9
    assert pme.grid_size_for_spacing(3.104, 0.1) == 32
