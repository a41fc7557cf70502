"""Pins the rank-1 (point dipole) PME and Ewald oracles (CPU)."""

import math

import numpy as np
import torch

from oracle import ewald as oe
from oracle import pme as op
from torchff.workloads import water_box


def _system(n_waters=12, L=1.3, seed=4):
    w = water_box(n_waters, L, seed=seed)
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, 0.02, (w["natoms"], 3))
    t = lambda a: torch.tensor(a, dtype=torch.float64)
    return t(w["coords"]), t(w["box"]), t(w["charges"]), t(p)


def test_pme_dipoles_converge_to_direct_ewald():
    x, box, q, p = _system()
    alpha = 3.0
    ref = float(oe.ewald_recip_energy(x, box, q, 10, alpha, p=p))
    errs = []
    for K, order in ((16, 6), (24, 8), (32, 8)):
        e = float(op.pme_energy_dipoles(x, box, q, p, K, alpha, order))
        errs.append(abs(e - ref) / abs(ref))
    assert errs[-1] < 1e-7 and errs[-1] < errs[0], errs


def test_zero_dipoles_reduce_to_charges():
    x, box, q, _ = _system()
    z = torch.zeros_like(x)
    assert abs(float(op.pme_energy_dipoles(x, box, q, z, 20, 3.0, 6)) - float(op.pme_energy(x, box, q, 20, 3.0, 6))) \
        < 1e-12 * abs(float(op.pme_energy(x, box, q, 20, 3.0, 6)))
    assert abs(float(oe.ewald_recip_energy(x, box, q, 8, 3.0, p=z)) - float(oe.ewald_recip_energy(x, box, q, 8, 3.0))) \
        < 1e-12


def test_dipole_is_the_limit_of_a_charge_pair():
    # a point dipole p at r equals charges +p/d at r + d/2 and -p/d at r - d/2 as d -> 0, once the
    # pair's own smooth mutual term -(p/d)^2 erf(alpha d)/d (which the reciprocal sum contains and
    # the self term of a point dipole removes) is added back
    x, box, q, _ = _system(8, 1.2, 7)
    alpha, kmax = 3.2, 10
    p0 = torch.tensor([0.011, -0.004, 0.007], dtype=torch.float64)
    r0 = torch.tensor([0.31, 0.52, 0.43], dtype=torch.float64)
    xd = torch.cat([x, r0[None]])
    qd = torch.cat([q, torch.zeros(1, dtype=torch.float64)])
    pd = torch.zeros_like(xd)
    pd[-1] = p0
    e_dip = float(oe.ewald_recip_energy(xd, box, qd, kmax, alpha, p=pd))
    errs = []
    for dl in (2e-3, 1e-3):
        dvec = p0 / torch.linalg.vector_norm(p0) * dl
        qq = float(torch.linalg.vector_norm(p0)) / dl
        xp = torch.cat([x, (r0 + 0.5 * dvec)[None], (r0 - 0.5 * dvec)[None]])
        qp = torch.cat([q, torch.tensor([qq, -qq], dtype=torch.float64)])
        e_pair = float(oe.ewald_recip_energy(xp, box, qp, kmax, alpha)) + qq * qq * math.erf(alpha * dl) / dl
        errs.append(abs(e_pair - e_dip))
    # the difference vanishes like d^2
    assert errs[1] < 0.3 * errs[0] and errs[1] < 1e-5 * abs(e_dip), errs
    # and PME reproduces the point-dipole energy
    e_pme = float(op.pme_energy_dipoles(xd, box, qd, pd, 40, alpha, 8))
    assert abs(e_pme - e_dip) < 1e-7 * abs(e_dip)


def test_dipole_energy_derivatives_by_finite_differences():
    x, box, q, p = _system(6, 1.1, 2)
    alpha, K, order = 3.0, 20, 6
    pt = p.clone().requires_grad_(True)
    xt = x.clone().requires_grad_(True)
    E = op.pme_energy_dipoles(xt, box, q, pt, K, alpha, order)
    gx, gp = torch.autograd.grad(E, [xt, pt])
    h = 1e-5
    for (arr, g) in ((x, gx), (p, gp)):
        for (i, a) in ((0, 0), (5, 2), (11, 1)):
            ap = arr.clone(); ap[i, a] += h
            am = arr.clone(); am[i, a] -= h
            if arr is x:
                ep = op.pme_energy_dipoles(ap, box, q, p, K, alpha, order)
                em = op.pme_energy_dipoles(am, box, q, p, K, alpha, order)
            else:
                ep = op.pme_energy_dipoles(x, box, q, ap, K, alpha, order)
                em = op.pme_energy_dipoles(x, box, q, am, K, alpha, order)
            fd = float(ep - em) / (2 * h)
            assert abs(fd - float(g[i, a])) < 1e-6 * max(1.0, abs(fd)), (i, a, fd, float(g[i, a]))
