"""Pins the multipolar (point dipole and quadrupole) PME and Ewald oracles (CPU)."""

import math

import numpy as np
import torch

from oracle import ewald as oe
from oracle import pme as op
from torchff.workloads import water_box


def _traceless(rng, n, scale):
    a = rng.normal(0.0, scale, (n, 3, 3))
    a = 0.5 * (a + a.transpose(0, 2, 1))
    tr = np.trace(a, axis1=1, axis2=2)
    a -= tr[:, None, None] * np.eye(3)[None] / 3.0
    return a


def _system(n_waters=12, L=1.3, seed=4, quad=False):
    w = water_box(n_waters, L, seed=seed)
    rng = np.random.default_rng(seed)
    p = rng.normal(0.0, 0.02, (w["natoms"], 3))
    t = lambda a: torch.tensor(a, dtype=torch.float64)
    th = t(_traceless(rng, w["natoms"], 0.01)) if quad else None
    return t(w["coords"]), t(w["box"]), t(w["charges"]), t(p), th


def test_pme_multipoles_converge_to_direct_ewald():
    for quad in (False, True):
        x, box, q, p, th = _system(quad=quad)
        alpha = 3.0
        ref = float(oe.ewald_recip_energy(x, box, q, 10, alpha, p=p, th=th))
        errs = []
        for K, order in ((16, 6), (24, 8), (32, 8)):
            e = float(op.pme_energy_multipoles(x, box, q, p, th, K, alpha, order))
            errs.append(abs(e - ref) / abs(ref))
        assert errs[-1] < 1e-7 and errs[-1] < errs[0], (quad, errs)


def test_pme_multipoles_triclinic_against_direct_ewald():
    x, box, q, p, th = _system(10, 1.4, 9, quad=True)
    box = torch.tensor([[1.4, 0.0, 0.0], [0.3, 1.3, 0.0], [-0.2, 0.25, 1.35]], dtype=torch.float64)
    ref = float(oe.ewald_recip_energy(x, box, q, 12, 3.0, p=p, th=th))
    e = float(op.pme_energy_multipoles(x, box, q, p, th, 36, 3.0, 8))
    assert abs(e - ref) < 1e-7 * abs(ref), (e, ref)


def test_zero_multipoles_reduce_to_charges():
    x, box, q, _, _ = _system()
    z = torch.zeros_like(x)
    z9 = torch.zeros(x.shape[0], 3, 3, dtype=torch.float64)
    e0 = float(op.pme_energy(x, box, q, 20, 3.0, 6))
    assert abs(float(op.pme_energy_multipoles(x, box, q, z, z9, 20, 3.0, 6)) - e0) < 1e-12 * abs(e0)
    assert abs(float(oe.ewald_recip_energy(x, box, q, 8, 3.0, p=z, th=z9)) - float(oe.ewald_recip_energy(x, box, q, 8, 3.0))) \
        < 1e-12


def _cluster_limit(charges_offsets, e_point, x, box, q, r0, alpha, kmax, ds):
    """E(recip + self) of small charge clusters (offsets scaled by d) minus their intra-cluster
    Gaussian pair terms sum q_i q_j erf(alpha r_ij)/r_ij, for each d in ds: the error against the
    point multipole must vanish like d^2."""
    errs = []
    for d in ds:
        qs = torch.tensor([c for c, _ in charges_offsets(d)], dtype=torch.float64)
        offs = torch.stack([o for _, o in charges_offsets(d)])
        xp = torch.cat([x, r0[None] + offs])
        qp = torch.cat([q, qs])
        e = float(oe.ewald_recip_energy(xp, box, qp, kmax, alpha))
        for i in range(len(qs)):
            for j in range(i):
                r = float(torch.linalg.vector_norm(offs[i] - offs[j]))
                e -= float(qs[i] * qs[j]) * math.erf(alpha * r) / r
        errs.append(abs(e - e_point))
    return errs


def test_dipole_is_the_limit_of_a_charge_pair():
    # a point dipole p at r equals charges +p/d at r + d/2 and -p/d at r - d/2 as d -> 0, once the
    # pair's own smooth mutual term (which the reciprocal sum contains and the self term of a point
    # dipole removes) is taken out
    x, box, q, _, _ = _system(8, 1.2, 7)
    alpha, kmax = 3.2, 10
    p0 = torch.tensor([0.011, -0.004, 0.007], dtype=torch.float64)
    r0 = torch.tensor([0.31, 0.52, 0.43], dtype=torch.float64)
    xd = torch.cat([x, r0[None]])
    qd = torch.cat([q, torch.zeros(1, dtype=torch.float64)])
    pd = torch.zeros_like(xd)
    pd[-1] = p0
    e_dip = float(oe.ewald_recip_energy(xd, box, qd, kmax, alpha, p=pd))
    n = p0 / torch.linalg.vector_norm(p0)
    pn = float(torch.linalg.vector_norm(p0))

    def pair(d):
        return [(pn / d, 0.5 * d * n), (-pn / d, -0.5 * d * n)]

    errs = _cluster_limit(pair, e_dip, x, box, q, r0, alpha, kmax, (2e-3, 1e-3))
    assert errs[1] < 0.3 * errs[0] and errs[1] < 1e-5 * abs(e_dip), errs
    e_pme = float(op.pme_energy_multipoles(xd, box, qd, pd, None, 40, alpha, 8))
    assert abs(e_pme - e_dip) < 1e-7 * abs(e_dip)


def test_quadrupole_is_the_limit_of_a_charge_cluster():
    # +Qc at +-d e1 and -Qc at +-d e2 (e1, e2 orthonormal): no charge, no dipole, second moment
    # sum q d d^T = 2 Qc d^2 (e1 e1^T - e2 e2^T), i.e. theta = (3/2) sum q d d^T
    # = 3 Qc d^2 (e1 e1^T - e2 e2^T), traceless; the hexadecapole vanishes like d^2
    x, box, q, _, _ = _system(8, 1.2, 11)
    alpha, kmax = 3.0, 10
    e1 = torch.tensor([0.6, 0.8, 0.0], dtype=torch.float64)
    e2 = torch.tensor([-0.48, 0.36, 0.8], dtype=torch.float64)
    c0 = 0.004  # theta = c0 (e1 e1^T - e2 e2^T)
    th0 = c0 * (torch.outer(e1, e1) - torch.outer(e2, e2))
    r0 = torch.tensor([0.41, 0.35, 0.66], dtype=torch.float64)
    xd = torch.cat([x, r0[None]])
    qd = torch.cat([q, torch.zeros(1, dtype=torch.float64)])
    thd = torch.zeros(xd.shape[0], 3, 3, dtype=torch.float64)
    thd[-1] = th0
    e_quad = float(oe.ewald_recip_energy(xd, box, qd, kmax, alpha, th=thd))
    # the same without the quadrupole differs by far more than the tolerance below
    e_none = float(oe.ewald_recip_energy(x, box, q, kmax, alpha))
    assert abs(e_quad - e_none) > 1e-4 * abs(e_quad)

    def cluster(d):
        Qc = c0 / (3.0 * d * d)
        return [(Qc, d * e1), (Qc, -d * e1), (-Qc, d * e2), (-Qc, -d * e2)]

    errs = _cluster_limit(cluster, e_quad, x, box, q, r0, alpha, kmax, (4e-3, 2e-3))
    assert errs[1] < 0.3 * errs[0] and errs[1] < 1e-6 * abs(e_quad), errs
    e_pme = float(op.pme_energy_multipoles(xd, box, qd, None, thd, 40, alpha, 8))
    assert abs(e_pme - e_quad) < 1e-7 * abs(e_quad)


def test_multipole_energy_derivatives_by_finite_differences():
    x, box, q, p, th = _system(6, 1.1, 2, quad=True)
    alpha, K, order = 3.0, 20, 6
    args = {"x": x, "q": q, "p": p, "th": th}

    def energy(a):
        return op.pme_energy_multipoles(a["x"], box, a["q"], a["p"], a["th"], K, alpha, order)

    leaves = {k: v.clone().requires_grad_(True) for k, v in args.items()}
    grads = dict(zip(leaves, torch.autograd.grad(energy(leaves), list(leaves.values()))))
    h = 1e-5
    for name in ("x", "p", "th", "q"):
        flat_idx = (0, 7, 13) if name != "q" else (0, 5, 11)
        for fi in flat_idx:
            idx = np.unravel_index(fi, args[name].shape)
            ap = {k: v.clone() for k, v in args.items()}
            am = {k: v.clone() for k, v in args.items()}
            ap[name][idx] += h
            am[name][idx] -= h
            fd = float(energy(ap) - energy(am)) / (2 * h)
            g = float(grads[name][idx])
            assert abs(fd - g) < 1e-6 * max(1.0, abs(fd)), (name, idx, fd, g)
