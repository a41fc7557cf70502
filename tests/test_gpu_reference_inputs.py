"""Parity on the reference's own test inputs and cases (GPU), against the float64 oracle.

* PME ranks 0-2 on the reference's random multipole sets (tests/water/random_water_{300,1000,10000}
  .npz, kept in tests/golden/): order 6, K = 10, alpha = 0.26525, the case of
  tests/test_pme.py:57-88.  Energy and the gradients of coordinates, charges, dipoles and
  quadrupoles against autograd of the oracle: 1e-10 relative energy, 1e-9 relative RMS.
* The direct Ewald sum on the 300-atom set, kmax by the reference's error estimate
  (tests/test_ewald.py:55-69).
* MultipolarInteraction rank 0 as in tests/test_multipoles.py:73-126: N = 200 random points in a
  cube of edge (10 N)^(1/3), charges 0.1 N(0, 1) (the test's numpy RandomState(42) draws),
  cutoff 9 (above half the edge), pairs within it, exclusions (i, i+1), alpha -1 (bare) and 0.4,
  energy alone and with fields (energy, potential, field), gradients of coordinates and charges
  through E + sum(potential^2) (the test's loss) against the oracle's double backward.
* Neighbour lists in a triclinic box, both algorithms, bit-exact sets.
* Nonbonded energy on the cluster-pair list against the oracle on the oracle's pair list.
* The drop-in pair kernel accumulates gradients as exact integer sums: two runs bitwise equal.
"""

import math
import os

import numpy as np
import pytest
import torch

from oracle import ewald as oe
from oracle import nblist as onb
from oracle import pme as opme
from oracle import realspace
from torchff.workloads import COULOMB_KJMOL, make_workload

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _rms_rel(a, b):
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def _load(n):
    d = np.load(os.path.join(GOLDEN, f"random_water_{n}.npz"), allow_pickle=False)
    return {k: d[k] for k in d.files}


def _leaves(d, rank, dev, dtype=torch.float64):
    t = lambda a: torch.tensor(a, dtype=dtype, device=dev, requires_grad=True)
    return t(d["coords"]), t(d["q"]), (t(d["p"]) if rank >= 1 else None), (t(d["t"]) if rank >= 2 else None)


@pytest.mark.parametrize("n", [300, 1000, 10000])
@pytest.mark.parametrize("rank", [0, 1, 2])
def test_pme_reference_inputs(n, rank):
    import torchff
    d = _load(n)
    K, alpha = int(d["max_hkl"]), float(d["alpha"])
    x, q, p, t = _leaves(d, rank, "cpu")
    box = torch.tensor(d["box"], dtype=torch.float64)
    E_ref = opme.pme_energy_multipoles(x, box, q, p, t, K, alpha, 6) if rank else opme.pme_energy(x, box, q, K, alpha, 6)
    leaves = [v for v in (x, q, p, t) if v is not None]
    g_ref = torch.autograd.grad(E_ref, leaves)
    xd, qd, pd, td = _leaves(d, rank, "cuda")
    E = torchff.PME(alpha, K, rank).to(device="cuda", dtype=torch.float64)(xd, box.cuda(), qd, pd, td)
    dl = [v for v in (xd, qd, pd, td) if v is not None]
    g = torch.autograd.grad(E, dl)
    errs = [_rms_rel(a.cpu().numpy(), b.numpy()) for a, b in zip(g, g_ref)]
    print(f"PME N={n} rank {rank}: E={float(E):.10f} ref={float(E_ref):.10f} gradient rms rel {errs}")
    assert abs(float(E) - float(E_ref)) <= 1e-10 * abs(float(E_ref))
    for e in errs:
        assert e < 1e-9


def _reference_kmax(box_len, alpha):
    # the reference test's rule (tests/test_ewald.py:55-69)
    for i in range(2, 50):
        err = (i * math.sqrt(box_len * alpha) / 20.0) * math.exp(-math.pi ** 2 * i * i / (box_len * alpha) ** 2)
        if err < 1e-6:
            return i
    return 50


@pytest.mark.parametrize("rank", [0, 1, 2])
def test_ewald_reference_inputs(rank):
    import torchff
    d = _load(300)
    alpha = math.sqrt(-math.log10(2 * 1e-6)) / 9.0
    kmax = _reference_kmax(float(d["box"][0, 0]), alpha)
    x, q, p, t = _leaves(d, rank, "cpu")
    E_ref = oe.ewald_recip_energy(x, torch.tensor(d["box"], dtype=torch.float64), q, kmax, alpha, p=p, th=t)
    leaves = [v for v in (x, q, p, t) if v is not None]
    g_ref = torch.autograd.grad(E_ref, leaves)
    xd, qd, pd, td = _leaves(d, rank, "cuda")
    E = torchff.Ewald(alpha, kmax, rank, return_fields=False)(xd, torch.tensor(d["box"], device="cuda"), qd, pd, td)
    g = torch.autograd.grad(E, [v for v in (xd, qd, pd, td) if v is not None])
    print(f"Ewald kmax={kmax} rank {rank}: E={float(E):.10f} ref={float(E_ref):.10f}")
    assert abs(float(E) - float(E_ref)) <= 1e-10 * abs(float(E_ref))
    for a, b in zip(g, g_ref):
        assert _rms_rel(a.cpu().numpy(), b.numpy()) < 1e-9


def _multipoles_case(n=200, cutoff=9.0):
    # the input draws of tests/test_multipoles.py:13-51 (RandomState(42): coordinates, then charges)
    box_len = float((n * 10.0) ** (1.0 / 3.0))
    rng = np.random.RandomState(42)
    coords = rng.rand(n, 3) * box_len
    q = rng.randn(n) * 0.1
    box = np.eye(3) * box_len
    dr = coords[:, None, :] - coords[None, :, :]
    ds = dr @ np.linalg.inv(box)
    ds = ds - np.floor(ds + 0.5)
    dist = np.linalg.norm(ds @ box, axis=2)
    ii, jj = np.triu_indices(n, k=1)
    mask = dist[ii, jj] < cutoff
    pairs = np.column_stack([ii[mask], jj[mask]]).astype(np.int64)
    excl = np.array([[i % n, (i + 1) % n] for i in range(n)], dtype=np.int64)
    return coords, box, q, pairs, excl, cutoff


@pytest.mark.parametrize("ewald,use_excl", [(-1.0, False), (0.4, True), (0.4, False)])
@pytest.mark.parametrize("return_fields", [False, True])
def test_multipolar_rank0_reference_case(ewald, use_excl, return_fields):
    import torchff
    coords, box, q, pairs, excl, cutoff = _multipoles_case()
    assert cutoff > box[0, 0] / 2
    pe = excl if use_excl else None
    x = torch.tensor(coords, dtype=torch.float64, requires_grad=True)
    qq = torch.tensor(q, dtype=torch.float64, requires_grad=True)
    E = realspace.ewald_real_energy(x, torch.tensor(box), torch.tensor(pairs), None if pe is None else torch.tensor(pe),
                                    qq, cutoff, ewald, 1.0)
    if return_fields:
        pot = torch.autograd.grad(E, qq, create_graph=True)[0]
        loss_ref = E + torch.sum(pot ** 2)
    else:
        loss_ref = E
    g_ref = torch.autograd.grad(loss_ref, [x, qq])
    xd = torch.tensor(coords, dtype=torch.float64, device="cuda", requires_grad=True)
    qd = torch.tensor(q, dtype=torch.float64, device="cuda", requires_grad=True)
    m = torchff.MultipolarInteraction(0, cutoff, ewald, 1.0, return_fields=return_fields)
    out = m(xd, torch.tensor(box, device="cuda"), torch.tensor(pairs, device="cuda"), qd,
            pairs_excl=None if pe is None else torch.tensor(pe, device="cuda"))
    if return_fields:
        e, potd, field = out
        assert e.shape == (1,) and potd.shape == (200,) and field.shape == (200, 3)
        assert float(field.abs().max()) == 0.0
        assert _rms_rel(potd.detach().cpu().numpy(), pot.detach().numpy()) < 1e-10
        loss = e.sum() + torch.sum(potd ** 2)
    else:
        e = out
        loss = e
    assert abs(float(e.sum()) - float(E)) <= 1e-10 * abs(float(E))
    g = torch.autograd.grad(loss, [xd, qd])
    for a, b in zip(g, g_ref):
        assert _rms_rel(a.cpu().numpy(), b.numpy()) < 1e-9


TRI = np.array([[2.6, 0.0, 0.0], [0.7, 2.4, 0.0], [-0.5, 0.6, 2.5]])


@pytest.mark.parametrize("algo", ["nsquared", "cell_list"])
@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_neighbor_list_triclinic(algo, dtype):
    import torchff
    rng = np.random.default_rng(71)
    n = 1500
    frac = rng.uniform(0, 1, (n, 3))
    coords = frac @ TRI
    rc = 0.9
    c = torch.tensor(coords, dtype=dtype, device="cuda")
    nl = torchff.NeighborList(n, algorithm=algo).to("cuda")
    pairs, npairs = nl(c, torch.tensor(TRI, dtype=dtype, device="cuda"), rc, -1, padding=True)
    got = onb.canonicalize(pairs.cpu().numpy(), int(npairs.item()))
    ref = onb.brute_force(c.double().cpu().numpy(), TRI, rc)
    assert len(ref) > 10000
    assert np.array_equal(got, ref)


def test_cluster_pair_energy_against_the_oracle():
    import torchff
    w = make_workload("water900")
    rc = w["cutoff"]
    dev = "cuda"
    x = torch.tensor(w["coords"], dtype=torch.float64, device=dev, requires_grad=True)
    box = torch.tensor(w["box"], dtype=torch.float64, device=dev)
    q, s, e = (torch.tensor(w[k], dtype=torch.float64, device=dev, requires_grad=True)
               for k in ("charges", "sigma", "epsilon"))
    excl = torch.tensor(np.asarray(w["exclusions"]).T.copy(), device=dev)  # (2, E) as the reference
    nbl = torchff.build_cluster_pairs(x.detach(), box, rc, excl, 0.6)
    E = torchff.compute_nonbonded_energy_from_cluster_pairs(x, box, s, e, q, COULOMB_KJMOL, rc, *nbl, True)
    g = torch.autograd.grad(E.sum(), [x, q, s, e])
    pairs = onb.neighbor_pairs(w["coords"], w["box"], rc, exclusions=w["exclusions"])
    t = lambda a: torch.tensor(a, dtype=torch.float64, requires_grad=True)
    xr, qr, sr, er = t(w["coords"]), t(w["charges"]), t(w["sigma"]), t(w["epsilon"])
    E_ref = realspace.nonbonded_energy(xr, torch.tensor(pairs), torch.tensor(w["box"]), sr, er, qr, COULOMB_KJMOL, rc,
                                       True)
    g_ref = torch.autograd.grad(E_ref, [xr, qr, sr])  # eps: the oracle's sqrt(eps_i eps_j) at eps = 0 is unbounded
    assert abs(float(E.sum()) - float(E_ref)) <= 1e-10 * abs(float(E_ref))
    for a, b in zip(g[:3], g_ref):
        assert _rms_rel(a.cpu().numpy(), b.numpy()) < 1e-10


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_pair_kernel_gradients_repeat_bitwise(dtype):
    import torchff
    w = make_workload("water96k")
    dev = "cuda"
    x = torch.tensor(w["coords"], dtype=dtype, device=dev)
    box = torch.tensor(w["box"], dtype=dtype, device=dev)
    nl = torchff.NeighborList(w["natoms"], exclusions=torch.tensor(w["exclusions"], device=dev), algorithm="cell_list")
    pairs, _ = nl(x, box, 1.1, 30_000_000, padding=True)
    q, s, e = (torch.tensor(w[k], dtype=dtype, device=dev) for k in ("charges", "sigma", "epsilon"))
    outs = []
    for _ in range(2):
        xx = x.clone().requires_grad_(True)
        qq, ss = q.clone().requires_grad_(True), s.clone().requires_grad_(True)
        E = torchff.Nonbonded()(xx, pairs, box, ss, e, qq, COULOMB_KJMOL, w["cutoff"], ewald_alpha=w["alpha"],
                                pairs_excl=torch.tensor(w["pairs_excl"], device=dev)).sum()
        outs.append([E.detach()] + list(torch.autograd.grad(E, [xx, qq, ss])))
    for a, b in zip(*outs):
        assert torch.equal(a, b)
