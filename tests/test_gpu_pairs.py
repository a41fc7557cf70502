"""Pair-term parity against the float64 oracle (GPU).

Tolerances: float64 energies/gradients to 1e-10 relative (the device and the
oracle evaluate the same formulas in different summation orders); float32 to
2e-5 relative on energies and 2e-4 on the force RMS.
"""

import numpy as np
import pytest
import torch

from oracle import nblist as onb
from oracle import realspace
from torchff.workloads import COULOMB_KJMOL, make_workload

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def water():
    w = make_workload("water900")
    pairs = onb.neighbor_pairs(w["coords"], w["box"], w["cutoff"] + 0.05, exclusions=w["exclusions"])
    # mix in padding rows and flip orientation of half the rows, as device lists have them
    p = pairs.copy()
    p[::2] = p[::2, ::-1]
    p = np.concatenate([p, -np.ones((37, 2), np.int64)])
    return w, p


def _t(x, dtype=torch.float64, grad=False, dev="cpu"):
    return torch.tensor(x, dtype=dtype, device=dev, requires_grad=grad)


def _rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def _rms_rel(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 2e-5)])
@pytest.mark.parametrize("do_shift", [True, False])
def test_nonbonded_energy_and_grads(water, dtype, tol, do_shift):
    import torchff
    w, p = water
    rc = w["cutoff"]
    # perturbed parameters so sigma/epsilon/charge gradients are all exercised (eps > 0 everywhere)
    rng = np.random.default_rng(0)
    sig = rng.uniform(0.25, 0.35, w["natoms"])
    eps = rng.uniform(0.1, 0.8, w["natoms"])
    q = w["charges"] + rng.normal(0, 0.02, w["natoms"])
    ref_in = [_t(w["coords"], grad=True), _t(sig, grad=True), _t(eps, grad=True), _t(q, grad=True)]
    E = realspace.nonbonded_energy(ref_in[0], torch.tensor(p), _t(w["box"]), ref_in[1], ref_in[2], ref_in[3],
                                   COULOMB_KJMOL, rc, do_shift)
    gref = torch.autograd.grad(E, ref_in)
    dev_in = [_t(w["coords"], dtype, True, "cuda"), _t(sig, dtype, True, "cuda"), _t(eps, dtype, True, "cuda"),
              _t(q, dtype, True, "cuda")]
    Ed = torchff.compute_nonbonded_energy_from_atom_pairs(dev_in[0], torch.tensor(p, device="cuda"),
                                                          _t(w["box"], dtype, dev="cuda"), dev_in[1], dev_in[2],
                                                          dev_in[3], COULOMB_KJMOL, rc, do_shift)
    assert Ed.shape == (1,)
    Ed.sum().backward()
    assert _rel(float(Ed.item()), float(E)) < tol
    for a, b in zip(dev_in, gref):
        assert _rms_rel(a.grad.double().cpu().numpy(), b.numpy()) < (tol if dtype == torch.float64 else 2e-4)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_nonbonded_forces_inplace(water, dtype):
    import torchff
    w, p = water
    x = _t(w["coords"], grad=True)
    E = realspace.nonbonded_energy(x, torch.tensor(p), _t(w["box"]), _t(w["sigma"]), _t(w["epsilon"]),
                                   _t(w["charges"]), COULOMB_KJMOL, w["cutoff"], True)
    F = -torch.autograd.grad(E, x)[0].numpy()
    f = torch.ones((w["natoms"], 3), dtype=dtype, device="cuda")
    torchff.compute_nonbonded_forces_from_atom_pairs(_t(w["coords"], dtype, dev="cuda"), torch.tensor(p, device="cuda"),
                                                     _t(w["box"], dtype, dev="cuda"), _t(w["sigma"], dtype, dev="cuda"),
                                                     _t(w["epsilon"], dtype, dev="cuda"),
                                                     _t(w["charges"], dtype, dev="cuda"), COULOMB_KJMOL, w["cutoff"], f)
    got = f.double().cpu().numpy() - 1.0
    assert _rms_rel(got, F) < (1e-10 if dtype == torch.float64 else 2e-4)


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 2e-5)])
@pytest.mark.parametrize("alpha", [2.92, -1.0, 0.0])
def test_multipolar_rank0(water, dtype, tol, alpha):
    import torchff
    w, p = water
    x, q = _t(w["coords"], grad=True), _t(w["charges"], grad=True)
    excl = torch.tensor(w["pairs_excl"])
    E = realspace.ewald_real_energy(x, _t(w["box"]), torch.tensor(p), excl, q, w["cutoff"], alpha, COULOMB_KJMOL)
    gx, gq = torch.autograd.grad(E, [x, q])
    xd, qd = _t(w["coords"], dtype, True, "cuda"), _t(w["charges"], dtype, True, "cuda")
    m = torchff.MultipolarInteraction(0, w["cutoff"], alpha, COULOMB_KJMOL)
    Ed = m(xd, _t(w["box"], dtype, dev="cuda"), torch.tensor(p, device="cuda"), qd,
           pairs_excl=excl.to("cuda"))
    assert Ed.shape == ()
    Ed.backward()
    assert _rel(float(Ed), float(E)) < tol
    assert _rms_rel(xd.grad.double().cpu().numpy(), gx.numpy()) < (tol if dtype == torch.float64 else 2e-4)
    assert _rms_rel(qd.grad.double().cpu().numpy(), gq.numpy()) < (tol if dtype == torch.float64 else 2e-4)


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 2e-5)])
def test_lj_type_table(water, dtype, tol):
    import torchff
    w, p = water
    x = _t(w["coords"], grad=True)
    st, et = _t(w["sigma_table"], grad=True), _t(w["epsilon_table"], grad=True)
    types = torch.tensor(w["atom_types"])
    E = realspace.lj_energy(x, torch.tensor(p), _t(w["box"]), st, et, w["cutoff"], atom_types=types)
    g = torch.autograd.grad(E, [x, st, et])
    d = [_t(w["coords"], dtype, True, "cuda"), _t(w["sigma_table"], dtype, True, "cuda"),
         _t(w["epsilon_table"], dtype, True, "cuda")]
    v = torchff.Vdw("LennardJones", use_type_pairs=True)
    Ed = v(d[0], torch.tensor(p, device="cuda"), _t(w["box"], dtype, dev="cuda"), d[1], d[2], w["cutoff"],
           atom_types=types.to("cuda"))
    Ed.backward()
    assert _rel(float(Ed), float(E)) < tol
    for a, b in zip(d, g):
        assert _rms_rel(a.grad.double().cpu().numpy(), b.numpy()) < (tol if dtype == torch.float64 else 2e-4)


def test_lj_per_pair(water):
    import torchff
    w, p = water
    valid = p[(p[:, 0] >= 0)]
    rng = np.random.default_rng(1)
    s = rng.uniform(0.25, 0.35, len(valid))
    e = rng.uniform(0.1, 0.8, len(valid))
    x, st, et = _t(w["coords"], grad=True), _t(s, grad=True), _t(e, grad=True)
    E = realspace.lj_energy(x, torch.tensor(valid), _t(w["box"]), st, et, w["cutoff"])
    g = torch.autograd.grad(E, [x, st, et])
    d = [_t(w["coords"], grad=True, dev="cuda"), _t(s, grad=True, dev="cuda"), _t(e, grad=True, dev="cuda")]
    Ed = torchff.compute_lennard_jones_energy(d[0], torch.tensor(valid, device="cuda"), _t(w["box"], dev="cuda"),
                                              d[1], d[2], w["cutoff"])
    Ed.backward()
    assert _rel(float(Ed), float(E)) < 1e-10
    for a, b in zip(d, g):
        assert _rms_rel(a.grad.cpu().numpy(), b.numpy()) < 1e-10


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-10), (torch.float32, 2e-5)])
def test_fused_ewald_lj(water, dtype, tol):
    import torchff
    w, p = water
    x, q = _t(w["coords"], grad=True), _t(w["charges"], grad=True)
    excl = torch.tensor(w["pairs_excl"])
    E = realspace.ewald_lj_energy(x, torch.tensor(p), _t(w["box"]), _t(w["sigma"]), _t(w["epsilon"]), q,
                                  COULOMB_KJMOL, w["cutoff"], w["alpha"], excl)
    gx, gq = torch.autograd.grad(E, [x, q])
    xd, qd = _t(w["coords"], dtype, True, "cuda"), _t(w["charges"], dtype, True, "cuda")
    nb = torchff.Nonbonded()
    Ed = nb(xd, torch.tensor(p, device="cuda"), _t(w["box"], dtype, dev="cuda"), _t(w["sigma"], dtype, dev="cuda"),
            _t(w["epsilon"], dtype, dev="cuda"), qd, COULOMB_KJMOL, w["cutoff"], ewald_alpha=w["alpha"],
            pairs_excl=excl.to("cuda"))
    Ed.sum().backward()
    assert _rel(float(Ed.item()), float(E)) < tol
    assert _rms_rel(xd.grad.double().cpu().numpy(), gx.numpy()) < (tol if dtype == torch.float64 else 2e-4)
    assert _rms_rel(qd.grad.double().cpu().numpy(), gq.numpy()) < (tol if dtype == torch.float64 else 2e-4)


def test_epsilon_gradient_zero_epsilon_convention(water):
    """Pairs whose mixed epsilon is zero contribute no epsilon gradient (documented convention)."""
    import torchff
    w, p = water
    eps = _t(w["epsilon"], grad=True, dev="cuda")  # hydrogens have eps = 0
    E = torchff.compute_nonbonded_energy_from_atom_pairs(_t(w["coords"], dev="cuda"), torch.tensor(p, device="cuda"),
                                                         _t(w["box"], dev="cuda"), _t(w["sigma"], dev="cuda"), eps,
                                                         _t(w["charges"], dev="cuda"), COULOMB_KJMOL, w["cutoff"])
    E.sum().backward()
    g = eps.grad.cpu().numpy()
    assert np.all(np.isfinite(g))
    assert np.all(g[1::3] == 0) and np.all(g[2::3] == 0)
    assert np.any(g[0::3] != 0)
