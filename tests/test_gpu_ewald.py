"""Direct Ewald op (point charges) against the float64 oracle's FFT-free sum (GPU).

Tolerances: float64 energy 1e-11 relative, potential and force RMS relative 1e-10 (both
from autograd of the oracle energy: dE/dq and -dE/dr); float32 1e-5 / 1e-4.
"""

import math

import numpy as np
import pytest
import torch

import torchff
from oracle import ewald as oe
from oracle import pme as op
from torchff.workloads import make_workload, water_box

pytestmark = pytest.mark.gpu


def _oracle(c, box, q, kmax, alpha):
    ct = torch.tensor(c, dtype=torch.float64, requires_grad=True)
    qt = torch.tensor(q, dtype=torch.float64, requires_grad=True)
    e = oe.ewald_recip_energy(ct, torch.tensor(box, dtype=torch.float64), qt, kmax, alpha)
    gc, gq = torch.autograd.grad(e, [ct, qt])
    return float(e), gq.numpy(), -gc.numpy()


def _rms_rel(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


CASES = [
    # water900 (cubic 2.08 nm), kmax 7
    ("water900", None, 7),
    # 30 waters in an orthorhombic box, kmax 10
    ("ortho", np.diag([1.3, 1.1, 0.95]), 10),
    # the same atoms in a triclinic box (k = B^-1 m)
    ("tri", np.array([[1.3, 0.0, 0.0], [0.2, 1.1, 0.0], [-0.15, 0.1, 0.95]]), 6),
]


def _inputs(name, box):
    if name == "water900":
        w = make_workload("water900")
        return w["coords"], w["box"], w["charges"], w["alpha"]
    w = water_box(30, 1.0, seed=9)
    return w["coords"], box, w["charges"], 3.2


@pytest.mark.parametrize("name,box,kmax", CASES, ids=[c[0] for c in CASES])
def test_ewald_matches_oracle(name, box, kmax):
    c, b, q, alpha = _inputs(name, box)
    E, pot_ref, F_ref = _oracle(c, b, q, kmax, alpha)
    dev = "cuda"
    ct = torch.tensor(c, device=dev)
    bt = torch.tensor(b, device=dev)
    qt = torch.tensor(q, device=dev)
    e, pot, field = torchff.ewald_long_range_with_fields(ct, bt, qt, None, None, kmax, alpha)
    print(f"{name}: dE={float(e) - E:.3e} pot_rms={_rms_rel(pot.cpu().numpy(), pot_ref):.2e}")
    assert abs(float(e) - E) < 1e-11 * abs(E)
    assert _rms_rel(pot.cpu().numpy(), pot_ref) < 1e-10
    assert float(field.abs().max()) == 0.0
    # forces through autograd of the energy op
    cg = ct.clone().requires_grad_(True)
    qg = qt.clone().requires_grad_(True)
    e2 = torchff.ewald_long_range(cg, bt, qg, None, None, kmax, alpha)
    gc, gq = torch.autograd.grad(e2 * 2.0, [cg, qg])
    assert _rms_rel(-gc.cpu().numpy() / 2.0, F_ref) < 1e-10
    assert _rms_rel(gq.cpu().numpy() / 2.0, pot_ref) < 1e-10


def test_ewald_float32():
    c, b, q, alpha = _inputs("water900", None)
    E, pot_ref, F_ref = _oracle(c, b, q, 7, alpha)
    m = torchff.Ewald(alpha, 7, 0, return_fields=False)
    ct = torch.tensor(c, dtype=torch.float32, device="cuda", requires_grad=True)
    e = m(ct, torch.tensor(b, dtype=torch.float32, device="cuda"), torch.tensor(q, dtype=torch.float32, device="cuda"))
    e.backward()
    assert abs(float(e) - E) < 1e-5 * abs(E)
    assert _rms_rel(-ct.grad.double().cpu().numpy(), F_ref) < 1e-4


def test_ewald_agrees_with_converged_pme():
    # the same reciprocal energy from the direct sum and from PME at a fine grid and high order
    c, b, q, alpha = _inputs("water900", None)
    dev = "cuda"
    ct, bt, qt = (torch.tensor(x, device=dev) for x in (c, b, q))
    e_ew = float(torchff.ewald_long_range(ct, bt, qt, None, None, 12, alpha))
    pme = torchff.PME(alpha, 64, 0, order=8).to(dev).to(torch.float64)
    e_pme = float(pme(ct, bt, qt))
    assert abs(e_ew - e_pme) < 1e-6 * abs(e_ew), (e_ew, e_pme)


def test_ewald_rejects():
    ct = torch.zeros((3, 3), dtype=torch.float64, device="cuda")
    bt = torch.eye(3, dtype=torch.float64, device="cuda")
    qt = torch.zeros(3, dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError):
        torchff.ewald_long_range(ct, bt, qt, None, None, 40, 3.0)
    with pytest.raises(ValueError):
        torchff.Ewald(3.0, 5, 3)
    with pytest.raises(ValueError):
        torchff.Ewald(3.0, 5, 0, use_customized_ops=False)
