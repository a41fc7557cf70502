"""PME parity against the float64 oracle (GPU).

Energies to 1e-10 relative in float64 (same discretisation, different
summation order and FFT), 1e-5 in float32; forces and charge derivatives by
RMS relative error (1e-9 float64, 1e-4 float32).
"""

import os

import numpy as np
import pytest
import torch

from oracle import pme as opme
from torchff.workloads import make_workload, water_box

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "random_water_300.npz")


def _ref(coords, box, q, K, alpha, order):
    x = torch.tensor(coords, dtype=torch.float64, requires_grad=True)
    qq = torch.tensor(q, dtype=torch.float64, requires_grad=True)
    E = opme.pme_energy(x, torch.tensor(box, dtype=torch.float64), qq, K, alpha, order)
    gx, gq = torch.autograd.grad(E, [x, qq])
    return float(E), gx.numpy(), gq.numpy()


def _dev(coords, box, q, K, alpha, order, dtype):
    import torchff
    m = torchff.PME(alpha, K, 0, order=order).to(device="cuda", dtype=dtype)
    x = torch.tensor(coords, dtype=dtype, device="cuda", requires_grad=True)
    qq = torch.tensor(q, dtype=dtype, device="cuda", requires_grad=True)
    E = m(x, torch.tensor(box, dtype=dtype, device="cuda"), qq)
    E.backward()
    return float(E), x.grad.double().cpu().numpy(), qq.grad.double().cpu().numpy()


def _rms_rel(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


TOL = {torch.float64: (1e-10, 1e-9), torch.float32: (1e-5, 1e-4)}


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("order,K", [(6, 24), (5, 24), (5, 25), (4, 20), (6, 27), (8, 32)])
def test_water900(dtype, order, K):
    w = make_workload("water900")
    ref = _ref(w["coords"], w["box"], w["charges"], K, w["alpha"], order)
    got = _dev(w["coords"], w["box"], w["charges"], K, w["alpha"], order, dtype)
    te, tf = TOL[dtype]
    assert abs(got[0] - ref[0]) < te * abs(ref[0])
    assert _rms_rel(got[1], ref[1]) < tf
    assert _rms_rel(got[2], ref[2]) < tf


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_reference_fixture_random_water_300(dtype):
    # the reference's own PME test inputs (tests/test_pme.py:57-88: K=10, order 6)
    d = np.load(GOLDEN, allow_pickle=False)
    K, alpha = int(d["max_hkl"]), float(d["alpha"])
    ref = _ref(d["coords"], d["box"], d["q"], K, alpha, 6)
    got = _dev(d["coords"], d["box"], d["q"], K, alpha, 6, dtype)
    te, tf = TOL[dtype]
    assert abs(got[0] - ref[0]) < te * abs(ref[0])
    assert _rms_rel(got[1], ref[1]) < tf
    assert _rms_rel(got[2], ref[2]) < tf


def test_triclinic_box():
    w = water_box(60, 1.6, seed=11)
    box = np.array([[1.6, 0.0, 0.0], [0.3, 1.5, 0.0], [-0.2, 0.25, 1.7]])
    ref = _ref(w["coords"], box, w["charges"], 20, 3.0, 5)
    got = _dev(w["coords"], box, w["charges"], 20, 3.0, 5, torch.float64)
    assert abs(got[0] - ref[0]) < 1e-10 * abs(ref[0])
    assert _rms_rel(got[1], ref[1]) < 1e-9


def test_water96k_energy_forces():
    w = make_workload("water96k")
    ref = _ref(w["coords"], w["box"], w["charges"], w["K"], w["alpha"], w["order"])
    for dtype in (torch.float64, torch.float32):
        got = _dev(w["coords"], w["box"], w["charges"], w["K"], w["alpha"], w["order"], dtype)
        te, tf = TOL[dtype]
        assert abs(got[0] - ref[0]) < te * abs(ref[0])
        assert _rms_rel(got[1], ref[1]) < tf


def test_bad_grid_raises():
    import torchff
    m = torchff.PME(3.0, 14 * 1, 0, order=5).cuda()  # 14 = 2 * 7 is not 2^a 3^b 5^c
    x = torch.zeros(3, 3, device="cuda", dtype=torch.float64)
    with pytest.raises(RuntimeError, match="2\\^a 3\\^b 5\\^c"):
        m(x, torch.eye(3, dtype=torch.float64, device="cuda"), torch.zeros(3, dtype=torch.float64, device="cuda"))


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_pme_return_fields(dtype):
    # pme_long_range_all: the same energy, the per-atom potential dE/dq (with the self term) and a
    # zero field for point charges; gradients through the energy
    import torchff
    w = make_workload("water900")
    E_ref, gx_ref, gq_ref = _ref(w["coords"], w["box"], w["charges"], 24, w["alpha"], 6)
    m = torchff.PME(w["alpha"], 24, 0, return_fields=True).to(device="cuda", dtype=dtype)
    x = torch.tensor(w["coords"], dtype=dtype, device="cuda", requires_grad=True)
    q = torch.tensor(w["charges"], dtype=dtype, device="cuda", requires_grad=True)
    E, pot, field = m(x, torch.tensor(w["box"], dtype=dtype, device="cuda"), q)
    te, tf = TOL[dtype]
    assert abs(float(E) - E_ref) < te * abs(E_ref)
    assert _rms_rel(pot.detach().double().cpu().numpy(), gq_ref) < tf
    assert field.shape == (w["natoms"], 3) and float(field.abs().max()) == 0.0
    E.backward()
    assert _rms_rel(x.grad.double().cpu().numpy(), gx_ref) < tf
    assert _rms_rel(q.grad.double().cpu().numpy(), gq_ref) < tf
