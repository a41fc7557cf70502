"""Multipolar PME (point dipoles and quadrupoles) against the float64 oracle (GPU).

The oracle's energy is differentiable, so every output of the op is checked against an
autograd derivative of it: potential = dE/dq, field = -dE/dp, quadrupole gradient dE/dtheta,
forces = -dE/dr, and the backward through the potential and field outputs (the loss
gE E + sum wq pot + sum wf field) against second derivatives of the oracle.  Tolerances:
energy 1e-10 relative in float64, 1e-5 in float32; vectors by RMS relative error 1e-9 / 1e-4.
"""

import numpy as np
import pytest
import torch

from oracle import pme as opme
from torchff.workloads import water_box

pytestmark = pytest.mark.gpu

TOL = {torch.float64: (1e-10, 1e-9), torch.float32: (1e-5, 1e-4)}
BOX_TRI = np.array([[1.6, 0.0, 0.0], [0.35, 1.5, 0.0], [-0.25, 0.3, 1.55]])


def _rms_rel(a, b):
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def _system(n_waters, L, seed, rank, box=None):
    w = water_box(n_waters, L, seed=seed)
    rng = np.random.default_rng(seed)
    n = w["natoms"]
    p = rng.normal(0.0, 0.02, (n, 3)) if rank >= 1 else None
    th = None
    if rank >= 2:
        a = rng.normal(0.0, 0.01, (n, 3, 3))
        a = 0.5 * (a + a.transpose(0, 2, 1))
        a -= np.trace(a, axis1=1, axis2=2)[:, None, None] * np.eye(3)[None] / 3.0
        th = a
    b = w["box"] if box is None else box
    return w["coords"], b, w["charges"], p, th


def _oracle(coords, box, q, p, th, K, alpha, order, wq=None, wf=None):
    t = lambda a: None if a is None else torch.tensor(a, dtype=torch.float64, requires_grad=True)
    x, qq, pp, tt = t(coords), t(q), t(p), t(th)
    bx = torch.tensor(box, dtype=torch.float64)
    E = opme.pme_energy_multipoles(x, bx, qq, pp, tt, K, alpha, order)
    leaves = [v for v in (x, qq, pp, tt) if v is not None]
    g = torch.autograd.grad(E, leaves, create_graph=True)
    gx, gq = g[0], g[1]
    gp = g[2] if pp is not None else None
    gt = g[-1] if tt is not None else None
    out = {"E": float(E.detach()), "pot": gq.detach().numpy(), "forces": -gx.detach().numpy()}
    if gp is not None:
        out["field"] = -gp.detach().numpy()
    if gt is not None:
        out["dtheta"] = gt.detach().numpy()
    if wq is not None:
        loss = 0.7 * E + (torch.tensor(wq) * gq).sum()
        if gp is not None and wf is not None:
            loss = loss + (torch.tensor(wf) * (-gp)).sum()
        gl = torch.autograd.grad(loss, leaves)
        out["loss_grads"] = [v.numpy() for v in gl]
    return out


def _dev_tensors(coords, box, q, p, th, dtype):
    t = lambda a: None if a is None else torch.tensor(a, dtype=dtype, device="cuda", requires_grad=True)
    return t(coords), torch.tensor(box, dtype=dtype, device="cuda"), t(q), t(p), t(th)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("rank,order,K,tri", [(1, 6, 20, False), (1, 5, 18, True), (1, 4, 16, False),
                                              (2, 6, 20, False), (2, 5, 18, True), (2, 4, 16, False),
                                              (1, 3, 16, False)])
def test_multipole_energy_and_gradients(dtype, rank, order, K, tri):
    import torchff
    coords, box, q, p, th = _system(16, 1.2, 5 + rank + order, rank, BOX_TRI if tri else None)
    alpha = 3.2
    ref = _oracle(coords, box, q, p, th, K, alpha, order)
    m = torchff.PME(alpha, K, rank, order=order).to(device="cuda", dtype=dtype)
    x, bx, qq, pp, tt = _dev_tensors(coords, box, q, p, th, dtype)
    E = m(x, bx, qq, pp, tt)
    E.backward()
    te, tf = TOL[dtype]
    print(f"rank {rank} order {order} K {K} tri {tri} {dtype}: E {float(E.detach()):.12f} ref {ref['E']:.12f}", flush=True)
    assert abs(float(E) - ref["E"]) < te * abs(ref["E"])
    assert _rms_rel(-x.grad.double().cpu().numpy(), ref["forces"]) < tf
    assert _rms_rel(qq.grad.double().cpu().numpy(), ref["pot"]) < tf
    assert _rms_rel(-pp.grad.double().cpu().numpy(), ref["field"]) < tf
    if th is not None:
        assert _rms_rel(tt.grad.double().cpu().numpy(), ref["dtheta"]) < tf


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("rank", [0, 1, 2])
def test_return_fields_and_backward_through_them(dtype, rank):
    # pme_long_range_all: (energy, potential, field) with the self terms; the backward takes
    # upstream gradients on all three (the reference's PMEAllFunction), checked against second
    # derivatives of the oracle
    import torchff
    coords, box, q, p, th = _system(14, 1.15, 31 + rank, rank)
    n = len(q)
    rng = np.random.default_rng(77)
    wq = rng.normal(size=n)
    wf = rng.normal(size=(n, 3)) if rank >= 1 else None
    K, order, alpha = 20, 6, 3.0
    ref = _oracle(coords, box, q, p, th, K, alpha, order, wq=wq, wf=wf)
    m = torchff.PME(alpha, K, rank, order=order, return_fields=True).to(device="cuda", dtype=dtype)
    x, bx, qq, pp, tt = _dev_tensors(coords, box, q, p, th, dtype)
    E, pot, field = m(x, bx, qq, pp, tt)
    te, tf = TOL[dtype]
    assert abs(float(E) - ref["E"]) < te * abs(ref["E"])
    assert _rms_rel(pot.detach().double().cpu().numpy(), ref["pot"]) < tf
    if rank == 0:
        assert float(field.abs().max()) == 0.0
    else:
        assert _rms_rel(field.detach().double().cpu().numpy(), ref["field"]) < tf
    loss = 0.7 * E + (torch.tensor(wq, dtype=dtype, device="cuda") * pot).sum()
    if rank >= 1:
        loss = loss + (torch.tensor(wf, dtype=dtype, device="cuda") * field).sum()
    loss.backward()
    got = [x.grad, qq.grad] + ([pp.grad] if pp is not None else []) + ([tt.grad] if tt is not None else [])
    for g, r in zip(got, ref["loss_grads"]):
        assert _rms_rel(g.double().cpu().numpy(), r) < tf


def test_quadrupole_layouts_agree():
    # (N, 3, 3), (N, 9) and the packed (N, 6) [xx, xy, xz, yy, yz, zz] give the same energy and
    # consistent gradients
    import torchff
    coords, box, q, p, th = _system(10, 1.1, 3, 2)
    m = torchff.PME(3.0, 16, 2, order=6).to(device="cuda", dtype=torch.float64)
    dev = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    t33 = dev(th).requires_grad_(True)
    t9 = dev(th.reshape(-1, 9)).requires_grad_(True)
    t6 = dev(th[:, [0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]).requires_grad_(True)
    es = []
    for t in (t33, t9, t6):
        e = m(dev(coords), dev(box), dev(q), dev(p), t)
        e.backward()
        es.append(float(e))
    assert abs(es[1] - es[0]) < 1e-13 * abs(es[0]) and abs(es[2] - es[0]) < 1e-12 * abs(es[0])
    g33 = t33.grad.cpu().numpy()
    assert np.allclose(t9.grad.cpu().numpy().reshape(-1, 3, 3), g33, rtol=1e-12, atol=1e-14)
    g6 = t6.grad.cpu().numpy()
    want6 = np.stack([g33[:, 0, 0], g33[:, 0, 1] + g33[:, 1, 0], g33[:, 0, 2] + g33[:, 2, 0], g33[:, 1, 1],
                      g33[:, 1, 2] + g33[:, 2, 1], g33[:, 2, 2]], axis=1)
    assert np.allclose(g6, want6, rtol=1e-12, atol=1e-14)


def test_quadrupoles_need_order_4():
    import torchff
    coords, box, q, p, th = _system(4, 1.0, 1, 2)
    m = torchff.PME(3.0, 16, 2, order=3).to(device="cuda", dtype=torch.float64)
    dev = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    with pytest.raises(RuntimeError, match="order >= 4"):
        m(dev(coords), dev(box), dev(q), dev(p), dev(th))


def test_rank0_path_unchanged_by_zero_multipoles():
    # explicit zero dipoles and quadrupoles reproduce the point-charge op
    import torchff
    coords, box, q, _, _ = _system(20, 1.3, 8, 0)
    dev = lambda a: torch.tensor(a, dtype=torch.float64, device="cuda")
    n = len(q)
    e0 = float(torchff.PME(3.0, 20, 0, order=6).to("cuda", torch.float64)(dev(coords), dev(box), dev(q)))
    e2 = float(torchff.PME(3.0, 20, 2, order=6).to("cuda", torch.float64)(
        dev(coords), dev(box), dev(q), torch.zeros(n, 3, dtype=torch.float64, device="cuda"),
        torch.zeros(n, 3, 3, dtype=torch.float64, device="cuda")))
    assert abs(e2 - e0) < 1e-12 * abs(e0)
