"""Direct Ewald op with point dipoles and quadrupoles against the float64 oracle (GPU).

Every output is checked against an autograd derivative of the oracle's energy (potential =
dE/dq, field = -dE/dp, dE/dtheta, forces = -dE/dr), and the backward through the potential and
field outputs against its second derivatives.  Tolerances: energy 1e-11 relative in float64,
1e-5 in float32; vectors by RMS relative error 1e-9 / 1e-4.
"""

import numpy as np
import pytest
import torch

from oracle import ewald as oe
from torchff.workloads import water_box

pytestmark = pytest.mark.gpu

TOL = {torch.float64: (1e-11, 1e-9), torch.float32: (1e-5, 1e-4)}
BOX_TRI = np.array([[1.5, 0.0, 0.0], [0.3, 1.45, 0.0], [-0.2, 0.25, 1.5]])


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
    return w["coords"], (w["box"] if box is None else box), w["charges"], p, th


def _oracle(coords, box, q, p, th, kmax, alpha, wq=None, wf=None):
    t = lambda a: None if a is None else torch.tensor(a, dtype=torch.float64, requires_grad=True)
    x, qq, pp, tt = t(coords), t(q), t(p), t(th)
    E = oe.ewald_recip_energy(x, torch.tensor(box, dtype=torch.float64), qq, kmax, alpha, p=pp, th=tt)
    leaves = [v for v in (x, qq, pp, tt) if v is not None]
    g = torch.autograd.grad(E, leaves, create_graph=True)
    out = {"E": float(E.detach()), "forces": -g[0].detach().numpy(), "pot": g[1].detach().numpy()}
    if pp is not None:
        out["field"] = -g[2].detach().numpy()
    if tt is not None:
        out["dtheta"] = g[-1].detach().numpy()
    if wq is not None:
        loss = 0.6 * E + (torch.tensor(wq) * g[1]).sum()
        if pp is not None:
            loss = loss - (torch.tensor(wf) * g[2]).sum()
        out["loss_grads"] = [v.numpy() for v in torch.autograd.grad(loss, leaves)]
    return out


def _dev(coords, box, q, p, th, dtype):
    t = lambda a: None if a is None else torch.tensor(a, dtype=dtype, device="cuda", requires_grad=True)
    return t(coords), torch.tensor(box, dtype=dtype, device="cuda"), t(q), t(p), t(th)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("rank,tri", [(1, False), (2, False), (1, True), (2, True)])
def test_ewald_multipoles(dtype, rank, tri):
    import torchff
    coords, box, q, p, th = _system(12, 1.2, 40 + rank, rank, BOX_TRI if tri else None)
    kmax, alpha = 7, 3.0
    ref = _oracle(coords, box, q, p, th, kmax, alpha)
    x, bx, qq, pp, tt = _dev(coords, box, q, p, th, dtype)
    E = torchff.Ewald(alpha, kmax, rank, return_fields=False)(x, bx, qq, pp, tt)
    E.backward()
    te, tf = TOL[dtype]
    assert abs(float(E.detach()) - ref["E"]) < te * abs(ref["E"]), (float(E.detach()), ref["E"])
    assert _rms_rel(-x.grad.double().cpu().numpy(), ref["forces"]) < tf
    assert _rms_rel(qq.grad.double().cpu().numpy(), ref["pot"]) < tf
    assert _rms_rel(-pp.grad.double().cpu().numpy(), ref["field"]) < tf
    if th is not None:
        assert _rms_rel(tt.grad.double().cpu().numpy(), ref["dtheta"]) < tf


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
@pytest.mark.parametrize("rank", [0, 1, 2])
def test_ewald_fields_and_backward_through_them(dtype, rank):
    import torchff
    coords, box, q, p, th = _system(10, 1.1, 50 + rank, rank)
    n = len(q)
    rng = np.random.default_rng(5)
    wq = rng.normal(size=n)
    wf = rng.normal(size=(n, 3))
    kmax, alpha = 6, 3.2
    ref = _oracle(coords, box, q, p, th, kmax, alpha, wq=wq, wf=wf)
    x, bx, qq, pp, tt = _dev(coords, box, q, p, th, dtype)
    E, pot, field = torchff.Ewald(alpha, kmax, rank, return_fields=True)(x, bx, qq, pp, tt)
    te, tf = TOL[dtype]
    assert abs(float(E.detach()) - ref["E"]) < te * abs(ref["E"])
    assert _rms_rel(pot.detach().double().cpu().numpy(), ref["pot"]) < tf
    if rank == 0:
        assert float(field.detach().abs().max()) == 0.0
    else:
        assert _rms_rel(field.detach().double().cpu().numpy(), ref["field"]) < tf
    loss = 0.6 * E + (torch.tensor(wq, dtype=dtype, device="cuda") * pot).sum()
    if rank >= 1:
        loss = loss + (torch.tensor(wf, dtype=dtype, device="cuda") * field).sum()
    loss.backward()
    got = [x.grad, qq.grad] + ([pp.grad] if pp is not None else []) + ([tt.grad] if tt is not None else [])
    for g, r in zip(got, ref["loss_grads"]):
        assert _rms_rel(g.double().cpu().numpy(), r) < tf


def test_ewald_multipoles_match_pme():
    # the same multipolar reciprocal energy and forces from the direct sum and from a fine PME grid
    import torchff
    coords, box, q, p, th = _system(12, 1.2, 61, 2)
    x, bx, qq, pp, tt = _dev(coords, box, q, p, th, torch.float64)
    e_ew = torchff.Ewald(3.0, 10, 2, return_fields=False)(x, bx, qq, pp, tt)
    e_ew.backward()
    f_ew = -x.grad.clone()
    x.grad = None
    pme = torchff.PME(3.0, 40, 2, order=8).to("cuda", torch.float64)
    e_pme = pme(x, bx, qq, pp, tt)
    e_pme.backward()
    assert abs(float(e_ew.detach()) - float(e_pme.detach())) < 1e-7 * abs(float(e_ew.detach()))
    assert _rms_rel((-x.grad).cpu().numpy(), f_ew.cpu().numpy()) < 1e-5
