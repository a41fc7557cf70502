"""torch-sim style consumer of the serving path (GPU): several periodic water boxes in one state,
one batch per call.  Energies (eV) and forces (eV/Angstrom) against the all-images float64 oracle
(energy within 1e-9 (float64) / 1e-6 (mixed) of the summed term magnitudes, force RMS 1e-9 / 1e-4
relative), each system alone bitwise equal to the batch."""

from types import SimpleNamespace

import numpy as np
import pytest
import torch

from oracle import images
from torchff.md import KJMOL_PER_EV, NM_PER_ANGSTROM, BatchedSimModel
from torchff.serving import pme_grid_size
from torchff.workloads import box_length_for_density, water_box

pytestmark = pytest.mark.gpu


def _systems(sizes, seeds):
    out = []
    for nw, sd in zip(sizes, seeds):
        L = box_length_for_density(nw)
        out.append(water_box(nw, L, seed=sd))
    return out


def _state(ws, order=None):
    P = np.concatenate([w["coords"] for w in ws]) / NM_PER_ANGSTROM
    I = np.concatenate([np.full(w["natoms"], k) for k, w in enumerate(ws)])
    cells = np.stack([np.asarray(w["box"]).T / NM_PER_ANGSTROM for w in ws])
    par = {k: np.concatenate([np.asarray(w[k]) for w in ws]) for k in ("charges", "sigma", "epsilon")}
    mol = np.concatenate([np.arange(w["natoms"]) // 3 for w in ws])
    if order is not None:
        P, I, mol = P[order], I[order], mol[order]
        par = {k: v[order] for k, v in par.items()}
    st = SimpleNamespace(positions=torch.tensor(P, device="cuda"), cell=torch.tensor(cells, device="cuda"),
                         system_idx=torch.tensor(I, device="cuda"))
    return st, par, mol


def _engine(precision):
    from torchff.serving import ServingEngine
    return ServingEngine(slots=1, max_systems=8, max_atoms=4096, max_grid_points=1 << 16, max_k=32,
                         precision=precision)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_state_against_the_oracle(precision):
    ws = _systems([20, 60, 150], [1, 2, 3])
    n = sum(w["natoms"] for w in ws)
    order = np.argsort(np.arange(n) % 5, kind="stable")  # systems interleaved in the state
    st, par, mol = _state(ws, order)
    eng = _engine(precision)
    model = BatchedSimModel(eng, par["charges"], par["sigma"], par["epsilon"], mol)
    out = model(st)
    assert out["energy"].shape == (3,) and out["forces"].shape == (n, 3)
    idx = st.system_idx.cpu().numpy()
    gidx = np.arange(n)[order]  # original (per-system concatenated) index of each state atom
    starts = np.cumsum([0] + [w["natoms"] for w in ws])
    te, tf = (1e-9, 1e-9) if precision == "fp64" else (1e-6, 1e-4)
    for k, w in enumerate(ws):
        K = [pme_grid_size(float(w["box"][a][a]), 0.1) for a in range(3)]
        E, F, terms = images.energy_forces(w, eng.cutoff, eng.alpha, K, eng.order)
        scale = sum(abs(v) for v in terms.values())  # mixed energies are bounded against the term sizes
        e = float(out["energy"][k]) * KJMOL_PER_EV
        sel = idx == k
        f = out["forces"].cpu().numpy()[sel] * KJMOL_PER_EV / NM_PER_ANGSTROM
        F = F[gidx[sel] - starts[k]]  # the oracle's forces in the state's order of this system's atoms
        print(f"system {k}: {w['natoms']} atoms E {e:.8f} ref {E:.8f}", flush=True)
        assert abs(e - E) < te * scale + 1e-6
        assert np.sqrt(np.mean((f - F) ** 2)) <= tf * np.sqrt(np.mean(F ** 2))


def test_each_system_alone_equals_the_batch():
    ws = _systems([30, 90], [7, 8])
    st, par, mol = _state(ws)
    eng = _engine("mixed")
    model = BatchedSimModel(eng, par["charges"], par["sigma"], par["epsilon"], mol)
    both = model(st)
    start = 0
    for k, w in enumerate(ws):
        sl = slice(start, start + w["natoms"])
        one = SimpleNamespace(positions=st.positions[sl], cell=st.cell[k:k + 1],
                              system_idx=torch.zeros(w["natoms"], dtype=torch.int64, device="cuda"))
        alone = BatchedSimModel(eng, par["charges"][sl], par["sigma"][sl], par["epsilon"][sl], mol[sl])(one)
        assert torch.equal(alone["energy"][0], both["energy"][k])
        assert torch.equal(alone["forces"], both["forces"][sl])
        start += w["natoms"]


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_resident_steps_match_the_request_path(precision):
    # positions move every step (device only); then the cell tensor is replaced (re-upload)
    ws = _systems([25, 80, 140], [11, 12, 13])
    n = sum(w["natoms"] for w in ws)
    order = np.argsort(np.arange(n) % 3, kind="stable")
    st, par, mol = _state(ws, order)
    model = BatchedSimModel(_engine(precision), par["charges"], par["sigma"], par["epsilon"], mol)
    ref_model = BatchedSimModel(_engine(precision), par["charges"], par["sigma"], par["epsilon"], mol)
    gen = torch.Generator(device="cuda").manual_seed(3)
    for step in range(4):
        if step == 3:
            st.cell = st.cell * 1.01  # a new cell tensor: the batch is rebuilt
        out = model(st)
        ref = ref_model.forward_via_requests(st)
        assert torch.equal(out["energy"], ref["energy"]), step
        assert torch.equal(out["forces"], ref["forces"]), step
        st.positions = st.positions + 0.01 * torch.randn(st.positions.shape, device="cuda", dtype=st.positions.dtype,
                                                         generator=gen)


def test_resident_model_keeps_positions_on_the_device():
    # after the first call (the upload), steps read no positions back to the host
    ws = _systems([40, 70], [21, 22])
    st, par, mol = _state(ws)
    model = BatchedSimModel(_engine("mixed"), par["charges"], par["sigma"], par["epsilon"], mol)
    model(st)
    key = model._res["key"]
    from unittest import mock
    with mock.patch("torchff.md.state_requests", side_effect=AssertionError("host path taken")):
        for _ in range(3):
            st.positions = st.positions + 0.001
            model(st)
    assert model._res["key"][3] == key[3]  # same cell: no rebuild
