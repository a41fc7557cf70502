"""Host side of the torch-sim style consumer: states to serving requests (CPU, no GPU calls)."""

from types import SimpleNamespace

import numpy as np
import pytest
import torch

from torchff.md import NM_PER_ANGSTROM, state_requests
from torchff.workloads import water_box


def _state(boxes_nm, seeds, interleave=False):
    pos, cells, idx, q, s, e, mol = [], [], [], [], [], [], []
    for k, (L, nw) in enumerate(boxes_nm):
        w = water_box(nw, L, seed=seeds[k])
        pos.append(w["coords"] / NM_PER_ANGSTROM)
        cells.append(np.eye(3) * L / NM_PER_ANGSTROM)  # columns = lattice vectors (diagonal here)
        idx.append(np.full(w["natoms"], k))
        q.append(w["charges"]); s.append(w["sigma"]); e.append(w["epsilon"])
        mol.append(np.arange(w["natoms"]) // 3)
    cat = lambda a: np.concatenate(a)
    P, I, Q, S, E, M = cat(pos), cat(idx), cat(q), cat(s), cat(e), cat(mol)
    if interleave:
        order = np.argsort(np.arange(len(I)) % 7, kind="stable")
        P, I, Q, S, E, M = P[order], I[order], Q[order], S[order], E[order], M[order]
    st = SimpleNamespace(positions=torch.tensor(P), cell=torch.tensor(np.stack(cells)), system_idx=torch.tensor(I))
    return st, Q, S, E, M


def test_requests_per_system_in_nm():
    st, q, s, e, m = _state([(1.3, 60), (0.9, 20)], [1, 2])
    reqs = state_requests(st, q, s, e, m, spacing=0.1)
    assert [r.rid for r in reqs] == [0, 1] and [r.natoms for r in reqs] == [180, 60]
    assert reqs[0].box_lengths == pytest.approx((1.3,) * 3) and reqs[1].box_lengths == pytest.approx((0.9,) * 3)
    assert reqs[0].grid == (15, 15, 15) and reqs[1].grid == (9, 9, 9)  # smallest 2^a 3^b 5^c >= L / 0.1
    w = water_box(60, 1.3, seed=1)
    np.testing.assert_allclose(reqs[0].records[:, :3], w["coords"], rtol=1e-15, atol=1e-15)
    np.testing.assert_array_equal(reqs[0].records[:, 3], w["charges"])
    np.testing.assert_array_equal(reqs[1].records[:, 6], np.arange(60) // 3)


def test_interleaved_atoms_keep_state_order_per_system():
    st, q, s, e, m = _state([(1.2, 40), (1.0, 30)], [3, 4], interleave=True)
    reqs = state_requests(st, q, s, e, m)
    idx = st.system_idx.numpy()
    for r in reqs:
        sel = np.nonzero(idx == r.rid)[0]
        np.testing.assert_allclose(r.records[:, :3], st.positions.numpy()[sel] * NM_PER_ANGSTROM)
        np.testing.assert_array_equal(r.records[:, 3], q[sel])


def test_rejects_triclinic_and_mismatched_parameters():
    st, q, s, e, m = _state([(1.2, 40)], [5])
    tri = st.cell.clone()
    tri[0, 0, 1] = 2.0
    with pytest.raises(ValueError, match="orthorhombic"):
        state_requests(SimpleNamespace(positions=st.positions, cell=tri, system_idx=st.system_idx), q, s, e, m)
    with pytest.raises(ValueError, match="charges"):
        state_requests(st, q[:-1], s, e, m)


def test_single_system_without_index():
    st, q, s, e, m = _state([(1.1, 30)], [6])
    reqs = state_requests(SimpleNamespace(positions=st.positions, cell=st.cell), q, s, e, m)
    assert len(reqs) == 1 and reqs[0].natoms == 90
