"""Host-side logic of the multipole and cluster-pair ops (CPU, no GPU calls)."""

import numpy as np
import pytest
import torch

from torchff import _ops
from torchff.nblist import NeighborList


def test_quadrupole_layouts_unpack_to_the_same_tensor():
    rng = np.random.default_rng(0)
    a = rng.normal(size=(5, 3, 3))
    a = torch.tensor(0.5 * (a + a.transpose(0, 2, 1)))
    packed = a[:, [0, 0, 0, 1, 1, 2], [0, 1, 2, 1, 2, 2]]
    for t in (a, a.reshape(5, 9), packed):
        assert torch.equal(_ops._as_quadrupole9(t, 5), a.reshape(5, 9))
    with pytest.raises(ValueError):
        _ops._as_quadrupole9(torch.zeros(5, 4), 5)


def test_packed_quadrupole_gradient_collects_both_off_diagonals():
    # the (N, 6) form's xy entry stands for both xy and yx of the tensor
    packed = torch.arange(12.0).reshape(2, 6).requires_grad_(True)
    g9 = torch.arange(18.0).reshape(2, 9)
    (_ops._as_quadrupole9(packed, 2) * g9).sum().backward()
    g = g9.reshape(2, 3, 3)
    want = torch.stack([g[:, 0, 0], g[:, 0, 1] + g[:, 1, 0], g[:, 0, 2] + g[:, 2, 0], g[:, 1, 1],
                        g[:, 1, 2] + g[:, 2, 1], g[:, 2, 2]], 1)
    assert torch.equal(packed.grad, want)


@pytest.mark.parametrize("layout", ["2xE", "Ex2"])
def test_cluster_pair_exclusions_match_the_neighbor_list_csr(layout):
    rng = np.random.default_rng(1)
    n = 40
    pairs = rng.integers(0, n, size=(60, 2))
    coo = torch.tensor(pairs)
    rp, col = _ops._exclusions_csr(coo.t() if layout == "2xE" else coo, n, torch.device("cpu"))
    rp_ref, col_ref = NeighborList.convert_pairs_coo_to_csr(coo, n)
    assert torch.equal(rp, rp_ref) and torch.equal(col, col_ref)
    # symmetric: j in row i iff i in row j
    rows = {i: set(col[rp[i]:rp[i + 1]].tolist()) for i in range(n)}
    assert all(i in rows[j] for i in range(n) for j in rows[i])


def test_cluster_pair_exclusions_none_and_padding():
    assert _ops._exclusions_csr(None, 10, torch.device("cpu")) == (None, None)
    e = torch.tensor([[0, 1], [-1, 3], [2, 2]])
    rp, col = _ops._exclusions_csr(e, 4, torch.device("cpu"))
    assert rp.tolist() == [0, 1, 2, 3, 3] and col.tolist() == [1, 0, 2]
