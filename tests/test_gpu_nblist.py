"""Neighbor-list parity: device pair sets equal the oracle's sorted set bit for bit (GPU)."""

import numpy as np
import pytest
import torch

from oracle import nblist as onb
from torchff.workloads import make_workload

pytestmark = pytest.mark.gpu

ALGOS = ["nsquared", "cell_list"]


def _uniform(n, density, seed):
    L = (n / density) ** (1 / 3)
    rng = np.random.default_rng(seed)
    return rng.uniform(0, L, (n, 3)), np.eye(3) * L


def _device_pairs(algo, coords, box, rc, dtype=torch.float64, excl=None, include_self=False, cap=None):
    import torchff
    c = torch.tensor(coords, dtype=dtype, device="cuda")
    b = None if box is None else torch.tensor(box, dtype=dtype, device="cuda")
    nl = torchff.NeighborList(len(coords), exclusions=None if excl is None else torch.tensor(excl, device="cuda"),
                              include_self=include_self, algorithm=algo).to("cuda")
    pairs, npairs = nl(c, b, rc, -1 if cap is None else cap, padding=True)
    torch.cuda.synchronize()
    return pairs.cpu().numpy(), int(npairs.item())


def _oracle(coords, box, rc, dtype=torch.float64, excl=None, include_self=False):
    # float32 inputs are judged on their float32 values (promoted exactly to float64)
    x = np.asarray(torch.tensor(coords, dtype=dtype).double().numpy())
    return onb.neighbor_pairs(x, box, rc, exclusions=excl, include_self=include_self) if box is not None \
        else onb.brute_force(x, None, rc, None if excl is None else onb.coo_to_csr(excl, len(x)), include_self)


@pytest.mark.parametrize("algo", ALGOS)
@pytest.mark.parametrize("n,rc,dtype,include_self", [
    (50, 0.5, torch.float64, False), (200, 0.8, torch.float64, False), (500, 0.5, torch.float32, False),
    (50, 0.5, torch.float64, True), (200, 0.8, torch.float64, True), (3000, 0.6, torch.float32, False),
])
def test_random_periodic(algo, n, rc, dtype, include_self):
    x, box = _uniform(n, 100.0, n + 17)
    got, npairs = _device_pairs(algo, x, box, rc, dtype, include_self=include_self)
    ref = _oracle(x, box, rc, dtype, include_self=include_self)
    assert npairs == len(ref)
    np.testing.assert_array_equal(onb.canonicalize(got, npairs), ref)
    assert np.all(got[npairs:] == -1)
    valid = got[:npairs]
    assert np.all(valid[:, 0] >= valid[:, 1])  # rows (a, b) with a > b (a == b only for self pairs)


@pytest.mark.parametrize("n,rc", [(50, 0.5), (400, 0.8)])
def test_non_periodic(n, rc):
    x, _ = _uniform(n, 100.0, 5)
    got, npairs = _device_pairs("nsquared", x, None, rc)
    np.testing.assert_array_equal(onb.canonicalize(got, npairs), _oracle(x, None, rc))


@pytest.mark.parametrize("algo", ALGOS)
def test_cutoff_edge_cases(algo):
    x, box = _uniform(30, 100.0, 3)
    got, npairs = _device_pairs(algo, x, box, 0.0)
    assert npairs == 0
    got, npairs = _device_pairs(algo, x, box, box[0, 0] * 10)
    assert npairs == 30 * 29 // 2
    x1, box1 = _uniform(1, 100.0, 3)
    got, npairs = _device_pairs(algo, x1, box1, 0.5)
    assert npairs == 0


@pytest.mark.parametrize("algo", ALGOS)
def test_water_exclusions(algo):
    w = make_workload("water900")
    got, npairs = _device_pairs(algo, w["coords"], w["box"], w["cutoff"], excl=w["exclusions"])
    ref = onb.neighbor_pairs(w["coords"], w["box"], w["cutoff"], exclusions=w["exclusions"])
    np.testing.assert_array_equal(onb.canonicalize(got, npairs), ref)


@pytest.mark.parametrize("algo", ALGOS)
def test_capacity_overflow_reports_count(algo):
    x, box = _uniform(400, 100.0, 9)
    ref = _oracle(x, box, 0.8)
    cap = len(ref) // 3
    got, npairs = _device_pairs(algo, x, box, 0.8, cap=cap)
    assert npairs == len(ref)
    assert got.shape[0] == cap and np.all(got >= 0)
    refset = set(map(tuple, ref.tolist()))
    assert set(map(tuple, onb.canonicalize(got).tolist())) <= refset
    import torchff
    nl = torchff.NeighborList(400, algorithm=algo)
    with pytest.raises(RuntimeError, match="Too many neighbor pairs"):
        nl(torch.tensor(x, device="cuda"), torch.tensor(box, device="cuda"), 0.8, cap)


def test_cell_list_water96k_with_skin():
    w = make_workload("water96k")
    rc = w["cutoff"] + w["skin"]
    got, npairs = _device_pairs("cell_list", w["coords"], w["box"], rc, excl=w["exclusions"], cap=40_000_000)
    ref = onb.neighbor_pairs(w["coords"], w["box"], rc, exclusions=w["exclusions"])
    assert npairs == len(ref)
    np.testing.assert_array_equal(onb.canonicalize(got, npairs), ref)


def test_cell_list_requires_box():
    import torchff
    with pytest.raises(RuntimeError, match="periodic box"):
        torch.ops.torchff.build_neighbor_list_cell_list(torch.zeros(4, 3, device="cuda"), None, 0.5, -1, None,
                                                        None, False)
    nl = torchff.NeighborList(4, algorithm="cell_list")
    with pytest.raises(ValueError):
        nl(torch.zeros(4, 3, device="cuda"), None, 0.5)
