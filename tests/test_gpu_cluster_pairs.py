"""Cluster-pair list ops (build_cluster_pairs, decode_cluster_pairs, nonbonded energy / forces on
cluster pairs) on the GPU.

Decoded pair sets equal the oracle's canonical pair set bit for bit (random and water boxes,
orthorhombic and triclinic, float32 and float64, small and oversized cells, with and without
exclusions); the format's own invariants are checked directly; energies and gradients on the
cluster pairs equal the atom-pair op on the same pair set (float64 to 1e-12, float32 to 1e-5).
"""

import numpy as np
import pytest
import torch

from oracle import nblist as onb
from torchff.workloads import make_workload

pytestmark = pytest.mark.gpu

TRI = np.array([[2.6, 0.0, 0.0], [0.5, 2.5, 0.0], [-0.4, 0.6, 2.4]])


def _uniform(n, density, seed, box=None):
    rng = np.random.default_rng(seed)
    if box is None:
        L = (n / density) ** (1 / 3)
        return rng.uniform(0, L, (n, 3)), np.eye(3) * L
    return rng.uniform(0, 1, (n, 3)) @ box, box


def _build(coords, box, rc, dtype=torch.float64, excl=None, cell=0.4, maxrows=-1):
    import torchff
    c = torch.tensor(coords, dtype=dtype, device="cuda")
    b = torch.tensor(box, dtype=dtype, device="cuda")
    e = None if excl is None else torch.tensor(np.asarray(excl).T.copy(), device="cuda")  # (2, E) as the reference
    nbl = torchff.build_cluster_pairs(c, b, rc, e, cell, maxrows)
    return c, b, nbl


def _decode(c, b, nbl, rc, cap=-1, padding=False):
    import torchff
    pairs, npairs = torchff.decode_cluster_pairs(c, b, *nbl, rc, cap, padding)
    torch.cuda.synchronize()
    return pairs.cpu().numpy(), int(npairs.item())


def _oracle(coords, box, rc, dtype, excl=None):
    x = torch.tensor(coords, dtype=dtype).double().numpy()
    return onb.neighbor_pairs(x, box, rc, exclusions=excl)


def _check_format(nbl, n):
    si, coo, mask, ic, ia = (t.cpu() for t in nbl)
    C = (n + 31) // 32
    assert sorted(si.tolist()) == list(range(n))  # a permutation
    x, y = coo[0].numpy(), coo[1].numpy()
    assert (x <= y).all() and (y < C).all()
    diag = set(x[x == y].tolist())
    assert diag == set(range(C))  # every diagonal block
    assert len(set(zip(x.tolist(), y.tolist()))) == len(x)  # no duplicate blocks
    assert mask.dtype == torch.uint32 and tuple(mask.shape) == (len(x), 32)
    icn, ian = ic.numpy(), ia.numpy()
    assert (np.diff(icn[icn >= 0]) >= 0).all()  # rows grouped by cluster, ascending
    # atoms in a cluster's rows belong to clusters above it and outside its exclusion blocks
    slot = np.empty(n, np.int64)
    slot[si.numpy()] = np.arange(n)
    blocks = set(zip(x.tolist(), y.tolist()))
    for r in range(0, len(icn), max(1, len(icn) // 200)):
        if icn[r] < 0:
            continue
        for a in ian[r][ian[r] >= 0]:
            cy = slot[a] // 32
            assert cy > icn[r] and (icn[r], cy) not in blocks


@pytest.mark.parametrize("n,rc,dtype,cell,tri", [
    (200, 0.6, torch.float64, 0.4, False), (1000, 0.6, torch.float32, 0.4, False),
    (3000, 0.7, torch.float64, 0.6, False), (2000, 0.9, torch.float64, 0.45, True),
    (700, 0.9, torch.float32, 5.0, False), (33, 0.3, torch.float64, 0.4, False),
])
def test_random_boxes(n, rc, dtype, cell, tri):
    x, box = _uniform(n, 100.0, n + 3, TRI if tri else None)
    c, b, nbl = _build(x, box, rc, dtype, cell=cell)
    _check_format(nbl, n)
    got, npairs = _decode(c, b, nbl, rc)
    ref = _oracle(x, box, rc, dtype)
    print(f"n {n} rc {rc} {dtype} cell {cell} tri {tri}: {npairs} pairs, rows {nbl[3].shape[0]}", flush=True)
    assert npairs == len(ref)
    np.testing.assert_array_equal(onb.canonicalize(got, npairs), ref)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_water_with_exclusions(dtype):
    w = make_workload("water900")
    c, b, nbl = _build(w["coords"], w["box"], w["cutoff"], dtype, excl=w["exclusions"], cell=0.6)
    _check_format(nbl, w["natoms"])
    got, npairs = _decode(c, b, nbl, w["cutoff"])
    ref = _oracle(w["coords"], w["box"], w["cutoff"], dtype, excl=w["exclusions"])
    assert npairs == len(ref)
    np.testing.assert_array_equal(onb.canonicalize(got, npairs), ref)
    # one-directional exclusions give the same list (they are symmetrised)
    e1 = w["exclusions"][w["exclusions"][:, 0] < w["exclusions"][:, 1]]
    _, _, nbl1 = _build(w["coords"], w["box"], w["cutoff"], dtype, excl=e1, cell=0.6)
    got1, np1 = _decode(c, b, nbl1, w["cutoff"])
    np.testing.assert_array_equal(onb.canonicalize(got1, np1), ref)


def test_build_is_deterministic_and_capacities():
    w = make_workload("water900")
    c, b, nbl = _build(w["coords"], w["box"], w["cutoff"], excl=w["exclusions"])
    _, _, nbl2 = _build(w["coords"], w["box"], w["cutoff"], excl=w["exclusions"])
    as_i32 = lambda t: t.view(torch.int32) if t.dtype == torch.uint32 else t
    for t1, t2 in zip(nbl, nbl2):
        assert torch.equal(as_i32(t1), as_i32(t2))
    M = nbl[3].shape[0]
    # a larger row capacity pads with -1; a smaller one raises
    _, _, nblp = _build(w["coords"], w["box"], w["cutoff"], excl=w["exclusions"], maxrows=M + 10)
    assert nblp[3].shape[0] == M + 10 and (nblp[3][M:] == -1).all() and (nblp[4][M:] == -1).all()
    assert torch.equal(nblp[4][:M], nbl[4])
    with pytest.raises(RuntimeError, match="max_num_interacting_clusters"):
        _build(w["coords"], w["box"], w["cutoff"], excl=w["exclusions"], maxrows=max(1, M // 2))
    # decode with a capacity: padding=True returns the full buffer, -1 after the pairs
    got, npairs = _decode(c, b, nblp, w["cutoff"])
    padded, np2 = _decode(c, b, nblp, w["cutoff"], cap=npairs + 7, padding=True)
    assert np2 == npairs and (padded[npairs:] == -1).all()
    np.testing.assert_array_equal(padded[:npairs], got)
    with pytest.raises(RuntimeError, match="Too many"):
        _decode(c, b, nblp, w["cutoff"], cap=npairs // 2)


@pytest.mark.parametrize("dtype", [torch.float64, torch.float32])
def test_nonbonded_on_cluster_pairs_matches_atom_pairs(dtype):
    import torchff
    w = make_workload("water900")
    rc = w["cutoff"]
    c, b, nbl = _build(w["coords"], w["box"], rc, dtype, excl=w["exclusions"], cell=0.6)
    pairs, npairs = torchff.decode_cluster_pairs(c, b, *nbl, rc)
    dev = lambda a: torch.tensor(a, dtype=dtype, device="cuda")
    q, sig, eps = dev(w["charges"]), dev(w["sigma"]), dev(w["epsilon"])
    k = 138.935456
    outs = []
    for path in ("cluster", "atom"):
        x = c.clone().requires_grad_(True)
        qq, ss, ee = (t.clone().requires_grad_(True) for t in (q, sig, eps))
        if path == "cluster":
            E = torchff.compute_nonbonded_energy_from_cluster_pairs(x, b, ss, ee, qq, k, rc, *nbl, True)
        else:
            E = torchff.compute_nonbonded_energy_from_atom_pairs(x, pairs.long(), b, ss, ee, qq, k, rc, True)
        E.sum().backward()
        outs.append([float(E.detach().sum())] + [t.grad.double().cpu().numpy() for t in (x, qq, ss, ee)])
    tol = 1e-12 if dtype == torch.float64 else 1e-5
    assert abs(outs[0][0] - outs[1][0]) < tol * abs(outs[1][0])
    for a, r in zip(outs[0][1:], outs[1][1:]):
        assert np.sqrt(np.mean((a - r) ** 2)) <= tol * np.sqrt(np.mean(r ** 2)) + 1e-300
    # forces op accumulates -dE/dr
    f = torch.full_like(c, 0.5)
    torchff.compute_nonbonded_forces_from_cluster_pairs(c, b, sig, eps, q, k, rc, *nbl, f)
    fr = torch.full_like(c, 0.5)
    torchff.compute_nonbonded_forces_from_atom_pairs(c, pairs.long(), b, sig, eps, q, k, rc, fr)
    d = (f - fr).double().cpu().numpy()
    assert np.sqrt(np.mean(d ** 2)) <= tol * np.sqrt(np.mean((fr - 0.5).double().cpu().numpy() ** 2))


def test_water96k_pair_count_matches_cell_list():
    import torchff
    w = make_workload("water96k")
    rc = w["cutoff"]
    c, b, nbl = _build(w["coords"], w["box"], rc, excl=w["exclusions"], cell=0.6)
    got, npairs = _decode(c, b, nbl, rc)
    nl = torchff.NeighborList(w["natoms"], exclusions=torch.tensor(w["exclusions"], device="cuda"),
                              algorithm="cell_list").to("cuda")
    ref, nref = nl(c, b, rc, 40_000_000, padding=True)
    nref = int(nref.item())
    print(f"water96k: {npairs} pairs from {nbl[3].shape[0]} rows and {nbl[1].shape[1]} exclusion blocks", flush=True)
    assert npairs == nref
    np.testing.assert_array_equal(onb.canonicalize(got, npairs), onb.canonicalize(ref.cpu().numpy(), nref))


def test_rejects():
    import torchff
    x, box = _uniform(100, 100.0, 1)
    c = torch.tensor(x, device="cuda")
    b = torch.tensor(box, device="cuda")
    with pytest.raises(RuntimeError, match="half the smallest box height"):
        torchff.build_cluster_pairs(c, b, float(box[0, 0]) * 0.6, None, 0.4)
    with pytest.raises(RuntimeError):
        torchff.build_cluster_pairs(c, b, 0.5, None, 0.0)
