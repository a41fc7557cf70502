"""Host-side checks that need no GPU: library exports, schemas, CSR conversion, moduli, workloads."""

import ctypes
import os
import re

import numpy as np
import pytest
import torch

import torchff
from torchff import _lib, bsplines
from oracle import bspline as obs
from oracle import nblist as onb
from torchff.workloads import WORKLOADS, make_workload

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    txt = open(os.path.join(ROOT, "include", "torchff_hip.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(tff_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_lib.LIB_PATH), "build the library first (python torchff-lib_amd/build_lib.py)"
    lib = ctypes.CDLL(_lib.LIB_PATH)
    declared = _header_functions()
    assert declared == _lib.exported_symbols()
    for name in declared:
        assert hasattr(lib, name), name


def test_workspace_queries_and_errors_without_gpu():
    lib = _lib.load(require_gpu=False)
    assert lib.tff_abi_version() == 2
    box = (ctypes.c_double * 9)(9.86, 0, 0, 0, 9.86, 0, 0, 0, 9.86)
    assert lib.tff_nblist_workspace_bytes(96000, box, 1.1, 1) > 96000 * 64
    assert lib.tff_pair_workspace_bytes(1, 96000, 10, 0) >= 96000 * 8 * 5
    assert lib.tff_pme_workspace_bytes(0, 96000, 96, 96, 96) >= 96 ** 3 * 16
    # argument validation happens before any device work
    rc = lib.tff_pme_energy(1, None, None, 10, box, 14, 14, 14, 5, 3.0, None, None, None, 1, 1.0, None, None, None,
                            0, None)
    assert rc != 0 and b"2^a 3^b 5^c" in lib.tff_last_error()
    rc = lib.tff_build_neighbor_list(1, None, 10, None, 1.0, 1, 0, None, None, None, 0, None, None, 0, None)
    assert rc != 0


def test_ops_registered_with_reference_schemas():
    names = ["build_neighbor_list_nsquared", "build_neighbor_list_nsquared_out", "build_neighbor_list_cell_list",
             "build_neighbor_list_cell_list_out", "compute_nonbonded_energy_from_atom_pairs",
             "compute_nonbonded_forces_from_atom_pairs", "compute_multipolar_energy_from_atom_pairs",
             "compute_lennard_jones_energy", "pme_long_range", "compute_nonbonded_ewald_energy_from_atom_pairs"]
    for n in names:
        assert hasattr(torch.ops.torchff, n)
    s = str(torch.ops.torchff.pme_long_range.default._schema)
    assert "Tensor xmoduli, Tensor ymoduli, Tensor zmoduli, int order=6" in s


def test_cpu_tensors_are_rejected():
    with pytest.raises(NotImplementedError):
        torch.ops.torchff.build_neighbor_list_nsquared(torch.zeros(4, 3), None, 1.0, -1, None, None, False)
    with pytest.raises(RuntimeError):
        torchff.compute_nonbonded_energy_from_atom_pairs(torch.zeros(2, 3), torch.tensor([[0, 1]]), torch.eye(3),
                                                         torch.ones(2), torch.ones(2), torch.ones(2), 1.0, 1.0)


def test_use_customized_ops_false_is_refused():
    with pytest.raises(ValueError):
        torchff.NeighborList(10, use_customized_ops=False)
    with pytest.raises(ValueError):
        torchff.PME(3.0, 10, 0, use_customized_ops=False)


def test_csr_conversion_matches_oracle():
    w = make_workload("water900")
    rp, col = torchff.NeighborList.convert_pairs_coo_to_csr(torch.tensor(w["exclusions"]), w["natoms"])
    orp, ocol = onb.coo_to_csr(w["exclusions"], w["natoms"])
    np.testing.assert_array_equal(rp.numpy(), orp)
    np.testing.assert_array_equal(col.numpy(), ocol)
    # one direction only, with duplicates, still symmetrised and deduplicated
    coo = torch.tensor([[0, 3], [3, 0], [2, 1], [0, 3]])
    rp, col = torchff.NeighborList.convert_pairs_coo_to_csr(coo, 5)
    assert rp.tolist() == [0, 1, 2, 3, 4, 4]
    assert col.tolist() == [3, 2, 1, 0]


@pytest.mark.parametrize("K,order", [(24, 6), (96, 5), (25, 5), (10, 6), (216, 5)])
def test_moduli_match_oracle(K, order):
    m = bsplines.compute_bspline_moduli_1d(K, order=order).numpy()
    np.testing.assert_allclose(m ** 2, obs.theta_sq(K, order), rtol=1e-13)
    z = bsplines.compute_bspline_moduli_z(K, order=order)
    assert z.shape == (K // 2 + 1,)


def test_workloads_are_reproducible():
    a = make_workload("water900")
    b = make_workload("water900")
    np.testing.assert_array_equal(a["coords"], b["coords"])
    assert a["natoms"] == 900 and abs(a["alpha"] - 2.92029) < 1e-5
    w = WORKLOADS["water96k"]
    assert w["n_waters"] * 3 == 96000 and w["K"] == 96 and w["order"] == 5
    # rigid geometry
    c = a["coords"].reshape(-1, 3, 3)
    d1 = np.linalg.norm(c[:, 1] - c[:, 0], axis=1)
    ang = np.degrees(np.arccos(np.sum((c[:, 1] - c[:, 0]) * (c[:, 2] - c[:, 0]), 1) / d1 ** 2))
    np.testing.assert_allclose(d1, 0.09572, atol=1e-12)
    np.testing.assert_allclose(ang, 104.52, atol=1e-9)


def test_host_box_is_read_on_every_call():
    # a host box is never cached: a write through .data (which autograd does not version) is seen
    b = torch.eye(3, dtype=torch.float64) * 3.0
    assert _lib.box_numpy(b)[2, 2] == 3.0
    b.data.mul_(1.01)
    assert _lib.box_numpy(b)[2, 2] == 3.0 * 1.01
    b[0, 0] = 5.0
    assert _lib.box_numpy(b)[0, 0] == 5.0
    # the cache switch and invalidation are plain host state
    torchff.set_box_cache(False)
    torchff.set_box_cache(True)
    torchff.invalidate_box_cache(b)
