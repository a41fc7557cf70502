"""Host side of the batched serving path: batch tables, capacities, grid sizes (CPU)."""

import ctypes

import numpy as np
import pytest
import torch

from torchff import _lib as L
from torchff.serving import ModuliTable, pme_grid_size, request_stream, smooth_size


@pytest.fixture(scope="module")
def lib():
    return L.load(require_gpu=False)


def _table(lib, max_sys, max_atoms=1 << 14):
    buf = np.zeros(lib.tff_serve_table_bytes(max_sys, max_atoms), dtype=np.uint8)
    return buf, ctypes.c_void_p(buf.ctypes.data)


def _totals(lib, tp):
    t = np.zeros(8, dtype=np.int64)
    assert lib.tff_serve_table_totals(tp, ctypes.c_void_p(t.ctypes.data)) == 0
    return t


def _add(lib, tp, caps, atom_off, natoms, Lb, K, cutoff=0.9):
    Ls = (ctypes.c_double * 3)(Lb, Lb, Lb)
    Ks = (ctypes.c_int * 3)(K, K, K)
    mo = (ctypes.c_int * 3)(0, 0, 0)
    return lib.tff_serve_table_add(tp, L.F32, *caps, atom_off, natoms, Ls, Ks, mo, cutoff, 3.47)


def _tiles(K):
    # z and x passes: K^2 lines in tiles of 16; y pass: K lines of K values per (x, z) tile of
    # CB = the largest power of two <= min(16, K) z values
    cb = 1
    while cb * 2 <= min(16, K):
        cb *= 2
    return [-(-K * K // 16), K * -(-K // cb), -(-K * K // 16)]


def test_grid_sizes():
    assert [pme_grid_size(x) for x in (0.843, 1.0, 1.44, 3.1)] == [9, 10, 15, 32]
    assert smooth_size(2 * 3 * 5 * 16) and not smooth_size(14) and not smooth_size(77)


def test_table_totals_and_offsets(lib):
    buf, tp = _table(lib, 8)
    caps = (8, 1 << 14, 32, 1 << 20)
    lib.tff_serve_table_reset(tp)
    sizes = [(60, 0.843, 9), (3000, 3.1, 32), (300, 1.44, 15)]
    off = 0
    for n, Lb, K in sizes:
        assert _add(lib, tp, caps, off, n, Lb, K) == 0
        off += n
    t = _totals(lib, tp)
    assert list(t[:3]) == [3, 3360, 1 + 47 + 5]
    tiles = np.sum([_tiles(K) for _, _, K in sizes], axis=0)
    assert list(t[3:6]) == list(tiles)
    assert t[6] == 9 ** 3 + 32 ** 3 + 15 ** 3
    assert t[7] == 3360
    lib.tff_serve_table_reset(tp)
    assert list(_totals(lib, tp)) == [0] * 8
    # the atom records start after the head and the records, 256-byte aligned
    off = lib.tff_serve_atoms_offset(8)
    assert off % 256 == 0 and lib.tff_serve_table_bytes(8, 1 << 14) == off + 64 * (1 << 14)


def test_table_refuses_what_does_not_fit(lib):
    buf, tp = _table(lib, 2, 4000)
    caps = (2, 4000, 32, 40000)
    lib.tff_serve_table_reset(tp)
    assert _add(lib, tp, caps, 0, 3000, 3.1, 32) == 0          # 3000 atoms, 32768 points
    before = _totals(lib, tp).copy()
    assert _add(lib, tp, caps, 3000, 1200, 1.6, 16) == L.SERVE_FULL  # atoms exceeded
    assert _add(lib, tp, caps, 3000, 900, 2.08, 24) == L.SERVE_FULL  # grid points exceeded
    assert list(_totals(lib, tp)) == list(before)                      # unchanged
    assert _add(lib, tp, caps, 3000, 60, 0.843, 9) == 0
    assert _add(lib, tp, caps, 3060, 60, 0.843, 9) == L.SERVE_FULL    # systems exceeded


def test_table_rejects_bad_systems(lib):
    buf, tp = _table(lib, 4)
    caps = (4, 1 << 14, 32, 1 << 16)
    lib.tff_serve_table_reset(tp)
    assert _add(lib, tp, caps, 0, 60, 0.843, 14) not in (0, L.SERVE_FULL)   # 14 = 2 x 7
    assert _add(lib, tp, caps, 0, 60, 0.843, 36) not in (0, L.SERVE_FULL)   # above max_k
    assert _add(lib, tp, caps, 0, 0, 0.843, 9) not in (0, L.SERVE_FULL)     # no atoms
    assert _add(lib, tp, caps, 0, 60, 0.05, 9) not in (0, L.SERVE_FULL)     # > 16 images per axis
    assert list(_totals(lib, tp)) == [0] * 8


@pytest.mark.parametrize("fill", [0xFF, 0x7F, 0x01])
def test_table_refuses_an_unreset_head(lib, fill):
    # a table whose head was never reset (pinned memory comes uninitialised) is refused,
    # not indexed by its garbage system count
    buf, tp = _table(lib, 4)
    buf[:] = fill
    caps = (4, 1 << 14, 32, 1 << 16)
    assert _add(lib, tp, caps, 0, 60, 0.843, 9) not in (0, L.SERVE_FULL)
    assert "reset" in lib.tff_last_error().decode()
    lib.tff_serve_table_reset(tp)
    assert _add(lib, tp, caps, 0, 60, 0.843, 9) == 0


def test_workspace_lds_limit(lib):
    # a slab transform (two power-of-two line sets of edge + 1 values) must fit 64 KiB of LDS:
    # max_k <= 62 (float32), 32 (float64)
    assert lib.tff_serve_workspace_bytes(L.F32, 4, 1000, 62, 1 << 20) > 0
    assert lib.tff_serve_workspace_bytes(L.F32, 4, 1000, 63, 1 << 20) == 0
    assert lib.tff_serve_workspace_bytes(L.F64, 4, 1000, 32, 1 << 20) > 0
    assert lib.tff_serve_workspace_bytes(L.F64, 4, 1000, 33, 1 << 20) == 0


def test_moduli_table_offsets():
    m = ModuliTable(5, 32, torch.float64, "cpu")
    ks = sorted(m.offsets)
    assert ks[:8] == [1, 2, 3, 4, 5, 6, 8, 9]
    for a, b in zip(ks, ks[1:]):
        assert m.offsets[b] - m.offsets[a] == a
    assert m.table.numel() == m.offsets[ks[-1]] + ks[-1]


def test_request_stream_statistics():
    reqs = request_stream(4096, seed=11)
    n = np.array([r.natoms // 3 for r in reqs])
    assert n.min() >= 20 and n.max() <= 1000
    # log-uniform: the median is near sqrt(20 * 1000) = 141
    assert 110 < np.median(n) < 180
    gaps = np.diff([r.arrival for r in reqs])
    assert abs(np.mean(gaps) * 2000.0 - 1.0) < 0.05
    r = reqs[0]
    assert r.grid[0] >= r.box_lengths[0] / 0.1 and smooth_size(r.grid[0])
