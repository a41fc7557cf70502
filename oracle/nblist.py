"""Neighbor-list oracle (test infrastructure only).

Canonical predicate (the one every list in this repository is checked
against): for atoms a != b (or a == b with include_self) that are not
excluded, keep the pair iff  |minimg(r_a - r_b)|^2 <= cutoff^2,  evaluated in
float64 on the (exactly promoted) input coordinates with the operation order
of csrc/common/pbc.cuh:9-17 and no fused multiply-add; cutoff^2 is formed in
float64.  This is the nsquared CUDA predicate (csrc/nblist/nblist_nsquared_cuda.cu:85-89)
promoted to float64; torchff/nblist.py:137-140 (``dr < cutoff``) and the
cell-list kernel (``norm <= cutoff``, nblist_clist_cuda.cu:293,337) disagree
with it only for pairs whose distance equals the cutoff to the last bit.

Exclusions follow ``NeighborList.convert_pairs_coo_to_csr``
(torchff/nblist.py:50-123): symmetrised, deduplicated CSR; a pair (i, j) is
excluded iff j appears in row i (csrc/nblist/exclusions.cuh:26-36).

Pair sets are returned canonically: int64 (P, 2) rows (i, j) with i < j
(i == j for self pairs), sorted lexicographically.

Pinned by tests/test_oracle_nblist.py: the cell-binned search equals the
all-pairs brute force on random and water boxes, cutoff 0 gives no pairs, a
cutoff above the half-diagonal gives all N(N-1)/2, counts match
(1/2) N rho (4/3) pi rc^3 for uniform density, exclusions are removed.
"""

import numpy as np

from .pbc import invert_box, min_image_np


def coo_to_csr(pairs_coo, natoms):
    """Symmetrise + deduplicate COO pairs into CSR (row_ptr (N+1,), col (M,)), int64.

    torchff/nblist.py:107-123.
    """
    p = np.asarray(pairs_coo, dtype=np.int64).reshape(-1, 2)
    allp = np.concatenate([p, p[:, ::-1]], axis=0)
    keys = np.unique(allp[:, 0] * natoms + allp[:, 1])
    rows = keys // natoms
    cols = keys % natoms
    counts = np.bincount(rows, minlength=natoms)
    row_ptr = np.zeros(natoms + 1, dtype=np.int64)
    row_ptr[1:] = np.cumsum(counts)
    return row_ptr, cols.astype(np.int64)


def _excluded_mask(i, j, row_ptr, col):
    """True where (i, j) is in the CSR exclusion set (exclusions.cuh:26-36)."""
    if col is None or len(col) == 0 or len(i) == 0:
        return np.zeros(len(i), dtype=bool)
    natoms = len(row_ptr) - 1
    rows = np.repeat(np.arange(natoms, dtype=np.int64), np.diff(row_ptr))
    ekeys = np.sort(rows * natoms + col)
    keys = i * natoms + j
    pos = np.minimum(np.searchsorted(ekeys, keys), len(ekeys) - 1)
    return ekeys[pos] == keys


def _predicate(coords, box, box_inv, cutoff2, i, j):
    d = coords[i] - coords[j]
    d = min_image_np(d, box, box_inv)
    d2 = d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1] + d[:, 2] * d[:, 2]
    return d2 <= cutoff2


def _canonical(i, j):
    lo = np.minimum(i, j)
    hi = np.maximum(i, j)
    out = np.stack([lo, hi], axis=1).astype(np.int64)
    if len(out) == 0:
        return out.reshape(0, 2)
    order = np.lexsort((out[:, 1], out[:, 0]))
    return out[order]


def brute_force(coords, box, cutoff, excl=None, include_self=False, chunk=2_000_000):
    """All-pairs O(N^2) list with the canonical predicate (torchff/nblist.py:39-48,137-140)."""
    coords = np.asarray(coords, dtype=np.float64)
    n = len(coords)
    box = None if box is None else np.asarray(box, dtype=np.float64)
    inv = None if box is None else invert_box(box)[0]
    cutoff2 = float(cutoff) * float(cutoff)
    row_ptr, col = (None, None) if excl is None else excl
    out_i, out_j = [], []
    # iterate rows in blocks so that memory stays bounded
    i0 = 0
    while i0 < n:
        i1 = i0 + 1
        budget = 0
        while i1 < n and budget + (n - i1) < chunk:
            budget += n - i1
            i1 += 1
        ii = [np.full(n - a - (0 if include_self else 1), a, dtype=np.int64) for a in range(i0, i1)]
        jj = [np.arange(a if include_self else a + 1, n, dtype=np.int64) for a in range(i0, i1)]
        i = np.concatenate(ii) if ii else np.zeros(0, np.int64)
        j = np.concatenate(jj) if jj else np.zeros(0, np.int64)
        keep = _predicate(coords, box, inv, cutoff2, i, j)
        selfp = i == j
        keep = np.where(selfp, include_self, keep)
        if col is not None:
            keep &= ~_excluded_mask(i, j, row_ptr, col)
        out_i.append(i[keep])
        out_j.append(j[keep])
        i0 = i1
    return _canonical(np.concatenate(out_i), np.concatenate(out_j))


def cell_search(coords, box, cutoff, excl=None, include_self=False):
    """Same set as :func:`brute_force`, found through cell binning (periodic box).

    Candidate generation only narrows the search: cells at least cutoff/reach
    wide along each reciprocal direction (reach 2 when 5 cells fit, else 1 when
    3 fit), a half shell of neighbour cells so every cell pair is visited once;
    every candidate passes through the identical canonical predicate.  Falls
    back to the brute force when fewer than 3 cells fit along an axis.
    """
    coords = np.asarray(coords, dtype=np.float64)
    box = np.asarray(box, dtype=np.float64)
    inv, _ = invert_box(box)
    invm = inv.reshape(3, 3)
    # perpendicular widths of the box: 1 / |column k of inv|
    widths = 1.0 / np.linalg.norm(invm, axis=0)
    if cutoff <= 0:
        return brute_force(coords, box, cutoff, excl, include_self)
    reach = None
    for R in (2, 1):
        nc = np.floor(widths * R / (cutoff * (1 + 1e-9))).astype(np.int64)
        if np.all(nc >= 2 * R + 1):
            reach = R
            break
    if reach is None:
        return brute_force(coords, box, cutoff, excl, include_self)
    frac = coords @ invm
    frac -= np.floor(frac)
    cell = np.minimum((frac * nc).astype(np.int64), nc - 1)
    cid = (cell[:, 0] * nc[1] + cell[:, 1]) * nc[2] + cell[:, 2]
    order = np.argsort(cid, kind="stable")
    ncell = int(nc.prod())
    start = np.searchsorted(cid[order], np.arange(ncell))
    end = np.searchsorted(cid[order], np.arange(ncell), side="right")
    cutoff2 = float(cutoff) * float(cutoff)
    row_ptr, col = (None, None) if excl is None else excl
    out_i, out_j = [], []
    side = 2 * reach + 1
    offs = [(a, b, c) for a in range(-reach, reach + 1) for b in range(-reach, reach + 1)
            for c in range(-reach, reach + 1)]
    half = offs[len(offs) // 2:]  # (0,0,0) first, then the positive half shell
    cx, cy, cz = np.unravel_index(np.arange(ncell), tuple(nc))
    cnt = end - start
    for (a, b, c) in half:
        same = (a, b, c) == (0, 0, 0)
        nid = (((cx + a) % nc[0]) * nc[1] + (cy + b) % nc[1]) * nc[2] + (cz + c) % nc[2]
        cnt_j = cnt[nid]
        tot = cnt * cnt_j
        if tot.sum() == 0:
            continue
        cells = np.repeat(np.arange(ncell), tot)
        base = np.repeat(np.cumsum(tot) - tot, tot)
        k = np.arange(tot.sum()) - base
        cj = cnt_j[cells]
        ia = order[start[cells] + k // cj]
        ja = order[start[nid[cells]] + k % cj]
        if same:
            sel = ia < ja if not include_self else ia <= ja
            ia, ja = ia[sel], ja[sel]
        keep = _predicate(coords, box, inv, cutoff2, ia, ja)
        keep = np.where(ia == ja, include_self, keep)
        ia, ja = ia[keep], ja[keep]
        if col is not None:
            ex = _excluded_mask(np.minimum(ia, ja), np.maximum(ia, ja), row_ptr, col)
            ia, ja = ia[~ex], ja[~ex]
        out_i.append(ia)
        out_j.append(ja)
    del side
    if not out_i:
        return np.zeros((0, 2), np.int64)
    return _canonical(np.concatenate(out_i), np.concatenate(out_j))


def neighbor_pairs(coords, box, cutoff, exclusions=None, include_self=False):
    """Oracle for ``NeighborList.forward`` (torchff/nblist.py:137-178): canonical sorted pair set."""
    n = len(coords)
    excl = None if exclusions is None else coo_to_csr(exclusions, n)
    if box is None:
        return brute_force(coords, None, cutoff, excl, include_self)
    return cell_search(coords, box, cutoff, excl, include_self)


def canonicalize(pairs, npairs=None):
    """Canonical form of a device/list output: drop -1 padding, (min, max), sort."""
    p = np.asarray(pairs, dtype=np.int64).reshape(-1, 2)
    if npairs is not None:
        p = p[: int(npairs)]
    p = p[(p[:, 0] >= 0) & (p[:, 1] >= 0)]
    return _canonical(p[:, 0], p[:, 1])
