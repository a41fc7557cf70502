"""Minimum-image convention and box inverse (oracle; test infrastructure only).

Follows csrc/common/pbc.cuh:6-18 (``apply_pbc_triclinic``: s = r.B^-1,
s -= round(s), r = s.B with round-half-away-from-zero) and pbc.cuh:38-64
(``invert_box_3x3``: cofactor / determinant).  torchff/pbc.py:86-97 uses
floor(s + 0.5) instead of round(s); the two differ only when a fractional
component is exactly +-0.5, which the canonical predicate below resolves the
CUDA way.

The arithmetic is written out operation by operation, in the order the HIP
kernels evaluate it with floating-point contraction disabled, so that the
float64 neighbor-list predicate is bit-identical on CPU and GPU.
"""

import numpy as np
import torch


def invert_box(box):
    """Row-major 3x3 box inverse via cofactors; returns (inv (9,), det).

    Restates csrc/common/pbc.cuh:38-64 exactly (same products, same order).
    ``box`` is a (3,3) array-like of float64.
    """
    m = np.asarray(box, dtype=np.float64).reshape(9)
    m00, m01, m02, m10, m11, m12, m20, m21, m22 = m
    c00 = m11 * m22 - m12 * m21
    c01 = m12 * m20 - m10 * m22
    c02 = m10 * m21 - m11 * m20
    c10 = m02 * m21 - m01 * m22
    c11 = m00 * m22 - m02 * m20
    c12 = m01 * m20 - m00 * m21
    c20 = m01 * m12 - m02 * m11
    c21 = m02 * m10 - m00 * m12
    c22 = m00 * m11 - m01 * m10
    det = m00 * c00 + m01 * c01 + m02 * c02
    inv_det = 1.0 / det
    inv = np.array([c00 * inv_det, c10 * inv_det, c20 * inv_det,
                    c01 * inv_det, c11 * inv_det, c21 * inv_det,
                    c02 * inv_det, c12 * inv_det, c22 * inv_det], dtype=np.float64)
    return inv, det


def round_half_away(x):
    """C ``round()``: nearest integer, halves away from zero (numpy rounds halves to even)."""
    t = np.trunc(x)
    return t + np.where(np.abs(x - t) >= 0.5, np.sign(x), 0.0)


def min_image_np(d, box, box_inv=None):
    """Minimum image of displacement rows d (P,3) float64, CUDA operation order.

    pbc.cuh:9-17: sa = d.x*inv0 + d.y*inv3 + d.z*inv6 (left to right, no FMA),
    sa -= round(sa); out.x = sa*b0 + sb*b3 + sc*b6.
    """
    if box is None:
        return np.asarray(d, dtype=np.float64)
    b = np.asarray(box, dtype=np.float64).reshape(9)
    inv = invert_box(b)[0] if box_inv is None else box_inv
    d = np.asarray(d, dtype=np.float64)
    x, y, z = d[:, 0], d[:, 1], d[:, 2]
    sa = x * inv[0] + y * inv[3] + z * inv[6]
    sb = x * inv[1] + y * inv[4] + z * inv[7]
    sc = x * inv[2] + y * inv[5] + z * inv[8]
    sa = sa - round_half_away(sa)
    sb = sb - round_half_away(sb)
    sc = sc - round_half_away(sc)
    ox = sa * b[0] + sb * b[3] + sc * b[6]
    oy = sa * b[1] + sb * b[4] + sc * b[7]
    oz = sa * b[2] + sb * b[5] + sc * b[8]
    return np.stack([ox, oy, oz], axis=1)


def min_image_torch(d, box):
    """Differentiable minimum image for autograd oracles (torchff/pbc.py:86-97).

    The integer shift is computed without gradient (as in the reference, where
    floor() has zero derivative), so d(out)/d(d) is the identity.
    """
    if box is None:
        return d
    box_inv = torch.linalg.inv(box)
    s = d @ box_inv
    shift = torch.round(s.detach())
    # torch.round is half-to-even; use half-away to match the CUDA predicate.
    frac = (s.detach() - torch.trunc(s.detach())).abs()
    shift = torch.where(frac == 0.5, torch.trunc(s.detach()) + torch.sign(s.detach()), shift)
    return (s - shift) @ box
