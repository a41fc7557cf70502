"""Cardinal B-splines and PME moduli of general order (oracle; test infrastructure only).

The reference hard-codes order 6: explicit piecewise quintics in
torchff/pme_ref.py:30-109 and csrc/pme/bsplines.cuh:9-100, moduli
``b(k) = sum_{m=-3}^{2} M6(m+3) cos(2 pi m k / K)`` in torchff/bsplines.py:20-54.
The configurations here also need order 5, so this module evaluates M_p by
the Cox-de Boor recursion for any p and the modulus in Essmann et al.
(J. Chem. Phys. 103, 8577 (1995), eq. 4.4):

    |theta(k)|^2 = | sum_{j=0}^{p-2} M_p(j+1) exp(2 pi i k j / K) |^2,

which for even p equals the square of the reference's real cosine sum (the
centred weights are symmetric) and for odd p keeps the sine part that a cosine
sum would drop.  For odd p and even K the modulus vanishes at the Nyquist
index; such entries (|theta|^2 < 1e-7) are replaced by the mean of their two
neighbours, the fix used by OpenMM's reference PME.

Pinned by tests/test_oracle_bspline.py: the order-6 recursion against the
reference's explicit polynomial pieces (an independent formulation),
partition of unity sum_k M_p(u+k) = 1, sum_k M_p'(u+k) = 0, the integer
values M_5(1..4) = (1, 11, 11, 1)/24, and the order-6 moduli against the
cosine-sum form of torchff/bsplines.py.
"""

import math

import numpy as np
import torch


def bspline_M(u, p):
    """M_p(u) by Cox-de Boor; works on numpy arrays or torch tensors (differentiable)."""
    lib = torch if isinstance(u, torch.Tensor) else np
    # M_1 = indicator of [0, 1)
    if lib is torch:
        m = [((u - k >= 0) & (u - k < 1)).to(u.dtype) for k in range(p)]
    else:
        m = [((u - k >= 0) & (u - k < 1)).astype(np.float64) for k in range(p)]
    # m[k] holds M_n(u - k) for the current n; recursion
    # M_n(x) = (x M_{n-1}(x) + (n - x) M_{n-1}(x - 1)) / (n - 1)
    for n in range(2, p + 1):
        new = []
        for k in range(p - n + 1):
            x = u - k
            new.append((x * m[k] + (n - x) * m[k + 1]) / (n - 1))
        m = new
    return m[0]


def bspline_dM(u, p):
    """dM_p/du = M_{p-1}(u) - M_{p-1}(u - 1)."""
    return bspline_M(u, p - 1) - bspline_M(u - 1, p - 1)


def bspline_d2M(u, p):
    """d^2 M_p/du^2 = M_{p-2}(u) - 2 M_{p-2}(u-1) + M_{p-2}(u-2)."""
    return bspline_M(u, p - 2) - 2 * bspline_M(u - 1, p - 2) + bspline_M(u - 2, p - 2)


def theta_sq(K, p):
    """|theta(k)|^2 for k = 0..K-1 (float64 numpy), with the zero-modulus fix."""
    j = np.arange(p - 1, dtype=np.float64)
    w = bspline_M(j + 1.0, p)
    k = np.arange(K, dtype=np.float64)
    arg = 2.0 * math.pi * np.outer(k, j) / K
    re = (np.cos(arg) * w).sum(1)
    im = (np.sin(arg) * w).sum(1)
    mod = re * re + im * im
    for i in range(K):
        if mod[i] < 1e-7:
            mod[i] = 0.5 * (mod[(i - 1) % K] + mod[(i + 1) % K])
    return mod


def stencil(p):
    """Grid offsets of the centred stencil: -floor(p/2) .. p-1-floor(p/2).

    Order 6 gives -3..2 (torchff/pme_ref.py:327-332, csrc/pme/kernels.cuh:82-97).
    """
    h = p // 2
    return np.arange(-h, p - h)


def anchor(s, p):
    """Grid anchor m0 and local spline argument u0 for scaled fractional coordinate s = K*f.

    Order 6 (even): m0 = ceil(s), u0 = m0 - s + 3 (torchff/pme_ref.py:19-27,
    csrc/pme/kernels.cuh:64-69).  General p: m0 = ceil(s - (p mod 2)/2),
    u0 = m0 - s + p/2, so the stencil covers the p grid points nearest to s and
    M_p(u0 + j) weights grid point m0 + j.
    """
    lib = torch if isinstance(s, torch.Tensor) else np
    m0 = lib.ceil(s - 0.5 * (p % 2))
    u0 = (m0 - s) + 0.5 * p
    return m0, u0
