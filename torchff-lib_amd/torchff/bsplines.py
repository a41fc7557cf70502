"""PME B-spline moduli of any order (drop-in for torchff/bsplines.py of the reference).

|theta(k)| = | sum_{j=0}^{p-2} M_p(j+1) exp(2 pi i k j / K) | (Essmann et al. 1995,
eq. 4.4), computed in float64.  For order 6 this equals the reference's
sum_{m=-3}^{2} M6(m+3) cos(2 pi m k / K) (torchff/bsplines.py:20-54); odd orders
keep the sine part, and a vanishing modulus (odd order, even K, Nyquist
index) is replaced by the mean of its neighbours, as OpenMM's PME does.
The reference stores these as float32 buffers; this build keeps float64 and
casts to the working dtype only when the module is moved.
"""

import math

import numpy as np
import torch


def cardinal_bspline(u, p):
    """M_p(u) by the Cox-de Boor recursion (numpy, float64)."""
    u = np.asarray(u, dtype=np.float64)
    m = [((u - k >= 0) & (u - k < 1)).astype(np.float64) for k in range(p)]
    for n in range(2, p + 1):
        m = [((u - k) * m[k] + (n - u + k) * m[k + 1]) / (n - 1) for k in range(p - n + 1)]
    return m[0]


def theta_sq(K, order=6):
    j = np.arange(order - 1, dtype=np.float64)
    w = cardinal_bspline(j + 1.0, order)
    k = np.arange(K, dtype=np.float64)
    arg = 2.0 * math.pi * np.outer(k, j) / K
    mod = (np.cos(arg) @ w) ** 2 + (np.sin(arg) @ w) ** 2
    for i in range(K):
        if mod[i] < 1e-7:
            mod[i] = 0.5 * (mod[(i - 1) % K] + mod[(i + 1) % K])
    return mod


def get_bspline_modulus(k, K, order=6):
    """|theta(k)| for integer tensor k (torchff/bsplines.py:20-54, any order)."""
    full = torch.as_tensor(np.sqrt(theta_sq(K, order)), dtype=torch.float64)
    return full[k.long().cpu()].to(k.device)


def compute_bspline_moduli_1d(K, dtype=torch.float64, device=None, order=6):
    """Moduli for k = 0..K-1 (x and y axes)."""
    return torch.as_tensor(np.sqrt(theta_sq(K, order)), dtype=torch.float64).to(dtype=dtype, device=device)


def compute_bspline_moduli_z(K3, dtype=torch.float64, device=None, order=6):
    """Moduli for k = 0..K3//2 (the half-spectrum axis)."""
    return compute_bspline_moduli_1d(K3, dtype, device, order)[: K3 // 2 + 1]
