"""FFT-free Ewald reciprocal sum for point charges (test infrastructure only).

Restates the rank-0 path of ``Ewald._forward_python`` (torchff/ewald.py:60-97):
half-space integer vectors m (ewald.py:35-44), structure factor
S(k) = sum_j q_j exp(2 pi i k.r_j), energy
(1/(pi V)) sum_half exp(-pi^2 k^2/alpha^2)/k^2 |S|^2 - alpha/sqrt(pi) sum q^2
(ewald.py:86-93).  Reciprocal vectors are taken as k = B^-1 m (the columns of
B^-1 are the reciprocal lattice vectors for row-major lattice vectors B);
torchff/ewald.py:66 forms m B^-1, which is the same for the cubic and
orthorhombic boxes used here.

Also :func:`madelung_nacl`, the textbook check that pins the complete Ewald
decomposition (real + reciprocal + self).
"""

import math

import numpy as np
import torch


def half_space_m(kmax):
    r = np.arange(-kmax, kmax + 1)
    mx, my, mz = np.meshgrid(np.arange(0, kmax + 1), r, r, indexing="ij")
    m = np.stack([mx.ravel(), my.ravel(), mz.ravel()], 1)
    keep = (m[:, 0] > 0) | ((m[:, 0] == 0) & (m[:, 1] > 0)) | ((m[:, 0] == 0) & (m[:, 1] == 0) & (m[:, 2] > 0))
    return m[keep]


def ewald_recip_energy(coords, box, q, kmax, alpha, chunk=4096, p=None, th=None):
    """Reciprocal + self energy (Gaussian units), float64 torch, differentiable in coords, q, p, th.

    With point dipoles p and Cartesian quadrupoles th (N, 3, 3) (ranks 1 and 2,
    torchff/ewald.py:75-97): L = q + i 2 pi k.p - (4 pi^2/3) k.th.k, S = sum_j L_j
    exp(2 pi i k.r_j), and the self terms (2 alpha^3 / (3 sqrt(pi))) sum |p|^2 and
    (8 alpha^5 / (45 sqrt(pi))) sum th:th (traceless quadrupoles)."""
    inv = torch.linalg.inv(box)
    V = torch.linalg.det(box)
    m = torch.as_tensor(half_space_m(kmax), dtype=coords.dtype)
    e = torch.zeros((), dtype=coords.dtype)
    for a in range(0, len(m), chunk):
        k = m[a:a + chunk] @ inv.T  # k_j = sum_i inv[j, i] m_i
        ksq = (k * k).sum(1)
        g = torch.exp(-math.pi ** 2 * ksq / (alpha * alpha)) / ksq
        kr = 2.0 * math.pi * (k @ coords.T)
        c, s = torch.cos(kr), torch.sin(kr)
        if th is not None:
            kk = torch.einsum("mx,nxy,my->mn", k, th, k)
            qe = q[None, :] - (4.0 * math.pi ** 2 / 3.0) * kk
            li = 2.0 * math.pi * (k @ p.T) if p is not None else torch.zeros_like(qe)
            sr = (c * qe - s * li).sum(1)
            si = (c * li + s * qe).sum(1)
        elif p is None:
            sr = (c * q[None, :]).sum(1)
            si = (s * q[None, :]).sum(1)
        else:
            li = 2.0 * math.pi * (k @ p.T)  # (M, N)
            sr = (c * q[None, :] - s * li).sum(1)
            si = (c * li + s * q[None, :]).sum(1)
        e = e + (g * (sr * sr + si * si)).sum()
    e = e / (math.pi * V)
    e = e - alpha / math.sqrt(math.pi) * torch.sum(q * q)
    if p is not None:
        e = e - 2.0 * alpha ** 3 / (3.0 * math.sqrt(math.pi)) * torch.sum(p * p)
    if th is not None:
        e = e - 8.0 * alpha ** 5 / (45.0 * math.sqrt(math.pi)) * torch.sum(th * th)
    return e


def real_space_all_images(coords, box, q, alpha, rmax):
    """Real-space Ewald sum over all periodic images within rmax (small systems only)."""
    coords = torch.as_tensor(coords, dtype=torch.float64)
    b = torch.as_tensor(box, dtype=torch.float64)
    inv = torch.linalg.inv(b)
    widths = 1.0 / torch.linalg.vector_norm(inv, dim=0)
    nimg = [int(math.ceil(rmax / float(w))) + 1 for w in widths]
    e = torch.zeros((), dtype=torch.float64)
    n = len(coords)
    d0 = coords[:, None, :] - coords[None, :, :]
    qq = q[:, None] * q[None, :]
    for a in range(-nimg[0], nimg[0] + 1):
        for bb in range(-nimg[1], nimg[1] + 1):
            for c in range(-nimg[2], nimg[2] + 1):
                shift = a * b[0] + bb * b[1] + c * b[2]
                d = d0 + shift
                r = torch.linalg.vector_norm(d, dim=-1)
                if a == 0 and bb == 0 and c == 0:
                    r = r + torch.eye(n, dtype=torch.float64) * 1e300
                mask = r <= rmax
                e = e + 0.5 * torch.sum(torch.where(mask, qq * torch.erfc(alpha * r) / r, torch.zeros_like(r)))
    return e


MADELUNG_NACL = 1.747564594633


def nacl_supercell(ncell, a=1.0):
    """Rock-salt supercell: ncell^3 conventional cells of edge a, 8 ions each; returns coords, q, box."""
    base = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0]])
    na = base
    cl = base + np.array([.5, 0, 0])
    pts, qs = [], []
    for i in range(ncell):
        for j in range(ncell):
            for k in range(ncell):
                o = np.array([i, j, k], dtype=np.float64)
                pts.append((na + o) * a)
                qs.append(np.ones(4))
                pts.append((cl + o) * a)
                qs.append(-np.ones(4))
    coords = np.concatenate(pts)
    q = np.concatenate(qs)
    box = np.eye(3) * a * ncell
    return coords, q, box
