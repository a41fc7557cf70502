"""Smooth-PME reciprocal-space oracle, any spline order, float64 torch (test infrastructure only).

Restates torchff/pme_ref.py for rank 0 (point charges), generalised from the
hard-wired order 6 to order p (ranks 1 and 2, point dipoles and quadrupoles, in
:func:`spread_multipoles` / :func:`pme_energy_multipoles`):

* scaled fractional coordinates s = K (r . B^-1) (pme_ref.py:11-17,24);
* anchor and spline arguments from :func:`oracle.bspline.anchor`
  (pme_ref.py:25-26 for order 6);
* charge spreading onto the K1 x K2 x K3 grid with periodic wrap
  (``spread_Q`` / ``Q_mesh_on_m``, pme_ref.py:111-130,145-167);
* S = FFT(Q), Phi = IFFT_{norm=forward}(C(k) S / |theta(k)|^2) with
  C(k) = 4 pi / (V k^2) exp(-k^2 / 4 alpha^2), C(0) = 0
  (``get_pme_recip``, pme_ref.py:250-282; k from signed indices with the
  Nyquist index taken positive, csrc/pme/kernels.cuh:506-512);
* potential at the atoms by interpolation (pme_ref.py:334-372) and the self
  term -2 alpha/sqrt(pi) q (pme_ref.py:434-445);
* energy = 1/2 sum_i q_i phi_i (torchff/pme.py:96-100).

Everything is differentiable, so forces (-dE/dr) and charge derivatives
(dE/dq = phi) come from autograd; :func:`pme_gather` gives the same
quantities analytically from the potential grid, as the reference's backward
kernel does (csrc/pme/kernels.cuh:191-470).

Pinned by tests/test_oracle_pme.py against the FFT-free Ewald sum of
:mod:`oracle.ewald` (relative error falls with K and p), the NaCl Madelung
constant, order-6 moduli against the reference's cosine form, the energy
identity 1/2 sum q phi = 1/2 sum_k G |S|^2, and finite-difference forces.
The multipolar energy is pinned by tests/test_oracle_dipoles.py against the
direct Ewald sum with L = q + 2 pi i k.p - (4 pi^2/3) k.theta.k, against the
limit of a charge pair (dipole) and of a four-charge cluster (quadrupole) with
their intra-cluster Gaussian terms removed in closed form, and by finite
differences.
"""

import math

import numpy as np
import torch

from . import bspline


def _ktuple(K):
    if isinstance(K, (tuple, list)):
        return tuple(int(k) for k in K)
    return (int(K),) * 3


def _weights(coords, box, K, order):
    """Per-atom, per-axis spline weights (N,3,p), derivative weights, and grid indices (N,3,p)."""
    Ks = torch.tensor(_ktuple(K), dtype=coords.dtype)
    inv = torch.linalg.inv(box)
    s = (coords @ inv) * Ks
    m0, u0 = bspline.anchor(s, order)
    sh = torch.as_tensor(bspline.stencil(order), dtype=coords.dtype)
    u = u0[:, :, None] + sh[None, None, :]
    M = bspline.bspline_M(u, order)
    dM = bspline.bspline_dM(u, order)
    idx = (m0.detach().to(torch.int64)[:, :, None] + sh.to(torch.int64)[None, None, :])
    idx = torch.remainder(idx, torch.tensor(_ktuple(K), dtype=torch.int64)[None, :, None])
    return M, dM, idx, inv


def _flat_index(idx, K):
    K0, K1, K2 = _ktuple(K)
    gx = idx[:, 0, :, None, None]
    gy = idx[:, 1, None, :, None]
    gz = idx[:, 2, None, None, :]
    return (gx * K1 + gy) * K2 + gz


def influence(box, K, alpha, order):
    """G(k) = C(k) / |theta|^2 on the full K1 x K2 x K3 grid (float64 torch), G(0) = 0."""
    K0, K1, K2 = _ktuple(K)
    inv = torch.linalg.inv(box.detach())
    V = torch.linalg.det(box.detach())

    def signed(n):
        m = torch.arange(n, dtype=torch.float64)
        return torch.where(m <= n // 2, m, m - n)

    mx, my, mz = signed(K0), signed(K1), signed(K2)
    m = torch.stack(torch.meshgrid(mx, my, mz, indexing="ij"), dim=-1)  # (K0,K1,K2,3)
    k = 2.0 * math.pi * torch.einsum("ji,xyzi->xyzj", inv, m)
    ksq = (k * k).sum(-1)
    ksq[0, 0, 0] = 1.0
    C = 4.0 * math.pi / (V * ksq) * torch.exp(-ksq / (4.0 * alpha * alpha))
    C[0, 0, 0] = 0.0
    th = [torch.as_tensor(bspline.theta_sq(n, order)) for n in (K0, K1, K2)]
    tsq = th[0][:, None, None] * th[1][None, :, None] * th[2][None, None, :]
    return C / tsq


def spread(coords, box, q, K, order):
    """Charge grid Q (K0,K1,K2): Q[m0+j] += q M(u0+j) (pme_ref.py:111-130, 164-167)."""
    M, _, idx, _ = _weights(coords, box, K, order)
    W = M[:, 0, :, None, None] * M[:, 1, None, :, None] * M[:, 2, None, None, :]
    flat = _flat_index(idx, K)
    K0, K1, K2 = _ktuple(K)
    Q = torch.zeros(K0 * K1 * K2, dtype=coords.dtype)
    Q = Q.index_add(0, flat.reshape(-1), (q[:, None, None, None] * W).reshape(-1))
    return Q.reshape(K0, K1, K2)


def potential_grid(Q, box, K, alpha, order):
    """Phi = IFFT_forward(G(k) FFT(Q)) real part (pme_ref.py:276-282)."""
    G = influence(box, K, alpha, order).to(Q.dtype)
    S = torch.fft.fftn(Q)
    return torch.fft.ifftn(G * S, norm="forward").real


def pme_energy(coords, box, q, K, alpha, order, return_phi=False):
    """Rank-0 PME energy (reciprocal + self), in Gaussian units (multiply by the Coulomb constant)."""
    Q = spread(coords, box, q, K, order)
    Phi = potential_grid(Q, box, K, alpha, order)
    M, _, idx, _ = _weights(coords, box, K, order)
    W = M[:, 0, :, None, None] * M[:, 1, None, :, None] * M[:, 2, None, None, :]
    flat = _flat_index(idx, K)
    phi = (Phi.reshape(-1)[flat] * W).sum(dim=(1, 2, 3))
    phi = phi - 2.0 * alpha / math.sqrt(math.pi) * q
    E = 0.5 * torch.sum(q * phi)
    if return_phi:
        return E, phi
    return E


def _lattice_R(box, K, dtype):
    """R_ax = K_a (B^-1)_{xa}: ds_a/dr_x, the reference's Nj_Aji_star (pme_ref.py:15-20)."""
    Ks = torch.tensor(_ktuple(K), dtype=dtype)
    return Ks[:, None] * torch.linalg.inv(box).T


def spread_multipoles(coords, box, q, p, th, K, order):
    """Charge grid of point charges, dipoles and Cartesian quadrupoles (pme_ref.py:145-235):
    Q(m) += q W(m) + p . grad_r W(m) + (1/3) theta : grad_r grad_r W(m).
    With u = m0 - s + const (du/ds = -1) and ds_a/dr_x = R_ax:
    p . grad_r W = -sum_a pu_a dW/du_a with pu = R p (the reference's W_tot -= p_u . grad_W_u), and
    theta : grad grad W = sum_ab (R theta R^T)_ab d2W/du_a du_b (its Theta_u contraction, scale 1/3).
    p (N, 3) and th (N, 3, 3) may be None; q may be zeros."""
    M, dM, idx, inv = _weights(coords, box, K, order)
    W = M[:, 0, :, None, None] * M[:, 1, None, :, None] * M[:, 2, None, None, :]
    val = q[:, None, None, None] * W
    R = _lattice_R(box, K, coords.dtype)
    if p is not None:
        pu = p @ R.T
        dW0 = dM[:, 0, :, None, None] * M[:, 1, None, :, None] * M[:, 2, None, None, :]
        dW1 = M[:, 0, :, None, None] * dM[:, 1, None, :, None] * M[:, 2, None, None, :]
        dW2 = M[:, 0, :, None, None] * M[:, 1, None, :, None] * dM[:, 2, None, None, :]
        val = val - (pu[:, 0, None, None, None] * dW0 + pu[:, 1, None, None, None] * dW1
                     + pu[:, 2, None, None, None] * dW2)
    if th is not None:
        Ks = torch.tensor(_ktuple(K), dtype=coords.dtype)
        s = (coords @ inv) * Ks
        _, u0 = bspline.anchor(s, order)
        sh = torch.as_tensor(bspline.stencil(order), dtype=coords.dtype)
        d2M = bspline.bspline_d2M(u0[:, :, None] + sh[None, None, :], order)
        U = torch.einsum("ax,nxy,by->nab", R, th, R)
        X = [M[:, a] for a in range(3)]
        D = [dM[:, a] for a in range(3)]
        D2 = [d2M[:, a] for a in range(3)]

        def outer(f0, f1, f2):
            return f0[:, :, None, None] * f1[:, None, :, None] * f2[:, None, None, :]

        H = {(0, 0): outer(D2[0], X[1], X[2]), (1, 1): outer(X[0], D2[1], X[2]), (2, 2): outer(X[0], X[1], D2[2]),
             (0, 1): outer(D[0], D[1], X[2]), (0, 2): outer(D[0], X[1], D[2]), (1, 2): outer(X[0], D[1], D[2])}
        quad = 0.0
        for (a, b_), h in H.items():
            c = U[:, a, b_] if a == b_ else U[:, a, b_] + U[:, b_, a]
            quad = quad + c[:, None, None, None] * h
        val = val + quad / 3.0
    flat = _flat_index(idx, K)
    K0, K1, K2 = _ktuple(K)
    Q = torch.zeros(K0 * K1 * K2, dtype=coords.dtype)
    Q = Q.index_add(0, flat.reshape(-1), val.reshape(-1))
    return Q.reshape(K0, K1, K2)


def pme_energy_multipoles(coords, box, q, p, th, K, alpha, order):
    """Multipolar PME energy (reciprocal + self) in Gaussian units: 1/2 sum_k G |FFT(Q)|^2 with Q
    from :func:`spread_multipoles`, minus alpha/sqrt(pi) sum (q^2 + (2 alpha^2/3) |p|^2
    + (8 alpha^4/45) theta : theta) (csrc/ewald/self_contribution.cuh:4-110; pme_ref.py:435-469;
    the quadrupole term is the self energy of traceless quadrupoles).  The reference's potential,
    field and field gradient are its derivatives: phi_i = dE/dq_i, E_i = -dE/dp_i,
    grad E_i = -3 dE/dtheta_i, F_i = -dE/dr_i."""
    Q = spread_multipoles(coords, box, q, p, th, K, order)
    G = influence(box, K, alpha, order)
    S = torch.fft.fftn(Q)
    e = 0.5 * torch.sum(G * (S.real ** 2 + S.imag ** 2)) - alpha / math.sqrt(math.pi) * torch.sum(q * q)
    if p is not None:
        e = e - 2.0 * alpha ** 3 / (3.0 * math.sqrt(math.pi)) * torch.sum(p * p)
    if th is not None:
        e = e - 8.0 * alpha ** 5 / (45.0 * math.sqrt(math.pi)) * torch.sum(th * th)
    return e


def pme_energy_kspace(coords, box, q, K, alpha, order):
    """Same energy from the k-space sum 1/2 sum_k G |S|^2 - alpha/sqrt(pi) sum q^2."""
    Q = spread(coords, box, q, K, order)
    G = influence(box, K, alpha, order)
    S = torch.fft.fftn(Q)
    return 0.5 * torch.sum(G * (S.real ** 2 + S.imag ** 2)) - alpha / math.sqrt(math.pi) * torch.sum(q * q)


def pme_gather(Phi, coords, box, q, K, alpha, order):
    """Analytic per-atom potential and force from the potential grid.

    phi_i = sum_g Phi(g) W_i(g) - 2 alpha/sqrt(pi) q_i and
    F_i = -q_i sum_g Phi(g) dW_i/dr (csrc/pme/kernels.cuh:256-470, rank 0;
    du/ds = -1, ds_a/dr_j = K_a (B^-1)_{j a}).
    """
    with torch.no_grad():
        M, dM, idx, inv = _weights(coords, box, K, order)
        flat = _flat_index(idx, K)
        ph = Phi.reshape(-1)[flat]
        W = M[:, 0, :, None, None] * M[:, 1, None, :, None] * M[:, 2, None, None, :]
        g0 = (ph * dM[:, 0, :, None, None] * M[:, 1, None, :, None] * M[:, 2, None, None, :]).sum((1, 2, 3))
        g1 = (ph * M[:, 0, :, None, None] * dM[:, 1, None, :, None] * M[:, 2, None, None, :]).sum((1, 2, 3))
        g2 = (ph * M[:, 0, :, None, None] * M[:, 1, None, :, None] * dM[:, 2, None, None, :]).sum((1, 2, 3))
        du = torch.stack([g0, g1, g2], 1)  # dphi/du_a
        Ks = torch.tensor(_ktuple(K), dtype=coords.dtype)
        # dphi/dr_j = sum_a dphi/du_a * (-1) * K_a inv[j, a]
        grad = -(du * Ks[None, :]) @ inv.T
        phi = (ph * W).sum((1, 2, 3)) - 2.0 * alpha / math.sqrt(math.pi) * q
        force = -q[:, None] * grad
    return phi, force


def ewald_error_alpha(cutoff, tol):
    """Ewald splitting parameter from a direct-space tolerance: sqrt(-ln(2 tol)) / rc (OpenMM rule)."""
    return math.sqrt(-math.log(2.0 * tol)) / cutoff


def grid_size_for_spacing(L, spacing):
    """Smallest K >= L/spacing of the form 2^a 3^b 5^c."""
    n = max(1, int(math.ceil(L / spacing - 1e-9)))
    while True:
        m = n
        for f in (2, 3, 5):
            while m % f == 0:
                m //= f
        if m == 1:
            return n
        n += 1


def as_numpy(x):
    return x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else np.asarray(x)
