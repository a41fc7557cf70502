"""Real-space pair-energy oracles in float64 torch (CPU), differentiable (test infrastructure only).

Every function takes torch tensors; gradients (forces, charge / sigma /
epsilon derivatives, and second derivatives for force matching) come from
torch.autograd on these restatements.  Pair rows with a negative index are
padding and are skipped, as every reference kernel does
(csrc/nonbonded/nonbonded_atom_pairs_cuda.cu:57, multipoles_forward.cuh:389).

Pinned by tests/test_oracle_realspace.py: the LJ pair minimum at
r = 2^(1/6) sigma equals -epsilon, erfc(a r) - 1 = -erf(a r) for the
exclusion term, Coulomb of an isolated pair equals k q_i q_j / r, autograd
forces agree with central finite differences (h = 1e-5), and the
Lorentz-Berthelot table built from per-atom parameters reproduces the
per-atom mixing.
"""

import math

import torch

from .pbc import min_image_torch


def _valid(pairs):
    pairs = pairs.to(torch.int64)
    keep = (pairs[:, 0] >= 0) & (pairs[:, 1] >= 0)
    return pairs[keep]


def _pair_r(coords, box, pairs, direction="ij"):
    i, j = pairs[:, 0], pairs[:, 1]
    d = coords[i] - coords[j] if direction == "ij" else coords[j] - coords[i]
    d = min_image_torch(d, box)
    return d, torch.linalg.vector_norm(d, dim=1)


def nonbonded_energy(coords, pairs, box, sigma, epsilon, charges, coul_constant, cutoff, do_shift=True):
    """Shifted Coulomb + LB-mixed LJ (``compute_nonbonded_energy_from_atom_pairs``).

    Coulomb: torchff/coulomb.py:22-44 (``k q_i q_j (1/r - [1/rc])`` for r <= rc);
    LJ: csrc/nonbonded/nonbonded_atom_pairs_cuda.cu:86-89
    (``4 eps_ij s^6 (s^6 - 1)``, s = sigma_ij / r, sigma_ij = (s_i + s_j)/2,
    eps_ij = sqrt(e_i e_j)); cutoff test ``r^2 > rc^2 -> skip`` (:67-70).
    """
    p = _valid(pairs)
    _, r = _pair_r(coords, box, p)
    mask = (r <= cutoff).to(r.dtype)
    i, j = p[:, 0], p[:, 1]
    rinv = 1.0 / r
    if do_shift:
        rinv = rinv - 1.0 / cutoff
    e_coul = charges[i] * charges[j] * rinv * coul_constant
    sij = 0.5 * (sigma[i] + sigma[j])
    eij = torch.sqrt(epsilon[i] * epsilon[j])
    s6 = (sij / r) ** 6
    e_lj = 4.0 * eij * s6 * (s6 - 1.0)
    return torch.sum((e_coul + e_lj) * mask)


def ewald_real_energy(coords, box, pairs, pairs_excl, q, cutoff, ewald_alpha, prefactor):
    """Rank-0 ``MultipolarInteraction`` energy (torchff/multipoles.py:244-341).

    pref * [ sum_{pairs, r<=rc} q_i q_j erfc(a r)/r
             + sum_{excl, r<=rc} q_i q_j (erfc(a r) - 1)/r ].
    Damping erfc(a r) from multipoles.py:106-110; exclusion correction
    ``damps - 1.0`` (multipoles.py:256-257); exclusions are added only when
    alpha > 0 (multipoles.py:315; at alpha = 0 the term vanishes identically,
    so this equals the CUDA rule alpha >= 0, multipoles.cu:54).  alpha < 0
    means plain 1/r.
    """
    def part(pp, is_excl):
        pp = _valid(pp)
        _, r = _pair_r(coords, box, pp, direction="ji")
        mask = (r <= cutoff).to(r.dtype)
        i, j = pp[:, 0], pp[:, 1]
        if ewald_alpha >= 0:
            damp = torch.erfc(ewald_alpha * r)
            if is_excl:
                damp = damp - 1.0
            it = damp / r
        else:
            it = 1.0 / r
        return prefactor * torch.sum(q[i] * q[j] * it * mask)

    e = part(pairs, False)
    if pairs_excl is not None and ewald_alpha > 0:
        e = e + part(pairs_excl, True)
    return e


def lj_energy(coords, pairs, box, sigma, epsilon, cutoff, atom_types=None):
    """``Vdw('LennardJones')`` energy (torchff/vdw.py:86-111, 247-261).

    With ``atom_types`` sigma/epsilon are (T, T) tables indexed by the pair's
    types (vdw.py:196-201); otherwise they are per-pair (P,) values.
    Pairs with r > cutoff contribute zero (vdw.py:256-257).
    """
    pairs = pairs.to(torch.int64)
    keep = (pairs[:, 0] >= 0) & (pairs[:, 1] >= 0)
    p = pairs[keep]
    _, r = _pair_r(coords, box, p, direction="ji")
    if atom_types is not None:
        ti, tj = atom_types[p[:, 0]], atom_types[p[:, 1]]
        sij, eij = sigma[ti, tj], epsilon[ti, tj]
    else:
        sij, eij = sigma[keep], epsilon[keep]
    s6 = (sij / r) ** 6
    e = 4.0 * eij * s6 * (s6 - 1.0)
    return torch.sum(torch.where(r <= cutoff, e, torch.zeros_like(e)))


def ewald_lj_energy(coords, pairs, box, sigma, epsilon, charges, coul_constant, cutoff,
                    ewald_alpha, pairs_excl=None):
    """Fused real-space erfc-Coulomb + LB-mixed LJ with exclusion correction.

    The additive extension of ``compute_nonbonded_energy_from_atom_pairs``:
    Coulomb part = :func:`ewald_real_energy` (prefactor = coul_constant),
    LJ part = per-atom Lorentz-Berthelot as in :func:`nonbonded_energy`.
    With ewald_alpha < 0 it reduces to the unshifted nonbonded energy.
    """
    e_c = ewald_real_energy(coords, box, pairs, pairs_excl, charges, cutoff, ewald_alpha, coul_constant)
    p = _valid(pairs)
    _, r = _pair_r(coords, box, p)
    mask = (r <= cutoff).to(r.dtype)
    i, j = p[:, 0], p[:, 1]
    sij = 0.5 * (sigma[i] + sigma[j])
    eij = torch.sqrt(epsilon[i] * epsilon[j])
    s6 = (sij / r) ** 6
    e_lj = torch.sum(4.0 * eij * s6 * (s6 - 1.0) * mask)
    return e_c + e_lj


def lb_tables(sigma_atom_types, eps_atom_types):
    """Lorentz-Berthelot (T, T) tables from per-type sigma/eps (examples/tip3p/md.py:202-214)."""
    s = torch.as_tensor(sigma_atom_types, dtype=torch.float64)
    e = torch.as_tensor(eps_atom_types, dtype=torch.float64)
    return 0.5 * (s[:, None] + s[None, :]), torch.sqrt(e[:, None] * e[None, :])


ERFC_SELF = 1.0 / math.sqrt(math.pi)
