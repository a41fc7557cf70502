"""TIP3P energy + forces composition oracle (test infrastructure only).

Restates ``Tip3pTorchFF.energy_components`` (examples/tip3p/model.py:100-123)
without the bonded terms (rigid waters): neighbor list at rc + skin
(model.py:102), LJ from the (2,2) Lorentz-Berthelot type table
(model.py:107-109, vdw.py:247-261), rank-0 Ewald real space with the
intramolecular exclusion correction scaled by 138.935456 (model.py:110-112),
and rank-0 PME times the same constant (model.py:113-114).  Forces are
-dE/dr by autograd in float64.
"""

import numpy as np
import torch

from . import nblist, pme, realspace
from .water import COULOMB_KJMOL


def pair_list(w, cutoff):
    """Canonical oracle neighbor list of a workload dict at the given cutoff (with its exclusions)."""
    return nblist.neighbor_pairs(w["coords"], w["box"], cutoff, exclusions=w["exclusions"])


def energy_terms(coords, box, w, pairs, cutoff, alpha, K, order):
    q = torch.as_tensor(w["charges"], dtype=torch.float64)
    types = torch.as_tensor(w["atom_types"])
    sig_t = torch.as_tensor(w["sigma_table"], dtype=torch.float64)
    eps_t = torch.as_tensor(w["epsilon_table"], dtype=torch.float64)
    pairs_t = torch.as_tensor(pairs)
    excl = torch.as_tensor(w["pairs_excl"])
    e_lj = realspace.lj_energy(coords, pairs_t, box, sig_t, eps_t, cutoff, atom_types=types)
    e_real = realspace.ewald_real_energy(coords, box, pairs_t, excl, q, cutoff, alpha, COULOMB_KJMOL)
    e_rec = pme.pme_energy(coords, box, q, K, alpha, order) * COULOMB_KJMOL
    return {"lj": e_lj, "coul_real": e_real, "coul_recip": e_rec}


def energy_forces(w, cutoff=None, alpha=None, K=None, order=None, skin=None, pairs=None):
    """Total energy (kJ/mol, float) and forces (N,3 float64 numpy) plus the per-term energies."""
    cutoff = w["cutoff"] if cutoff is None else cutoff
    alpha = w["alpha"] if alpha is None else alpha
    K = w["K"] if K is None else K
    order = w["order"] if order is None else order
    skin = w.get("skin", 0.0) if skin is None else skin
    if pairs is None:
        pairs = pair_list(w, cutoff + skin)
    coords = torch.tensor(w["coords"], dtype=torch.float64, requires_grad=True)
    box = torch.tensor(w["box"], dtype=torch.float64)
    terms = energy_terms(coords, box, w, pairs, cutoff, alpha, K, order)
    E = terms["lj"] + terms["coul_real"] + terms["coul_recip"]
    (g,) = torch.autograd.grad(E, coords)
    return float(E), -g.numpy(), {k: float(v) for k, v in terms.items()}, np.asarray(pairs)


def _lj_per_atom(coords, box, pairs, sigma, epsilon, cutoff):
    """Lorentz-Berthelot LJ from per-atom sigma / epsilon (nonbonded_atom_pairs_cuda.cu:80-134);
    eps_ij = sqrt(eps_i eps_j) written so that its gradient is 0 where eps_i eps_j = 0 (the
    convention of the build's kernels; the reference's sqrt gives inf there)."""
    p = realspace._valid(pairs)
    _, r = realspace._pair_r(coords, box, p)
    mask = (r <= cutoff).to(r.dtype)
    i, j = p[:, 0], p[:, 1]
    sij = 0.5 * (sigma[i] + sigma[j])
    prod = epsilon[i] * epsilon[j]
    pos = prod > 0
    eij = torch.sqrt(torch.where(pos, prod, torch.ones_like(prod))) * pos.to(prod.dtype)
    s6 = (sij / r) ** 6
    return torch.sum(4.0 * eij * s6 * (s6 - 1.0) * mask)


def force_matching(w, charges, sigma, epsilon, target, cutoff=None, alpha=None, K=None, order=None, skin=None,
                   pairs=None):
    """Force-matching loss L = mean((F - target)^2) over all 3N components, F = -dE/dr with E the
    TIP3P composition (per-atom LB LJ + rank-0 Ewald real space with the exclusion correction +
    PME, as :func:`energy_terms`), and its gradients with respect to coords, charges, sigma and
    epsilon by float64 double backward.  Returns (L, F, {"coords", "charges", "sigma", "epsilon"})."""
    cutoff = w["cutoff"] if cutoff is None else cutoff
    alpha = w["alpha"] if alpha is None else alpha
    K = w["K"] if K is None else K
    order = w["order"] if order is None else order
    skin = w.get("skin", 0.0) if skin is None else skin
    if pairs is None:
        pairs = pair_list(w, cutoff + skin)
    coords = torch.tensor(w["coords"], dtype=torch.float64, requires_grad=True)
    box = torch.tensor(w["box"], dtype=torch.float64)
    q = torch.tensor(charges, dtype=torch.float64, requires_grad=True)
    sig = torch.tensor(sigma, dtype=torch.float64, requires_grad=True)
    eps = torch.tensor(epsilon, dtype=torch.float64, requires_grad=True)
    pairs_t = torch.as_tensor(pairs)
    excl = torch.as_tensor(w["pairs_excl"])
    E = (_lj_per_atom(coords, box, pairs_t, sig, eps, cutoff)
         + realspace.ewald_real_energy(coords, box, pairs_t, excl, q, cutoff, alpha, COULOMB_KJMOL)
         + pme.pme_energy(coords, box, q, K, alpha, order) * COULOMB_KJMOL)
    (g,) = torch.autograd.grad(E, coords, create_graph=True)
    F = -g
    L = torch.mean((F - torch.as_tensor(target, dtype=torch.float64)) ** 2)
    grads = torch.autograd.grad(L, [coords, q, sig, eps])
    names = ["coords", "charges", "sigma", "epsilon"]
    return float(L), F.detach().numpy(), {k: v.detach().numpy() for k, v in zip(names, grads)}


def energy_gradients(w, charges, sigma, epsilon, cutoff=None, alpha=None, K=None, order=None, skin=None,
                     pairs=None):
    """E of the per-atom parameter composition of :func:`force_matching` and its first derivatives
    with respect to coords, charges, sigma and epsilon (float64 autograd)."""
    cutoff = w["cutoff"] if cutoff is None else cutoff
    alpha = w["alpha"] if alpha is None else alpha
    K = w["K"] if K is None else K
    order = w["order"] if order is None else order
    skin = w.get("skin", 0.0) if skin is None else skin
    if pairs is None:
        pairs = pair_list(w, cutoff + skin)
    coords = torch.tensor(w["coords"], dtype=torch.float64, requires_grad=True)
    box = torch.tensor(w["box"], dtype=torch.float64)
    q = torch.tensor(charges, dtype=torch.float64, requires_grad=True)
    sig = torch.tensor(sigma, dtype=torch.float64, requires_grad=True)
    eps = torch.tensor(epsilon, dtype=torch.float64, requires_grad=True)
    pairs_t = torch.as_tensor(pairs)
    excl = torch.as_tensor(w["pairs_excl"])
    E = (_lj_per_atom(coords, box, pairs_t, sig, eps, cutoff)
         + realspace.ewald_real_energy(coords, box, pairs_t, excl, q, cutoff, alpha, COULOMB_KJMOL)
         + pme.pme_energy(coords, box, q, K, alpha, order) * COULOMB_KJMOL)
    grads = torch.autograd.grad(E, [coords, q, sig, eps])
    names = ["coords", "charges", "sigma", "epsilon"]
    return float(E), {k: v.numpy() for k, v in zip(names, grads)}

