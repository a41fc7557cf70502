"""Energy + forces of one small periodic water box with every periodic image in range
(test infrastructure only).

The reference composes the TIP3P energy from a minimum-image neighbor list
(examples/tip3p/model.py:100-123, torchff/nblist.py), which needs every box edge
to be at least twice the cutoff.  Serving evaluates boxes down to 20 waters
(edge 0.84 nm < 2 x 0.9 nm), where one pair interacts through several images.
This restatement keeps the reference's pair terms and widens the sum to all
images, the periodic-lattice sum that the minimum-image form truncates:

    E_real = 1/2 sum_i sum_j sum_n [ (j, n) != (i, 0), |d_ijn| <= rc ] phi(i, j, n)
    d_ijn  = minimg(r_i - r_j) - n L,   n in Z^3
    phi    = k q_i q_j erfc(a r) / r + 4 eps_ij (s^12 - s^6)          (other pairs)
           = k q_i q_j (erfc(a r) - 1) / r                            (same molecule, n = 0)

with Lorentz-Berthelot mixing from per-atom sigma / epsilon
(csrc/nonbonded/nonbonded_atom_pairs_cuda.cu:86-89), the exclusion correction of
torchff/multipoles.py:256-257 on the nearest image of each intramolecular pair,
and PME (reciprocal + self) times the same Coulomb constant
(examples/tip3p/model.py:113-114).  Forces are -dE/dr by float64 autograd.

Pinned by tests/test_oracle_images.py:
* for edges >= 2 rc it equals composite.energy_forces (the neighbor-list
  composition restating the reference) on the same box;
* for an edge below 2 rc (and below rc), the energy of the box equals 1/27 of
  the energy of its 3 x 3 x 3 supercell computed by composite.energy_forces
  (edge 3 L >= 2 rc, grid 3 K: the PME sum is then exactly 27 times the
  cell's), and each force equals that of every copy of the atom in the
  supercell; this fixes the image range and the nearest-image exclusion;
* widening the image range by one changes nothing.
"""

import itertools
import math

import torch

from . import pme
from .water import COULOMB_KJMOL


def image_range(L, cutoff, extra=0):
    """Largest |n| per axis with an image in range: the minimum-image difference lies in
    [-L/2, L/2], so |n| <= 1/2 + rc / L."""
    return [int(math.floor(0.5 + cutoff / float(l))) + extra for l in L]


def real_space(coords, L, q, sigma, epsilon, mol, cutoff, alpha, coul=COULOMB_KJMOL, extra_images=0):
    """E_real above (float64 torch, differentiable in every tensor argument)."""
    Lt = torch.as_tensor(L, dtype=torch.float64)
    n = coords.shape[0]
    d0 = coords[:, None, :] - coords[None, :, :]
    d0 = d0 - Lt * torch.round(d0 / Lt)
    same_atom = torch.eye(n, dtype=torch.bool)
    mol = torch.as_tensor(mol)
    same_mol = mol[:, None] == mol[None, :]
    qq = coul * q[:, None] * q[None, :]
    sij = 0.5 * (sigma[:, None] + sigma[None, :])
    prod = epsilon[:, None] * epsilon[None, :]
    pos = prod > 0
    eij = torch.sqrt(torch.where(pos, prod, torch.ones_like(prod))) * pos.to(prod.dtype)
    nimg = image_range(L, cutoff, extra_images)
    e = torch.zeros((), dtype=torch.float64)
    for img in itertools.product(*(range(-m, m + 1) for m in nimg)):
        zero = all(c == 0 for c in img)
        d = d0 - torch.tensor(img, dtype=torch.float64) * Lt
        r2 = (d * d).sum(-1)
        valid = r2 <= cutoff * cutoff
        if zero:
            valid = valid & ~same_atom
        r = torch.sqrt(torch.where(valid, r2, torch.ones_like(r2)))
        damp = torch.erfc(alpha * r)
        if zero:
            damp = damp - same_mol.to(damp.dtype)
        s6 = (sij / r) ** 6
        lj = 4.0 * eij * s6 * (s6 - 1.0)
        if zero:
            lj = torch.where(same_mol, torch.zeros_like(lj), lj)
        phi = qq * damp / r + lj
        e = e + 0.5 * torch.sum(torch.where(valid, phi, torch.zeros_like(phi)))
    return e


def molecule_ids(natoms, atoms_per_molecule=3):
    return torch.arange(natoms, dtype=torch.int64) // atoms_per_molecule


def energy_forces(w, cutoff, alpha, K, order, coul=COULOMB_KJMOL, extra_images=0):
    """Total energy, forces (N, 3) and terms {"real", "recip_self"} of a water-box dict
    (torchff.workloads.water_box layout: per-atom charges / sigma / epsilon, 3 atoms per
    molecule, cubic or orthorhombic box)."""
    coords = torch.tensor(w["coords"], dtype=torch.float64, requires_grad=True)
    box = torch.tensor(w["box"], dtype=torch.float64)
    L = [float(box[a, a]) for a in range(3)]
    q = torch.tensor(w["charges"], dtype=torch.float64)
    sig = torch.tensor(w["sigma"], dtype=torch.float64)
    eps = torch.tensor(w["epsilon"], dtype=torch.float64)
    mol = molecule_ids(len(q))
    e_real = real_space(coords, L, q, sig, eps, mol, cutoff, alpha, coul, extra_images)
    e_pme = pme.pme_energy(coords, box, q, K, alpha, order) * coul
    E = e_real + e_pme
    (g,) = torch.autograd.grad(E, coords)
    return float(E), -g.numpy(), {"real": float(e_real), "recip_self": float(e_pme)}
