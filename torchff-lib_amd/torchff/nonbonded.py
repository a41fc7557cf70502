"""Fused nonbonded pair energy (drop-in for torchff/nonbonded.py of the reference).

``compute_nonbonded_energy_from_atom_pairs`` / ``..._forces_...`` keep the
reference signatures (torchff/nonbonded.py:11-96): shifted Coulomb
k q_i q_j (1/r - [1/rc]) + Lorentz-Berthelot LJ over a pair list.
``Nonbonded.forward`` additionally accepts ``ewald_alpha`` and ``pairs_excl``:
with ``ewald_alpha`` given it evaluates the fused real-space Ewald term
k q_i q_j erfc(alpha r)/r + LJ plus the k q_i q_j (erfc(alpha r) - 1)/r
correction over ``pairs_excl`` (the additive extension; the reference splits
these between MultipolarInteraction rank 0 and Vdw).
"""

import torch
import torch.nn as nn

from . import _ops  # noqa: F401
from ._common import require_custom


@torch._dynamo.disable
def compute_nonbonded_energy_from_atom_pairs(coords, pairs, box, sigma, epsilon, charges, coul_constant, cutoff,
                                             do_shift=True):
    if pairs.dtype != torch.int64:
        pairs = pairs.to(torch.int64)
    return torch.ops.torchff.compute_nonbonded_energy_from_atom_pairs(coords, pairs, box, sigma, epsilon, charges,
                                                                      coul_constant, cutoff, do_shift)


def compute_nonbonded_forces_from_atom_pairs(coords, pairs, box, sigma, epsilon, charges, coul_constant, cutoff,
                                             forces):
    if pairs.dtype != torch.int64:
        pairs = pairs.to(torch.int64)
    return torch.ops.torchff.compute_nonbonded_forces_from_atom_pairs(coords, pairs, box, sigma, epsilon, charges,
                                                                      coul_constant, cutoff, forces)


@torch._dynamo.disable
def compute_nonbonded_ewald_energy_from_atom_pairs(coords, pairs, box, sigma, epsilon, charges, coul_constant,
                                                   cutoff, ewald_alpha, pairs_excl=None):
    if pairs.dtype != torch.int64:
        pairs = pairs.to(torch.int64)
    return torch.ops.torchff.compute_nonbonded_ewald_energy_from_atom_pairs(
        coords, pairs, box, sigma, epsilon, charges, coul_constant, cutoff, ewald_alpha, pairs_excl)


def compute_nonbonded_energy_from_cluster_pairs(coords, box, sigma, epsilon, charges, coul_constant, cutoff,
                                                sorted_atom_indices, cluster_exclusions, bitmask_exclusions,
                                                interacting_clusters, interacting_atoms, do_shift=True):
    """The atom-pair nonbonded energy evaluated on a cluster-pair list (the schema commented at
    csrc/nonbonded/nonbonded_interface.cpp:7), differentiable in coords, sigma, epsilon, charges."""
    return torch.ops.torchff.compute_nonbonded_energy_from_cluster_pairs(
        coords, box, sigma, epsilon, charges, coul_constant, cutoff, sorted_atom_indices, cluster_exclusions,
        bitmask_exclusions, interacting_clusters, interacting_atoms, do_shift)


def compute_nonbonded_forces_from_cluster_pairs(coords, box, sigma, epsilon, charges, coul_constant, cutoff,
                                                sorted_atom_indices, cluster_exclusions, bitmask_exclusions,
                                                interacting_clusters, interacting_atoms, forces):
    """Adds the nonbonded forces on a cluster-pair list into ``forces`` (nonbonded_interface.cpp:8)."""
    torch.ops.torchff.compute_nonbonded_forces_from_cluster_pairs(
        coords, box, sigma, epsilon, charges, coul_constant, cutoff, sorted_atom_indices, cluster_exclusions,
        bitmask_exclusions, interacting_clusters, interacting_atoms, forces)


class Nonbonded(nn.Module):
    """Fused fixed-charge nonbonded interaction (torchff/nonbonded.py:99-141)."""

    def __init__(self, use_customized_ops: bool = True):
        super().__init__()
        require_custom(use_customized_ops, "Nonbonded")
        self.use_customized_ops = use_customized_ops

    def forward(self, coords, pairs, box, sigma, epsilon, charges, coul_constant, cutoff, do_shift=True,
                ewald_alpha=None, pairs_excl=None):
        if ewald_alpha is not None:
            return compute_nonbonded_ewald_energy_from_atom_pairs(coords, pairs, box, sigma, epsilon, charges,
                                                                  coul_constant, cutoff, ewald_alpha, pairs_excl)
        return compute_nonbonded_energy_from_atom_pairs(coords, pairs, box, sigma, epsilon, charges, coul_constant,
                                                        cutoff, do_shift)
