"""Rank-0 (point-charge) ``MultipolarInteraction`` (torchff/multipoles.py:215-355 of the reference).

Real-space Ewald Coulomb prefactor * sum q_i q_j erfc(alpha r)/r over the pair
list (bare 1/r when ewald_alpha < 0) plus the (erfc(alpha r) - 1)/r
correction over ``pairs_excl``.  ``return_fields=True`` returns (energy, per-atom
potential dE/dq, field) with the field zero for point charges, as the reference's
CUDA op (csrc/multipoles/multipoles_field.cu), differentiable through energy and
potential.  Dipoles and quadrupoles (ranks 1-2) in real space are not part of this
build and raise.
"""

import torch
import torch.nn as nn

from . import _ops  # noqa: F401
from ._common import require_custom


class MultipolarInteraction(nn.Module):
    def __init__(self, rank: int, cutoff: float, ewald_alpha: float = -1.0, prefactor: float = 1.0,
                 return_fields: bool = False, use_customized_ops: bool = True, cuda_graph_compat: bool = True):
        super().__init__()
        if rank != 0:
            raise NotImplementedError("MultipolarInteraction: only rank 0 is provided by the gfx950 build")
        require_custom(use_customized_ops, "MultipolarInteraction")
        self.rank = rank
        self.cutoff = cutoff
        self.ewald_alpha = ewald_alpha
        self.prefactor = prefactor
        self.return_fields = return_fields
        self.use_customized_ops = use_customized_ops
        self.cuda_graph_compat = cuda_graph_compat

    @torch._dynamo.disable
    def forward(self, coords, box, pairs, q, p=None, t=None, pairs_excl=None):
        if self.return_fields:
            return torch.ops.torchff.compute_multipolar_energy_and_fields_from_atom_pairs(
                coords, box, pairs, pairs_excl, q, p, t, self.cutoff, self.ewald_alpha, self.prefactor)
        return torch.ops.torchff.compute_multipolar_energy_from_atom_pairs(
            coords, box, pairs, pairs_excl, q, p, t, self.cutoff, self.ewald_alpha, self.prefactor)
