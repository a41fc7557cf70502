"""Lennard-Jones 12-6 pair energy (drop-in for torchff/vdw.py of the reference, LJ form).

``Vdw('LennardJones')`` with per-pair sigma/epsilon or a (T, T) type table
indexed by ``atom_types`` (torchff/vdw.py:142-261); cutoff r <= rc, no shift.
The AMOEBA buffered 14-7 form is not part of this build.
"""

from typing import Optional

import torch
import torch.nn as nn

from . import _ops  # noqa: F401
from ._common import require_custom


@torch._dynamo.disable
def compute_lennard_jones_energy(coords, pairs, box, sigma, epsilon, cutoff, atom_types=None):
    return torch.ops.torchff.compute_lennard_jones_energy(coords, pairs, box, sigma, epsilon, cutoff, atom_types)


class Vdw(nn.Module):
    def __init__(self, function: str = "LennardJones", cutoff: Optional[float] = None,
                 use_customized_ops: bool = True, use_type_pairs: bool = False, sum_output: bool = True,
                 cuda_graph_compat: bool = True):
        super().__init__()
        if function != "LennardJones":
            raise NotImplementedError(f"Vdw: function {function!r} is not provided by the gfx950 build")
        require_custom(use_customized_ops, "Vdw")
        if not sum_output:
            raise ValueError("sum_output must be True when use_customized_ops is True "
                             "(the kernels only compute total energy, not per-pair terms).")
        self.function = function
        self.cutoff = cutoff
        self.use_customized_ops = use_customized_ops
        self.use_type_pairs = use_type_pairs
        self.sum_output = sum_output
        self.cuda_graph_compat = cuda_graph_compat

    def forward(self, coords, pairs, box, sigma, epsilon, cutoff, atom_types=None):
        return compute_lennard_jones_energy(coords, pairs, box, sigma, epsilon, cutoff,
                                            atom_types if self.use_type_pairs else None)
