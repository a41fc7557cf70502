"""Smooth PME for point charges, dipoles and quadrupoles (drop-in for torchff/pme.py).

``PME(alpha, max_hkl, rank)`` keeps the reference's constructor and forward
(torchff/pme.py:13-127) and adds ``order`` (B-spline order, default 6 as in
the reference; 5 for the 96k- and 999k-atom workloads).  ``forward(coords,
box, q, p=None, t=None)`` returns the reciprocal-space plus self energy in
Gaussian units (multiply by the Coulomb constant), differentiable in coords, q,
p and t; with ``return_fields=True`` it returns (energy, per-atom potential,
field), both including the self terms, the field being zero for point charges
as in the reference (``pme_long_range_all``).  Quadrupoles t may be (N, 3, 3),
(N, 9) or the packed (N, 6) [xx, xy, xz, yy, yz, zz].
"""

from typing import Optional

import torch
import torch.nn as nn

from . import _ops  # noqa: F401
from . import bsplines
from ._common import require_custom


class PME(nn.Module):
    def __init__(self, alpha: float, max_hkl: int, rank: int, use_customized_ops: bool = True,
                 return_fields: bool = False, order: int = 6):
        super().__init__()
        if rank not in (0, 1, 2):
            raise ValueError(f"PME: rank must be 0, 1 or 2 (got {rank})")
        require_custom(use_customized_ops, "PME")
        self.alpha = float(alpha)
        self.max_hkl = int(max_hkl)
        self.rank = int(rank)
        self.order = int(order)
        self.use_customized_ops = use_customized_ops
        self.return_fields = return_fields
        K = self.max_hkl
        self.register_buffer("xmoduli", bsplines.compute_bspline_moduli_1d(K, order=self.order))
        self.register_buffer("ymoduli", bsplines.compute_bspline_moduli_1d(K, order=self.order))
        self.register_buffer("zmoduli", bsplines.compute_bspline_moduli_z(K, order=self.order))

    @torch._dynamo.disable
    def forward(self, coords: torch.Tensor, box: torch.Tensor, q: torch.Tensor, p: Optional[torch.Tensor] = None,
                t: Optional[torch.Tensor] = None):
        if self.return_fields:
            # (energy, potential incl. the self term, field: zero for point charges)
            return torch.ops.torchff.pme_long_range_all(coords, box, q, p, t, self.max_hkl, self.alpha,
                                                        self.xmoduli, self.ymoduli, self.zmoduli, self.order)
        return torch.ops.torchff.pme_long_range(coords, box, q, p, t, self.max_hkl, self.alpha, self.xmoduli,
                                                self.ymoduli, self.zmoduli, self.order)
