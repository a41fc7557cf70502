"""Direct (FFT-free) Ewald reciprocal sum (drop-in for torchff/ewald.py of the reference).

Same functions and module as the reference (torchff/ewald.py:9-59): ``ewald_long_range``,
``ewald_long_range_with_fields`` and ``Ewald(alpha, kmax, rank, use_customized_ops=True,
return_fields=True)``, backed by ``tff_ewald`` / ``tff_ewald_multipole`` (gfx950), for point
charges, dipoles p (N, 3) and quadrupoles t ((N, 3, 3), (N, 9) or packed (N, 6)).  The
pure-PyTorch path (``use_customized_ops=False``) is not shipped: the float64 restatement lives in
the oracle used by the tests.
"""

import torch
import torch.nn as nn

from . import _ops  # noqa: F401  (registers torch.ops.torchff.ewald_*)
from ._common import require_custom


@torch._dynamo.disable
def ewald_long_range(coords: torch.Tensor, box: torch.Tensor, q: torch.Tensor, p, t, kmax: int, alpha: float):
    return torch.ops.torchff.ewald_long_range(coords, box, q, p, t, kmax, alpha)


@torch._dynamo.disable
def ewald_long_range_with_fields(coords: torch.Tensor, box: torch.Tensor, q: torch.Tensor, p, t, kmax: int,
                                 alpha: float):
    return torch.ops.torchff.ewald_long_range_all(coords, box, q, p, t, kmax, alpha)


class Ewald(nn.Module):
    def __init__(self, alpha: float, kmax: int, rank: int, use_customized_ops: bool = True,
                 return_fields: bool = True):
        super().__init__()
        require_custom(use_customized_ops, "Ewald")
        if rank not in (0, 1, 2):
            raise ValueError(f"Ewald: rank must be 0, 1 or 2 (got {rank})")
        self.alpha = float(alpha)
        self.kmax = int(kmax)
        self.rank = rank
        self.return_fields = return_fields
        self.use_customized_ops = use_customized_ops

    def forward(self, coords, box, q, p=None, t=None):
        if self.return_fields:
            return ewald_long_range_with_fields(coords, box, q, p, t, self.kmax, self.alpha)
        return ewald_long_range(coords, box, q, p, t, self.kmax, self.alpha)
