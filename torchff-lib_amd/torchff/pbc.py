"""Minimum-image utility module (torchff/pbc.py:6-97 of the reference).

Pure tensor arithmetic on whatever device the inputs live on, exactly as the
reference's module; the HIP kernels apply the same convention internally
(s = d . B^-1, s -= round(s), d = s . B).
"""

from typing import Optional

import torch
import torch.nn as nn


class PBC(nn.Module):
    def __init__(self, box: Optional[torch.Tensor] = None, box_inv: Optional[torch.Tensor] = None):
        super().__init__()
        if box is not None and box_inv is not None:
            self.register_buffer("box", box)
            self.register_buffer("box_inv", box_inv)
        else:
            self.box = None
            self.box_inv = None

    def forward(self, dr_vecs, box=None, box_inv=None):
        box = box if box is not None else getattr(self, "box", None)
        box_inv = box_inv if box_inv is not None else getattr(self, "box_inv", None)
        return applyPBC(dr_vecs, box, box_inv)


def applyPBC(dr_vecs, box=None, box_inv=None):
    if box is not None and box_inv is None:
        box_inv, _ = torch.linalg.inv_ex(box)
    if box is None or box_inv is None:
        return dr_vecs
    ds = torch.matmul(dr_vecs, box_inv)
    ds = ds - torch.floor(ds + 0.5)
    return torch.matmul(ds, box)
