"""Neighbor list (drop-in for torchff/nblist.py of the reference, ``use_customized_ops=True``).

``NeighborList`` keeps the reference's constructor and forward signature
(torchff/nblist.py:10-178).  Both algorithms return the same pair set: the
float64 predicate |minimg(r_a - r_b)|^2 <= cutoff^2, minus CSR exclusions;
rows are (a, b) with a > b, padded with -1 up to the capacity.
"""

from typing import Optional, Tuple

import torch
import torch.nn as nn

from . import _ops  # noqa: F401
from ._common import require_custom


class NeighborList(nn.Module):
    def __init__(self, natoms: int, exclusions: Optional[torch.Tensor] = None, include_self: bool = False,
                 use_customized_ops: bool = True, algorithm: str = "nsquared"):
        super().__init__()
        if algorithm not in ("nsquared", "cell_list"):
            raise ValueError(f"algorithm must be 'nsquared' or 'cell_list', got '{algorithm}'")
        require_custom(use_customized_ops, "NeighborList")
        self.natoms = natoms
        self.include_self = include_self
        self.use_customized_ops = use_customized_ops
        self.algorithm = algorithm
        if exclusions is not None:
            self.register_buffer("exclusions_coo", exclusions)
            row_ptr, col = self.convert_pairs_coo_to_csr(exclusions, natoms)
            self.register_buffer("excl_row_ptr", row_ptr)
            self.register_buffer("excl_col_indices", col)
        else:
            self.exclusions_coo = None
            self.excl_row_ptr = None
            self.excl_col_indices = None

    @classmethod
    def convert_pairs_coo_to_csr(cls, pairs_coo: torch.Tensor, num_atoms: Optional[int] = None
                                 ) -> Tuple[torch.Tensor, torch.Tensor]:
        """Symmetrised, deduplicated CSR of a COO pair list (torchff/nblist.py:50-123)."""
        allp = torch.cat([pairs_coo, pairs_coo.flip(1)], dim=0).to(torch.long)
        if num_atoms is None:
            num_atoms = int(allp.max().item()) + 1
        keys = torch.unique(allp[:, 0] * num_atoms + allp[:, 1])
        rows = keys // num_atoms
        col = keys % num_atoms
        counts = torch.zeros(num_atoms, dtype=torch.long, device=pairs_coo.device)
        counts.scatter_add_(0, rows, torch.ones_like(rows))
        row_ptr = torch.zeros(num_atoms + 1, dtype=torch.long, device=pairs_coo.device)
        row_ptr[1:] = torch.cumsum(counts, dim=0)
        return row_ptr, col

    def forward(self, coords: torch.Tensor, box: torch.Tensor, cutoff: float, max_npairs: int = -1,
                padding: bool = False):
        if self.algorithm == "cell_list":
            if box is None:
                raise ValueError("Cell-list algorithm requires a periodic box")
            pairs, npairs = torch.ops.torchff.build_neighbor_list_cell_list(
                coords, box, cutoff, max_npairs, self.excl_row_ptr, self.excl_col_indices, self.include_self)
        else:
            pairs, npairs = torch.ops.torchff.build_neighbor_list_nsquared(
                coords, box, cutoff, max_npairs, self.excl_row_ptr, self.excl_col_indices, self.include_self)
        if padding:
            return pairs, npairs
        found = int(npairs.item())
        if found > pairs.shape[0]:
            raise RuntimeError(f"Too many neighbor pairs found. Maximum is {pairs.shape[0]} but found {found}")
        return pairs[:found]


def build_neighbor_list_nsquared(coords, box, cutoff, max_npairs=-1, padding=False, excl_row_ptr=None,
                                 excl_col_indices=None, include_self=False, out=None):
    """Functional O(N^2) list (torchff/nblist.py:183-250): returns (pairs, npairs)."""
    return _functional("nsquared", coords, box, cutoff, max_npairs, padding, excl_row_ptr, excl_col_indices,
                       include_self, out)


def build_neighbor_list_cell_list(coords, box, cutoff, max_npairs=-1, padding=False, excl_row_ptr=None,
                                  excl_col_indices=None, include_self=False, out=None):
    """Functional cell-list variant with the same arguments and result as the O(N^2) one."""
    return _functional("cell_list", coords, box, cutoff, max_npairs, padding, excl_row_ptr, excl_col_indices,
                       include_self, out)


def _functional(algo, coords, box, cutoff, max_npairs, padding, excl_row_ptr, excl_col_indices, include_self, out):
    ops = torch.ops.torchff
    if out is not None:
        if algo == "nsquared":
            npairs = ops.build_neighbor_list_nsquared_out(coords, box, cutoff, out, excl_row_ptr, excl_col_indices,
                                                          include_self)
        else:
            _, npairs = ops.build_neighbor_list_cell_list_out(coords, box, cutoff, out, excl_row_ptr,
                                                              excl_col_indices, include_self)
        pairs = out
    elif algo == "nsquared":
        pairs, npairs = ops.build_neighbor_list_nsquared(coords, box, cutoff, max_npairs, excl_row_ptr,
                                                         excl_col_indices, include_self)
    else:
        pairs, npairs = ops.build_neighbor_list_cell_list(coords, box, cutoff, max_npairs, excl_row_ptr,
                                                          excl_col_indices, include_self)
    if not padding:
        found = int(npairs.item())
        if found > pairs.shape[0]:
            raise RuntimeError(f"Too many neighbor pairs found. Maximum is {pairs.shape[0]} but found {found}")
        pairs = pairs[:found]
    return pairs, npairs


def build_cluster_pairs(coords: torch.Tensor, box: torch.Tensor, cutoff: float,
                        exclusions: Optional[torch.Tensor] = None, cell_size: float = 0.4,
                        max_num_interacting_clusters: int = -1):
    """Cluster-pair list (torchff/nblist.py:260-286): (sorted_atom_indices, cluster_exclusions,
    bitmask_exclusions, interacting_clusters, interacting_atoms); exclusions as a (2, E) COO list."""
    return torch.ops.torchff.build_cluster_pairs(coords, box, cutoff, exclusions, cell_size,
                                                 max_num_interacting_clusters)


def decode_cluster_pairs(coords: torch.Tensor, box: torch.Tensor, sorted_atom_indices, cluster_exclusions,
                         bitmask_exclusions, interacting_clusters, interacting_atoms, cutoff: float,
                         max_npairs: int = -1, padding: bool = False):
    """Atom pairs of a cluster-pair list within the cutoff (torchff/nblist.py:289-316): (pairs, npairs)."""
    return torch.ops.torchff.decode_cluster_pairs(coords, box, sorted_atom_indices, cluster_exclusions,
                                                  bitmask_exclusions, interacting_clusters, interacting_atoms,
                                                  cutoff, max_npairs, padding)


def build_neighbor_list_cluster_pairs(coords: torch.Tensor, box: torch.Tensor, cutoff: float,
                                      exclusions: Optional[torch.Tensor], cell_size: float = 0.45,
                                      max_num_interacting_clusters: int = -1, max_npairs: int = -1,
                                      padding: bool = False):
    """Pair list through the cluster pairs (torchff/nblist.py:319-344, without its debug prints)."""
    nbl = build_cluster_pairs(coords, box, cutoff, exclusions, cell_size, max_num_interacting_clusters)
    return decode_cluster_pairs(coords, box, *nbl, cutoff, max_npairs, padding)
