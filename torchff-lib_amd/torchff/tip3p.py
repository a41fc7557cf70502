"""The nonbonded energy of the reference's TIP3P model as one TorchScript-compatible module.

``Tip3pNonbonded`` composes what ``Tip3pTorchFF.energy_components`` sums for the nonbonded
terms (examples/tip3p/model.py:100-123): the neighbour list (nsquared as the reference's model,
or cell list), Lennard-Jones from the (T, T) Lorentz-Berthelot type tables
(examples/tip3p/md.py:202-214), rank-0 Ewald real space with the (erfc - 1)/r correction on the
intramolecular pairs (md.py:152-158), and PME scaled by the Coulomb constant.  It calls only
``torch.ops.torchff.*`` operators that ``libtorchff_ops.so`` registers in C++, so
``torch.jit.trace`` (the reference's route to OpenMM's TorchForce, md.py:451-520) or
``torch.jit.script`` gives a module that runs, with its forces by autograd, in a process where
only libtorch and ``libtorchff_ops.so`` are loaded.  Bonded terms are not part of it.
"""

import torch
import torch.nn as nn

from . import _ops  # noqa: F401  (loads the operator library)
from . import bsplines
from .nblist import NeighborList
from .workloads import COULOMB_KJMOL


class Tip3pNonbonded(nn.Module):
    def __init__(self, natoms: int, charges, atom_types, sigma_table, epsilon_table, exclusions,
                 coulomb_excl_pairs, cutoff: float, ewald_alpha: float, max_hkl: int, order: int = 6,
                 algorithm: str = "nsquared", max_npairs: int = -1, prefactor: float = COULOMB_KJMOL):
        super().__init__()
        if algorithm not in ("nsquared", "cell_list"):
            raise ValueError("algorithm must be 'nsquared' or 'cell_list'")
        self.natoms = int(natoms)
        self.cutoff = float(cutoff)
        self.alpha = float(ewald_alpha)
        self.K = int(max_hkl)
        self.order = int(order)
        self.prefactor = float(prefactor)
        self.cell_list = algorithm == "cell_list"
        # the reference model's capacity (model.py:97) unless given
        self.max_npairs = int(max_npairs) if max_npairs > 0 else \
            int(natoms * 4 / 3 * 3.1416 * cutoff ** 3 * 100 / 2 * 1.2)
        rp, col = NeighborList.convert_pairs_coo_to_csr(torch.as_tensor(exclusions), self.natoms)
        self.register_buffer("excl_row_ptr", rp)
        self.register_buffer("excl_col_indices", col)
        self.register_buffer("charges", torch.as_tensor(charges, dtype=torch.float64).clone())
        self.register_buffer("atom_types", torch.as_tensor(atom_types, dtype=torch.int64).clone())
        self.register_buffer("sigma_table", torch.as_tensor(sigma_table, dtype=torch.float64).clone())
        self.register_buffer("epsilon_table", torch.as_tensor(epsilon_table, dtype=torch.float64).clone())
        self.register_buffer("coulomb_excl_pairs", torch.as_tensor(coulomb_excl_pairs, dtype=torch.int64).clone())
        self.register_buffer("xmoduli", bsplines.compute_bspline_moduli_1d(self.K, order=self.order))
        self.register_buffer("ymoduli", bsplines.compute_bspline_moduli_1d(self.K, order=self.order))
        self.register_buffer("zmoduli", bsplines.compute_bspline_moduli_z(self.K, order=self.order))

    def energy_components(self, coords: torch.Tensor, box: torch.Tensor):
        """(Lennard-Jones, real-space Coulomb incl. the exclusion correction, reciprocal Coulomb)."""
        if self.cell_list:
            pairs, npairs = torch.ops.torchff.build_neighbor_list_cell_list(
                coords, box, self.cutoff, self.max_npairs, self.excl_row_ptr, self.excl_col_indices, False)
        else:
            pairs, npairs = torch.ops.torchff.build_neighbor_list_nsquared(
                coords, box, self.cutoff, self.max_npairs, self.excl_row_ptr, self.excl_col_indices, False)
        q = self.charges.to(coords.dtype)
        e_lj = torch.ops.torchff.compute_lennard_jones_energy(
            coords, pairs, box, self.sigma_table.to(coords.dtype), self.epsilon_table.to(coords.dtype), self.cutoff,
            self.atom_types)
        e_coul = torch.ops.torchff.compute_multipolar_energy_from_atom_pairs(
            coords, box, pairs, self.coulomb_excl_pairs, q, None, None, self.cutoff, self.alpha, self.prefactor)
        e_recip = torch.ops.torchff.pme_long_range(
            coords, box, q, None, None, self.K, self.alpha, self.xmoduli.to(coords.dtype),
            self.ymoduli.to(coords.dtype), self.zmoduli.to(coords.dtype), self.order) * self.prefactor
        return e_lj, e_coul, e_recip

    def forward(self, coords: torch.Tensor, box: torch.Tensor) -> torch.Tensor:
        e_lj, e_coul, e_recip = self.energy_components(coords, box)
        return e_lj + e_coul + e_recip


def for_workload(w, algorithm="nsquared", max_npairs=-1):
    """A Tip3pNonbonded for a water workload dict (torchff.workloads)."""
    return Tip3pNonbonded(w["natoms"], w["charges"], w["atom_types"], w["sigma_table"], w["epsilon_table"],
                          w["exclusions"], w["pairs_excl"], w["cutoff"], w["alpha"], w["K"], order=w["order"],
                          algorithm=algorithm, max_npairs=max_npairs)
