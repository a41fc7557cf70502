"""Molecular-dynamics consumers of the batched serving path.

``BatchedSimModel`` follows the torch-sim ``ModelInterface`` protocol that the reference's TIP3P
example implements for one system (examples/tip3p/md_torchsim.py:98-143): ``forward(state)`` takes
a state with ``positions`` (N, 3) in Angstrom, ``cell`` (n_systems, 3, 3) with the lattice vectors
in columns (ASE / torch-sim convention) and the atom-to-system map ``system_idx`` (``batch`` in
older torch-sim), and returns ``{"energy": (n_systems,) eV, "forces": (N, 3) eV/Angstrom}``.
Where the reference raises for ``n_systems != 1``, every system of the state here goes into one
batch of the serving engine (real-space erfc-Coulomb + Lennard-Jones over all periodic images,
intramolecular exclusions, smooth PME), so a multi-replica or multi-box simulation costs one
launch sequence per step.  torch-sim itself is not imported: any object with those attributes
works.  Boxes must be orthorhombic (the serving engine's geometry).
"""

from typing import List, Optional, Sequence

import numpy as np
import torch
import torch.nn as nn

from .serving import Request, ServingEngine, pme_grid_size

KJMOL_PER_EV = 96.48533212331014  # kJ/mol per eV (examples/tip3p/md_torchsim.py:49)
NM_PER_ANGSTROM = 0.1


def _system_index(state, n_atoms: int) -> np.ndarray:
    idx = getattr(state, "system_idx", None)
    if idx is None:
        idx = getattr(state, "batch", None)
    if idx is None:
        return np.zeros(n_atoms, dtype=np.int64)
    return np.asarray(torch.as_tensor(idx).detach().cpu().numpy(), dtype=np.int64).reshape(-1)


def state_requests(state, charges: Sequence[float], sigma: Sequence[float], epsilon: Sequence[float],
                   molecule: Sequence[int], spacing: float = 0.1, tol: float = 1e-9) -> List[Request]:
    """One serving request per system of ``state`` (positions in Angstrom, cells with lattice
    vectors in columns), with the per-atom parameters of the whole state (charges in e, sigma in
    nm, epsilon in kJ/mol, molecule ids for the intramolecular exclusions) and the PME grid from
    the spacing rule.  The request of system s holds its atoms in state order."""
    pos = torch.as_tensor(state.positions).detach().to(torch.float64).cpu().numpy().reshape(-1, 3)
    cell = torch.as_tensor(state.cell).detach().to(torch.float64).cpu().numpy().reshape(-1, 3, 3)
    n = pos.shape[0]
    sys_idx = _system_index(state, n)
    if sys_idx.shape[0] != n:
        raise ValueError(f"system index has {sys_idx.shape[0]} entries for {n} atoms")
    q, s, e, mol = (np.asarray(a).reshape(-1) for a in (charges, sigma, epsilon, molecule))
    for name, a in (("charges", q), ("sigma", s), ("epsilon", e), ("molecule", mol)):
        if a.shape[0] != n:
            raise ValueError(f"{name} has {a.shape[0]} entries for {n} atoms")
    reqs = []
    for k in range(cell.shape[0]):
        atoms = np.nonzero(sys_idx == k)[0]
        if atoms.size == 0:
            raise ValueError(f"system {k} has no atoms")
        box = cell[k].T * NM_PER_ANGSTROM  # rows = lattice vectors, nm
        off = box - np.diag(np.diag(box))
        if np.abs(off).max() > tol * max(1.0, np.abs(box).max()):
            raise ValueError(f"system {k}: the serving engine takes orthorhombic boxes only")
        L = np.diag(box)
        K = [pme_grid_size(float(x), spacing) for x in L]
        reqs.append(Request(k, pos[atoms] * NM_PER_ANGSTROM, q[atoms], s[atoms], e[atoms], mol[atoms], L, K))
    return reqs


class BatchedSimModel(nn.Module):
    """torch-sim style model over the serving engine (see the module docstring).

    ``charges``, ``sigma``, ``epsilon``, ``molecule``: per atom of the states this model will see
    (the same atom order); ``engine``: a ``ServingEngine`` whose capacities fit the batch."""

    def __init__(self, engine: ServingEngine, charges, sigma, epsilon, molecule, spacing: float = 0.1,
                 device: Optional[torch.device] = None, dtype: torch.dtype = torch.float64):
        super().__init__()
        self.engine = engine
        self.charges = np.asarray(charges, dtype=np.float64).reshape(-1)
        self.sigma = np.asarray(sigma, dtype=np.float64).reshape(-1)
        self.epsilon = np.asarray(epsilon, dtype=np.float64).reshape(-1)
        self.molecule = np.asarray(molecule, dtype=np.int64).reshape(-1)
        self.spacing = float(spacing)
        self._device = torch.device(device) if device is not None else engine.device
        self._dtype = dtype

    # the torch-sim ModelInterface properties
    @property
    def device(self) -> torch.device:
        return self._device

    @property
    def dtype(self) -> torch.dtype:
        return self._dtype

    @property
    def compute_stress(self) -> bool:
        return False

    @property
    def compute_forces(self) -> bool:
        return True

    def forward(self, state, **kwargs):
        reqs = state_requests(state, self.charges, self.sigma, self.epsilon, self.molecule, self.spacing)
        for r in reqs:
            if not self.engine.fits_alone(r):
                raise RuntimeError(f"system {r.rid} ({r.natoms} atoms, grid {r.grid}) exceeds the engine's "
                                   "capacities")
        done = sorted(self.engine.evaluate(reqs), key=lambda r: r.rid)
        n = sum(r.natoms for r in done)
        sys_idx = _system_index(state, n)
        forces = np.zeros((n, 3), dtype=np.float64)
        for r in done:
            forces[np.nonzero(sys_idx == r.rid)[0]] = r.forces
        energy = np.array([r.energies[4] for r in done], dtype=np.float64) / KJMOL_PER_EV
        forces *= NM_PER_ANGSTROM / KJMOL_PER_EV
        return {"energy": torch.as_tensor(energy, dtype=self._dtype, device=self._device),
                "forces": torch.as_tensor(forces, dtype=self._dtype, device=self._device)}
