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
    (the same atom order); ``engine``: a ``ServingEngine`` whose capacities fit the batch.

    The batch stays resident on the device in the engine's first slot: the host builds and
    uploads its table only when the batch composition changes (another atom count, system
    index tensor or cell tensor, or the slot used for other work in between).  Each call then
    writes the state's positions into the slot's device atom records, replays the captured
    evaluation on the slot's stream and gathers energies and forces on the device: no
    device-to-host copy of positions and no host synchronisation per step.  A cell or system
    index changed in place without a version bump (through ``.data`` or a view) is not seen;
    give the state a new tensor (or call :meth:`invalidate`).  Calls on one engine from several
    models take turns on that slot (each change of owner re-uploads the table)."""

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
        self._res = None

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

    def invalidate(self):
        """Forget the resident batch: the next call rebuilds and uploads the table."""
        self._res = None

    @staticmethod
    def _tensor_key(t):
        return None if t is None else (t.data_ptr(), t._version, tuple(t.shape), t.device)

    def _key(self, state):
        idx = getattr(state, "system_idx", None)
        if idx is None:
            idx = getattr(state, "batch", None)
        pos = state.positions
        return (id(self), tuple(pos.shape), pos.device, self._tensor_key(torch.as_tensor(state.cell)),
                self._tensor_key(idx if isinstance(idx, torch.Tensor) else None),
                None if isinstance(idx, torch.Tensor) or idx is None else tuple(np.asarray(idx).reshape(-1)))

    def _setup(self, state, key):
        """Host side, once per batch composition: one request per system, the table built in the
        slot and uploaded, and the maps between state order and the slot's atom records
        (systems in order, each system's atoms in state order)."""
        slot = self.engine.slots[0]
        if slot.requests:
            raise RuntimeError("BatchedSimModel needs an idle engine slot")
        reqs = state_requests(state, self.charges, self.sigma, self.epsilon, self.molecule, self.spacing)
        slot.reset()
        for r in reqs:
            if not slot.try_add(r):
                slot.reset()
                raise RuntimeError(f"system {r.rid} ({r.natoms} atoms, grid {r.grid}) does not fit the engine's "
                                   "capacities next to the others")
        n = sum(r.natoms for r in reqs)
        with torch.cuda.stream(slot.stream):
            slot.upload()
        slot.stream.synchronize()  # the pinned host table is reused by the next upload
        slot.reset()
        sys_idx = _system_index(state, n)
        order = np.concatenate([np.nonzero(sys_idx == r.rid)[0] for r in reqs]).astype(np.int64)
        dev = slot.ws.device
        perm = torch.as_tensor(order, device=dev)
        inv = torch.empty_like(perm)
        inv[perm] = torch.arange(n, device=dev)
        slot.resident_key = key
        # a device divisor: torch divides by a host scalar through its reciprocal, the request path
        # (numpy) by true division
        ev = torch.tensor(KJMOL_PER_EV, dtype=torch.float64, device=dev)
        self._res = dict(key=key, perm=perm, inv=inv, n=n, nsys=len(reqs), slot=slot, kjmol_per_ev=ev)

    def forward(self, state, **kwargs):
        key = self._key(state)
        res = self._res
        if res is None or res["key"] != key or res["slot"].resident_key != key:
            self._setup(state, key)
            res = self._res
        slot, n, nsys = res["slot"], res["n"], res["nsys"]
        cur = torch.cuda.current_stream(slot.ws.device)
        pos = torch.as_tensor(state.positions).detach().to(device=slot.ws.device, dtype=torch.float64)
        slot.device_atoms(n)[:, 0:3] = pos.index_select(0, res["perm"]) * NM_PER_ANGSTROM
        slot.stream.wait_stream(cur)
        with torch.cuda.stream(slot.stream):
            slot.run()
        cur.wait_stream(slot.stream)
        d = slot.d_out
        f0 = 5 * self.engine.caps["max_systems"]
        energy = d[4: 5 * nsys: 5] / res["kjmol_per_ev"]
        forces = d[f0: f0 + 3 * n].view(n, 3).index_select(0, res["inv"]) * (NM_PER_ANGSTROM / KJMOL_PER_EV)
        return {"energy": energy.to(device=self._device, dtype=self._dtype),
                "forces": forces.to(device=self._device, dtype=self._dtype)}

    def forward_via_requests(self, state):
        """The same evaluation through the request path (host tables, one upload and one read-back
        per call): the reference for the resident path."""
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
