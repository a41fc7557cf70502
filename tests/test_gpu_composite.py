"""TIP3P composition (LJ type table + rank-0 Ewald real space with exclusions + PME) on device vs oracle (GPU)."""

import numpy as np
import pytest
import torch

from oracle import composite
from torchff.workloads import COULOMB_KJMOL, make_workload

pytestmark = pytest.mark.gpu


def _device_energy_forces(w, dtype):
    import torchff
    dev = "cuda"
    n = w["natoms"]
    cut = w["cutoff"]
    x = torch.tensor(w["coords"], dtype=dtype, device=dev, requires_grad=True)
    box = torch.tensor(w["box"], dtype=dtype, device=dev)
    q = torch.tensor(w["charges"], dtype=dtype, device=dev)
    nl = torchff.NeighborList(n, exclusions=torch.tensor(w["exclusions"], device=dev), algorithm="cell_list")
    cap = int(n * 4 / 3 * 3.1416 * (cut + w["skin"]) ** 3 * 100 / 2 * 1.3) + 1000
    pairs, npairs = nl(x.detach(), box, cut + w["skin"], cap, padding=True)
    lj = torchff.Vdw("LennardJones", use_type_pairs=True)
    coul = torchff.MultipolarInteraction(0, cut, w["alpha"], COULOMB_KJMOL)
    pme = torchff.PME(w["alpha"], w["K"], 0, order=w["order"]).to(device=dev, dtype=dtype)
    e_lj = lj(x, pairs, box, torch.tensor(w["sigma_table"], dtype=dtype, device=dev),
              torch.tensor(w["epsilon_table"], dtype=dtype, device=dev), cut,
              atom_types=torch.tensor(w["atom_types"], device=dev))
    e_c = coul(x, box, pairs, q, pairs_excl=torch.tensor(w["pairs_excl"], device=dev))
    e_r = pme(x, box, q) * COULOMB_KJMOL
    E = e_lj + e_c + e_r
    E.backward()
    return float(E), -x.grad.double().cpu().numpy(), int(npairs.item())


@pytest.mark.parametrize("name", ["water900", "water96k"])
def test_energy_forces_match_oracle(name):
    w = make_workload(name)
    E, F, terms, pairs = composite.energy_forces(w)
    for dtype, te, tf in ((torch.float64, 1e-9, 1e-9), (torch.float32, 1e-5, 1e-4)):
        Ed, Fd, npairs = _device_energy_forces(w, dtype)
        if dtype == torch.float64:
            assert npairs == len(pairs)
        else:
            # float32 coordinates move pairs within ~1e-7 relative of the list cutoff across it; the exact
            # float32 set is checked against the oracle on the rounded values in test_gpu_nblist.py
            assert abs(npairs - len(pairs)) <= 1e-6 * len(pairs) + 2
        # energy: relative to the sum of |terms| (the total is a near-cancellation)
        scale = sum(abs(v) for v in terms.values())
        assert abs(Ed - E) < te * scale, (dtype, Ed, E)
        rms = float(np.sqrt(np.mean((Fd - F) ** 2)) / np.sqrt(np.mean(F ** 2)))
        assert rms < tf, (dtype, rms)
