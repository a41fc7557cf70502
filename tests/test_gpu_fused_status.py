"""Status reporting of the fused path (GPU): each capacity flag is raised with its own cause, on
a later call without check() and at once with check(), and a box changed behind the host cache
is detected on the device.

Inputs that set each flag (all coordinates finite, every kernel exits normally):
  * pair capacity: the 96k-atom water box with z squeezed three-fold (the per-cluster neighbour
    count is three times what the box average sizes the segments for);
  * brick candidates (float64 spread): the 96k-atom box on a 16^3 grid (one 8^3 brick covers half
    the box, so most of the 3000 clusters reach it, above the 1024 kept per brick);
  * brick list (float spread): a 6 x 6 x 6 lattice of 1/6 nm spacing inside one 1 nm brick of a
    6 nm box on a 48^3 grid, the rest of the box a 0.43 nm lattice: 216 stencil bases in one brick
    against a capacity of 4 x 14 + 64 = 120;
  * exclusion partners: atom 0 of the 900-atom box given nine extra partners;
  * column size: a 12 x 12 x 560 lattice (0.208 nm spacing) in a 2.5 x 2.5 x 116.7 nm box,
    8960 atoms per x-y column (above the 8192 sorted in LDS: the column stays unsorted in z).
"""

import numpy as np
import pytest
import torch

import torchff
from torchff.fused import FusedNonbondedPME, for_workload
from torchff.workloads import ewald_alpha, make_workload

pytestmark = pytest.mark.gpu


def _args(w, dt=torch.float64):
    return [torch.tensor(w[k], dtype=dt, device="cuda") for k in ("coords", "box", "charges", "sigma", "epsilon")]


def _squeezed96k():
    w = make_workload("water96k")
    w["coords"] = np.array(w["coords"])
    w["coords"][:, 2] /= 3.0
    return w


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_pair_overflow_raises_with_check(precision):
    w = _squeezed96k()
    m = for_workload(w, precision)
    with pytest.raises(RuntimeError, match="cluster-pair capacity"):
        m(*_args(w), check=True)


def test_pair_overflow_raises_on_a_later_call():
    w = _squeezed96k()
    m = for_workload(w, "mixed")
    args = _args(w)
    m(*args)  # no check: returns
    torch.cuda.synchronize()
    with pytest.raises(RuntimeError, match="earlier forward call returned invalid results: cluster-pair capacity"):
        m(*args)


def test_brick_overflow_has_its_own_message():
    w = make_workload("water96k", K=16)
    m = for_workload(w, "fp64")
    with pytest.raises(RuntimeError, match="charge-spreading brick") as ei:
        m(*_args(w), check=True)
    assert "cluster-pair capacity" not in str(ei.value)


def _dense_brick_box():
    """216 atoms on a 1/6 nm lattice inside the brick [2.25, 3.25)^3 nm of a 6 nm box (grid 48^3,
    bricks of 8 points = 1 nm), the rest of the box a 14^3 lattice of 0.43 nm spacing."""
    L = 6.0
    dense = (np.stack(np.meshgrid(*[np.arange(6)] * 3, indexing="ij"), -1).reshape(-1, 3) + 0.5) / 6.0 + 2.25
    sparse = (np.stack(np.meshgrid(*[np.arange(14)] * 3, indexing="ij"), -1).reshape(-1, 3) + 0.5) * (L / 14)
    inside = np.all((sparse >= 2.0) & (sparse < 3.5), axis=1)
    x = np.concatenate([dense, sparse[~inside]])
    n = len(x)
    q = np.where(np.arange(n) % 2 == 0, 0.4, -0.4)
    return dict(coords=x, box=np.diag([L, L, L]), charges=q, sigma=np.full(n, 0.15), epsilon=np.full(n, 0.3))


def test_brick_list_overflow_float_spread():
    w = _dense_brick_box()
    n = len(w["coords"])
    assert 4 * -(-n // 216) + 64 < 216
    m = FusedNonbondedPME(n, 0.9, ewald_alpha(0.9), (48, 48, 48), order=5, precision="mixed")
    with pytest.raises(RuntimeError, match="charge-spreading brick") as ei:
        m(*_args(w), check=True)
    assert "cluster-pair capacity" not in str(ei.value)
    # the same system on the float64 spread (no lists) is within its capacities
    m64 = FusedNonbondedPME(n, 0.9, ewald_alpha(0.9), (48, 48, 48), order=5, precision="fp64")
    E, F = m64(*_args(w), check=True)[:2]
    assert torch.isfinite(E).all() and torch.isfinite(F).all()


def test_exclusion_overflow():
    w = make_workload("water900")
    extra = np.array([[0, k] for k in range(3, 12)], dtype=np.int64)
    w["exclusions"] = np.concatenate([w["exclusions"], extra])
    m = for_workload(w, "mixed")
    with pytest.raises(RuntimeError, match="more than 8 exclusion partners"):
        m(*_args(w), check=True)


def test_column_overflow():
    n_xy, n_z, a = 12, 560, 2.5 / 12
    g = (np.stack(np.meshgrid(np.arange(n_xy), np.arange(n_xy), np.arange(n_z), indexing="ij"), -1)
         .reshape(-1, 3) + 0.5) * a
    n = len(g)
    q = np.where(np.arange(n) % 2 == 0, 0.4, -0.4)
    w = dict(coords=g, box=np.diag([2.5, 2.5, n_z * a]), charges=q, sigma=np.full(n, 0.3),
             epsilon=np.full(n, 0.3))
    m = FusedNonbondedPME(n, 0.9, ewald_alpha(0.9), (16, 16, 600), order=5, precision="mixed")
    args = _args(w)
    m(*args)
    torch.cuda.synchronize()
    with pytest.raises(RuntimeError, match="earlier forward call returned invalid results: .*more than 8192 atoms"):
        m(*args)
    with pytest.raises(RuntimeError, match="more than 8192 atoms"):
        m(*args, check=True)


def test_stale_box_is_detected_and_invalidation_recovers():
    w = make_workload("water900")
    m = for_workload(w, "fp64")
    args = _args(w)
    box = args[1]
    m(*args, check=True)
    box.data[2, 2] += 0.05  # bypasses the version counter: the cached host copy is stale
    with pytest.raises(RuntimeError, match="box tensor was changed"):
        m(*args, check=True)
    torchff.invalidate_box_cache(box)
    E1, F1, _ = m(*args, check=True)
    fresh = [a.clone() for a in args]
    E2, F2, _ = for_workload(w, "fp64")(*fresh, check=True)
    assert float(E1) == float(E2) and torch.equal(F1, F2)


def test_dropin_reads_the_box_every_call_without_the_cache():
    w = make_workload("water900")
    x, box, q = (torch.tensor(w[k], dtype=torch.float64, device="cuda") for k in ("coords", "box", "charges"))
    pme = torchff.PME(w["alpha"], w["K"], 0, order=w["order"]).to(device="cuda", dtype=torch.float64)
    pme(x, box, q)
    box.data[0, 0] += 0.03
    ref = float(pme(x, box.clone(), q))
    torchff.set_box_cache(False)
    try:
        got = float(pme(x, box, q))
    finally:
        torchff.set_box_cache(True)
    assert got == ref
