"""Fused cluster-pair path (energy + forces in one call) against the float64 oracle (GPU).

Tolerances: fp64 mode energy terms to 1e-9 of the summed |terms| and force RMS
1e-9; mixed mode (float32 arithmetic, int64 fixed-point accumulation) energy to
1e-7 of the summed |terms| and force RMS relative error <= 1e-4 (the stated
target for the 96k-atom box).  The mixed energy bound is relative to the summed
magnitudes, not to the net energy: float32 rounding of the charge products and of
s^6 / s^12 is the same for every pair of one atom-type pair, so those per-pair
errors (a few units of 2^-24) add with one sign instead of averaging out.  The
errors seen are about 5e-8 (900 atoms) and 1e-8 (96 000 atoms) of that sum.
"""

import numpy as np
import pytest
import torch

from oracle import composite
from torchff.workloads import make_workload, water_box, ewald_alpha

pytestmark = pytest.mark.gpu


def _run(w, precision, dtype=torch.float64):
    from torchff.fused import for_workload
    m = for_workload(w, precision)
    dev = "cuda"
    E, F, terms = m(torch.tensor(w["coords"], dtype=dtype, device=dev), torch.tensor(w["box"], dtype=dtype, device=dev),
                    torch.tensor(w["charges"], dtype=dtype, device=dev), torch.tensor(w["sigma"], dtype=dtype, device=dev),
                    torch.tensor(w["epsilon"], dtype=dtype, device=dev), check=True)
    return float(E), F.double().cpu().numpy(), terms.cpu().numpy()


@pytest.fixture(scope="module", params=["water900", "water96k"])
def case(request):
    w = make_workload(request.param)
    E, F, terms, _ = composite.energy_forces(w)
    return w, E, F, terms


def _rms_rel(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


@pytest.mark.parametrize("precision,te,tf", [("fp64", 1e-9, 1e-9), ("mixed", 1e-7, 1e-4)])
def test_fused_matches_oracle(case, precision, te, tf):
    w, E, F, terms = case
    Ed, Fd, td = _run(w, precision)
    scale = sum(abs(v) for v in terms.values())
    assert abs(Ed - E) < te * scale, (Ed, E, td, terms)
    assert abs(td[0] - terms["lj"]) < te * scale
    assert abs(td[1] - terms["coul_real"]) < te * scale
    assert abs(td[2] + td[3] - terms["coul_recip"]) < te * scale
    rms = _rms_rel(Fd, F)
    print(f"{w['name']} {precision}: dE={Ed - E:.3e} (x{1.0 / scale:.2e} = {abs(Ed - E) / scale:.2e} of sum|terms|) "
          f"dLJ={td[0] - terms['lj']:.2e} dCreal={td[1] - terms['coul_real']:.2e} "
          f"dCrecip={td[2] + td[3] - terms['coul_recip']:.2e} force_rms_rel={rms:.3e}")
    assert rms < tf


def test_fused_float32_io():
    w = make_workload("water900")
    E, F, terms, _ = composite.energy_forces(w)
    Ed, Fd, _ = _run(w, "mixed", torch.float32)
    assert _rms_rel(Fd, F) < 2e-4


def test_fused_deterministic():
    w = make_workload("water900")
    a = _run(w, "mixed")
    b = _run(w, "mixed")
    assert a[0] == b[0]
    assert np.array_equal(a[1], b[1])


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_fused_schedules_identical(precision):
    # reciprocal chain serial (0), by size (1), spread before the fork (2), all on the
    # auxiliary stream (3): integer accumulation makes every schedule bitwise identical
    from torchff import _lib as L
    w = make_workload("water900")
    lib = L.load()
    outs = {}
    try:
        for mode in (1, 0, 2, 3):
            lib.tff_fused_set_overlap(mode)
            outs[mode] = _run(w, precision)
    finally:
        lib.tff_fused_set_overlap(1)
    for mode in (0, 2, 3):
        assert outs[mode][0] == outs[1][0]
        assert np.array_equal(outs[mode][1], outs[1][1])
        assert np.array_equal(outs[mode][2], outs[1][2])


def test_fused_small_odd_box():
    # 200 waters at 1 g/cm3 (L ~ 1.81 nm would violate 2 rc; use rc 0.8) with an odd grid
    from torchff.workloads import box_length_for_density
    n = 216
    L = box_length_for_density(n)
    w = water_box(n, L, seed=21)
    w.update(cutoff=0.8, skin=0.05, K=25, order=5, alpha=ewald_alpha(0.8))
    w["name"] = "water648"
    E, F, terms, _ = composite.energy_forces(w)
    Ed, Fd, _ = _run(w, "fp64")
    scale = sum(abs(v) for v in terms.values())
    assert abs(Ed - E) < 1e-9 * scale
    assert _rms_rel(Fd, F) < 1e-9


@pytest.mark.parametrize("precision,te,tf", [("fp64", 1e-9, 1e-9), ("mixed", 1e-7, 1e-4)])
def test_fused_shared_z_planes(precision, te, tf):
    # one 30 x 30 layer of waters whose oxygens all sit at exactly the same z (no jitter), in a
    # 9.3 x 9.3 x 6.0 nm box: about 36 oxygens of every x-y column fall into one z bucket of the
    # column sort, above its 32-key bucket limit, so the columns take the bitonic fallback
    w = water_box(900, 9.3, n_lat=30, seed=41, jitter=0.0)
    w["coords"] = np.ascontiguousarray(w["coords"][:, [1, 2, 0]])
    w["box"] = np.diag([9.3, 9.3, 6.0])
    assert np.sum(w["coords"][0::3, 2] == w["coords"][0, 2]) == 900
    w.update(cutoff=0.9, skin=0.0, K=64, order=5, alpha=ewald_alpha(0.9))
    w["name"] = "water_layer2700"
    E, F, terms, _ = composite.energy_forces(w)
    Ed, Fd, _ = _run(w, precision)
    scale = sum(abs(v) for v in terms.values())
    assert abs(Ed - E) < te * scale, (Ed, E)
    assert _rms_rel(Fd, F) < tf


@pytest.mark.parametrize("precision,te,tf", [("fp64", 1e-9, 1e-9), ("mixed", 1e-7, 1e-4)])
def test_fused_tiny_grid(precision, te, tf):
    # 64 waters at 1 g/cm3 (L ~ 1.24 nm), rc 0.6, a 10^3 grid: every axis is shorter than a
    # spreading brick plus a stencil, so the stencil wraps into both ends of a brick
    from torchff.workloads import box_length_for_density
    n = 64
    L = box_length_for_density(n)
    w = water_box(n, L, seed=33)
    w.update(cutoff=0.6, skin=0.0, K=10, order=5, alpha=ewald_alpha(0.6))
    w["name"] = "water192"
    E, F, terms, _ = composite.energy_forces(w)
    Ed, Fd, _ = _run(w, precision)
    scale = sum(abs(v) for v in terms.values())
    assert abs(Ed - E) < te * scale, (Ed, E)
    assert _rms_rel(Fd, F) < tf


def _tensors(w, dev="cuda"):
    return [torch.tensor(w[k], dtype=torch.float64, device=dev)
            for k in ("coords", "box", "charges", "sigma", "epsilon")]


def _same(a, b):
    assert float(a[0]) == float(b[0])
    assert torch.equal(a[1], b[1])
    assert torch.equal(a[2], b[2])


@pytest.mark.parametrize("precision", ["mixed", "fp64"])
def test_fused_graph_replay_matches_one_by_one(precision):
    """The forward launches captured once into a graph and replayed give, bitwise, what the same
    kernels launched one by one give: with the coordinates changed in place between calls (same
    addresses, so the same graph), with other work and synchronisations between replays, and
    after a box change (a new key, so a new capture)."""
    from torchff import _lib as L
    from torchff.fused import for_workload
    lib = L.load()
    w = make_workload("water900")
    m = for_workload(w, precision)
    t = _tensors(w)
    gen = torch.Generator(device="cuda").manual_seed(5)
    try:
        for step in range(4):
            if step:
                t[0].add_(0.01 * torch.randn(t[0].shape, generator=gen, device="cuda", dtype=torch.float64))
            lib.tff_fused_set_graphs(2)
            g1 = m(*t)  # the first call of a key captures (then launches the graph), later ones replay
            g2 = m(*t)
            g3 = m(*t)
            torch.randn(4096, device="cuda").sum().item()  # other work and a synchronisation
            lib.tff_fused_set_graphs(0)
            e = m(*t)
            for g in (g1, g2, g3):
                _same(g, e)
        t[1].mul_(1.01)
        t[0].mul_(1.01)
        lib.tff_fused_set_graphs(2)
        g = m(*t)
        g2 = m(*t)
        lib.tff_fused_set_graphs(0)
        e = m(*t)
        _same(g, e)
        _same(g2, e)
    finally:
        lib.tff_fused_set_graphs(1)
    m.check()


def test_fused_inside_torch_graph_capture():
    """A forward call captured into a torch CUDA graph (the library then launches its kernels one by
    one into the capture) replays to the one-by-one results of the same inputs, bitwise, with the
    coordinates changed in place between replays; check() reads the last replay's status."""
    from torchff.fused import for_workload
    w = make_workload("water900")
    m = for_workload(w, "mixed")
    t = _tensors(w)
    ref = for_workload(w, "mixed")
    m(*t)  # eager first call: workspace, status ring and box cache
    torch.cuda.synchronize()
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        m(*t)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        out = m(*t)
    gen = torch.Generator(device="cuda").manual_seed(7)
    for rep in range(3):
        if rep:
            t[0].add_(0.01 * torch.randn(t[0].shape, generator=gen, device="cuda", dtype=torch.float64))
        torch.randn(4096, device="cuda").sum().item()
        g.replay()
        torch.cuda.synchronize()
        _same(out, ref(*t))
    m.check()
    ref.check()
