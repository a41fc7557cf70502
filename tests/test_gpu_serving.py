"""Batched serving path against the all-images float64 oracle, and its batching invariants (GPU).

Tolerances (per system): fp64 energy terms to 1e-9 of the summed |terms| and force
RMS relative error 1e-9; mixed (float32 arithmetic, int64 fixed-point charge grid,
float64 sums) energy to 1e-6 of the summed |terms| and force RMS 1e-4.  Results
of one system are bitwise the same alone, in any batch, and from a graph replay.
"""

import numpy as np
import pytest
import torch

from oracle import images
from torchff.serving import ServingEngine, latency_stats, request_stream, water_request
from torchff.workloads import ewald_alpha

pytestmark = pytest.mark.gpu

RC = 0.9


def _oracle(r, alpha, order=5):
    rec = r.records
    w = dict(coords=rec[:, :3], box=np.diag(r.box_lengths), charges=rec[:, 3], sigma=rec[:, 4], epsilon=rec[:, 5])
    return images.energy_forces(w, RC, alpha, r.grid, order)


def _rms_rel(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


# 20 waters: edge 0.84 nm < rc; 35: rc < edge < 2 rc; 100: edge 1.44 < 2 rc; 300: edge 2.08 >= 2 rc;
# 1000: the largest request size of the serving stream (edge 3.10 nm, 32^3 grid)
SIZES = [(20, 101), (35, 102), (100, 103), (300, 104), (1000, 105)]


@pytest.fixture(scope="module")
def oracle_cases():
    reqs = [water_request(i, n, seed) for i, (n, seed) in enumerate(SIZES)]
    return reqs, [_oracle(r, ewald_alpha(RC)) for r in reqs]


def _fresh(reqs):
    return [water_request(r.rid, r.natoms // 3, 101 + r.rid) for r in reqs]


@pytest.mark.parametrize("precision,te,tf", [("fp64", 1e-9, 1e-9), ("mixed", 1e-6, 1e-4)])
def test_serve_matches_oracle(oracle_cases, precision, te, tf):
    reqs, refs = oracle_cases
    eng = ServingEngine(slots=1, max_k=32, precision=precision)
    out = eng.evaluate(_fresh(reqs))
    for r, (E, F, terms) in zip(out, refs):
        e = r.energies
        scale = abs(terms["real"]) + abs(terms["recip_self"])
        print(f"{r.natoms} atoms {precision}: dE={e[4] - E:.3e} ({abs(e[4] - E) / scale:.2e} of sum|terms|) "
              f"dreal={e[0] - terms['real']:.2e} drecip={e[2] + e[3] - terms['recip_self']:.2e} "
              f"force_rms_rel={_rms_rel(r.forces, F):.3e}")
        assert abs(e[0] - terms["real"]) < te * scale
        assert abs(e[2] + e[3] - terms["recip_self"]) < te * scale
        assert abs(e[4] - E) < te * scale
        assert _rms_rel(r.forces, F) < tf


def test_same_alone_and_in_mix():
    reqs = request_stream(24, min_waters=20, max_waters=400, seed=5)
    eng = ServingEngine(slots=2, max_k=32)
    mixed = eng.evaluate(reqs)
    for r in mixed:
        alone = eng.evaluate([water_request(r.rid, r.natoms // 3, 5 * 100003 + r.rid)])[0]
        assert np.array_equal(alone.energies, r.energies), r.rid
        assert np.array_equal(alone.forces, r.forces), r.rid


def test_graph_replays_match_direct_launches():
    reqs = request_stream(12, min_waters=20, max_waters=300, seed=6)
    direct = ServingEngine(slots=1, max_k=32, use_graph=False).evaluate(reqs)
    ref = [(r.energies.copy(), r.forces.copy()) for r in direct]
    eng = ServingEngine(slots=1, max_k=32, use_graph=True)
    for rep in range(3):
        # eager device work between replays
        x = torch.randn(1 << 20, device="cuda")
        torch.cuda.synchronize()
        assert float(x.abs().sum()) > 0
        got = eng.evaluate(request_stream(12, min_waters=20, max_waters=300, seed=6))
        for (e, f), r in zip(ref, got):
            assert np.array_equal(e, r.energies), (rep, r.rid)
            assert np.array_equal(f, r.forces), (rep, r.rid)


@pytest.mark.parametrize("use_graph", [True, False])
def test_results_stored_to_host_match_device_output(use_graph):
    """The request path's results, stored into pinned host memory by the last kernel, equal bit for
    bit those of the same uploaded batch evaluated into device memory (the device-resident path),
    over two batches in a row."""
    eng = ServingEngine(slots=1, max_k=32, use_graph=use_graph)
    s = eng.slots[0]
    for seed in (8, 9):
        for r in request_stream(16, min_waters=20, max_waters=600, seed=seed):
            assert s.try_add(r)
        f0, nsys = 5 * eng.caps["max_systems"], len(s.requests)
        s.launch()
        s.event.synchronize()
        host = s.h_out.clone()
        with torch.cuda.stream(s.stream):
            s.run()  # the same uploaded batch, results into d_out
        s.stream.synchronize()
        dev = s.d_out.cpu()
        # the written parts: 5 energies per system, 3 forces per atom
        assert torch.equal(host[:5 * nsys], dev[:5 * nsys]), seed
        assert torch.equal(host[f0: f0 + 3 * s.natoms], dev[f0: f0 + 3 * s.natoms]), seed
        s.finish(0.0)


def test_serve_stream_realtime():
    reqs = request_stream(200, rate=2000.0, min_waters=20, max_waters=1000, seed=7)
    eng = ServingEngine(slots=2, max_k=32)
    done, nb = eng.serve(reqs, realtime=True)
    assert all(r.t_done is not None and r.t_done >= r.t_submit for r in done)
    st = latency_stats(done)
    print(f"200 requests in {nb} batches: p50 {st['p50_ms']:.3f} ms p99 {st['p99_ms']:.3f} ms "
          f"{st['evals_per_s']:.0f} evals/s")
    # spot-check results against single evaluations
    check = [done[i] for i in (0, 57, 199)]
    for r in check:
        alone = eng.evaluate([water_request(r.rid, r.natoms // 3, 7 * 100003 + r.rid)])[0]
        assert np.array_equal(alone.forces, r.forces)


def test_oversized_request_is_refused():
    eng = ServingEngine(slots=1, max_atoms=1024, max_k=32, use_graph=False)
    with pytest.raises(RuntimeError):
        eng.evaluate([water_request(0, 400, 1)])
