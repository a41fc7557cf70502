"""Energy + force evaluations per second on the 96,000-atom PME water box (one MI355X).

One step = one full evaluation of the TIP3P nonbonded energy and forces of
32,000 rigid waters (L = 9.86 nm): spatial clustering / neighbour search rebuilt
every call at rc + skin = 1.1 nm with the intramolecular exclusions,
erfc-Coulomb + Lennard-Jones at rc = 1.0 nm with the (erfc - 1)/r exclusion
correction, order-5 PME on a 96^3 grid (spread, FFTs, influence function,
gather), self energy; forces = -dE/dr.  Inputs are resident on the device
before timing.

Paths measured (same inputs):
  mixed    FusedNonbondedPME(precision="mixed")  -> the headline value
  fp64     FusedNonbondedPME(precision="fp64")
  dropin   the reference-shaped modules in float64 (NeighborList cell_list with an
           explicit (P, 2) pair list + Nonbonded(ewald) + PME, autograd forces):
           the plain FP64 drop-in the speed-up is quoted against.
Errors are stated against one float64 CPU oracle evaluation, which is also the
timed CPU baseline.  The line also carries the batched-serving results ("serving":
p50/p99 latency of 4096 Poisson requests at 2000 req/s of 20-1000-water boxes,
batched against one system per launch, and saturation throughput).  Prints one
JSON line.  --gpus above 1 is refused.
"""

import argparse
import json
import math
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "torchff-lib_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402
import torch  # noqa: E402

HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6
FP32_PEAK_TFLOPS = 157.3
# floating-point operations per in-cutoff pair, a fixed tally: minimum image 9, r^2 5, rsqrt 1,
# erfc + exp 10, Coulomb 4, LJ 8, force scalar 8, force vector 3, accumulations 12
PAIR_FLOPS = 60


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="water96k")
    ap.add_argument("--dropin-steps", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-samples", type=int, default=3, help="timed CPU oracle evaluations (median reported)")
    ap.add_argument("--only", default="", help="comma list of paths to run (mixed,fp64,dropin)")
    ap.add_argument("--differentiation", action="store_true",
                    help="time force-matching steps (E, F, loss, gradients wrt coords, q, sigma, eps) instead")
    ap.add_argument("--widened", action="store_true",
                    help="multipolar PME (ranks 0-2) and the direct Ewald op, energy + gradients per call")
    ap.add_argument("--md", action="store_true",
                    help="only time the torch-sim style consumer (16 boxes per call against one call per box)")
    ap.add_argument("--serving", action="store_true",
                    help="serve 4096 Poisson requests (20-1000 waters) batched and one system per launch instead")
    ap.add_argument("--requests", type=int, default=4096)
    ap.add_argument("--no-serving", action="store_true", help="leave the serving results out of the main line")
    ap.add_argument("--serving-sizes", default="", metavar="N,N,...",
                    help="only time single-system serving batches of these water counts")
    ap.add_argument("--rate", type=float, default=2000.0)
    ap.add_argument("--calibrate-copy", type=float, default=0.0, metavar="MB",
                    help="only run a plain device copy of MB megabytes (to calibrate HBM byte counters)")
    return ap.parse_args()


def calibrate_copy(mb, steps):
    """A plain device-to-device copy of known size: the FETCH_SIZE / WRITE_SIZE counters read
    for it give the factor between counter KiB and bytes moved."""
    n = int(mb * 1e6) // 4
    x = torch.ones(n, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(2):
        y.copy_(x)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        y.copy_(x)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / steps
    print(json.dumps({"calibration": "device copy", "bytes_read": 4 * n, "bytes_written": 4 * n,
                      "ms": sec * 1e3, "GB/s": 8 * n / sec / 1e9}))


TRAFFIC_PROFILE = os.path.join("profiles", "fused_tiles_traffic_water96k.md")
KERNEL_SOURCE = os.path.join("torchff-lib_amd", "csrc", "fused.hip")


def source_sha256(rel):
    import hashlib
    with open(os.path.join(ROOT, rel), "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def tile_traffic():
    """HBM bytes per launch of the pair kernel from the committed counter summary (FETCH_SIZE +
    WRITE_SIZE, scaled by the copy-kernel calibration), and where it came from.  The summary
    records the sha256 of the kernel source it was measured on; a summary of another source
    gives traffic None (stale)."""
    path = os.path.join(ROOT, TRAFFIC_PROFILE)
    info = {"profile": TRAFFIC_PROFILE, "kernel_source_sha256": source_sha256(KERNEL_SOURCE)}
    if not os.path.exists(path):
        return None, info
    vals = {}
    with open(path) as f:
        for line in f:
            line = line.strip().lstrip("- ")
            if line.startswith(("fetch_bytes_per_launch:", "write_bytes_per_launch:", "fused_hip_sha256:")):
                k, v = line.split(":", 1)
                vals[k.strip()] = v.strip().split()[0]
    info["measured_on_this_source"] = vals.get("fused_hip_sha256") == info["kernel_source_sha256"]
    if not info["measured_on_this_source"] or "fetch_bytes_per_launch" not in vals \
            or "write_bytes_per_launch" not in vals:
        return None, info
    info["fetch_bytes"] = float(vals["fetch_bytes_per_launch"])
    info["write_bytes"] = float(vals["write_bytes_per_launch"])
    return info["fetch_bytes"] + info["write_bytes"], info


def pme_stage_bytes(w):
    """Algorithmic HBM bytes per evaluation of the reciprocal-space stages (N atoms, K^3 grid,
    float32 grid, int64 fixed-point charge grid, real-input half spectrum along z)."""
    n = w["natoms"]
    k = w["K"]
    half = k * k * (k // 2 + 1) * 8          # complex float half spectrum
    grid64, grid32 = 8 * k ** 3, 4 * k ** 3
    return {
        "spread": n * 16 + grid64,                                   # read x,y,z,q; write the int64 grid
        "fft": (grid64 + half) + 4 * half + 4 * half + (half + grid32),  # z r2c, y, x+conv+x, y, z c2r
        "gather": grid32 + n * 16 + n * 24,                          # read potential + atoms; write forces
    }


class DropIn:
    """Reference-shaped modules (NeighborList + Nonbonded(ewald) + PME) in float64, autograd forces."""

    def __init__(self, w, dev):
        import torchff
        from torchff.workloads import COULOMB_KJMOL
        dt = torch.float64
        self.k = COULOMB_KJMOL
        n = w["natoms"]
        self.cut = w["cutoff"]
        self.rl = w["cutoff"] + w["skin"]
        self.x = torch.tensor(w["coords"], dtype=dt, device=dev)
        self.box = torch.tensor(w["box"], dtype=dt, device=dev)
        self.q = torch.tensor(w["charges"], dtype=dt, device=dev)
        self.sig = torch.tensor(w["sigma"], dtype=dt, device=dev)
        self.eps = torch.tensor(w["epsilon"], dtype=dt, device=dev)
        self.excl = torch.tensor(w["pairs_excl"], device=dev)
        self.nl = torchff.NeighborList(n, exclusions=torch.tensor(w["exclusions"], device=dev),
                                       algorithm="cell_list")
        self.cap = int(n * 4.0 / 3.0 * math.pi * self.rl ** 3 * 100.0 / 2.0 * 1.25)
        self.nb = torchff.Nonbonded()
        self.pme = torchff.PME(w["alpha"], w["K"], 0, order=w["order"]).to(device=dev, dtype=dt)
        self.alpha = w["alpha"]

    def step(self):
        x = self.x.detach().requires_grad_(True)
        pairs, _ = self.nl(x.detach(), self.box, self.rl, self.cap, padding=True)
        e = self.nb(x, pairs, self.box, self.sig, self.eps, self.q, self.k, self.cut, ewald_alpha=self.alpha,
                    pairs_excl=self.excl).sum() + self.pme(x, self.box, self.q) * self.k
        e.backward()
        return e.detach(), -x.grad


class Fused:
    def __init__(self, w, dev, precision):
        from torchff.fused import for_workload
        dt = torch.float64
        self.m = for_workload(w, precision)
        self.args = [torch.tensor(w[k], dtype=dt, device=dev) for k in ("coords", "box", "charges", "sigma",
                                                                          "epsilon")]

    def step(self):
        E, F, _ = self.m(*self.args)
        return E, F


def time_steps(fn, steps, warmup):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    out = None
    for _ in range(steps):
        out = fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, out


def kernel_times(fn, steps, names):
    from torchff import _lib
    _lib.prof_collect()  # drop anything pending
    _lib.prof_enable(True)
    time_steps(fn, steps, 0)
    _lib.prof_enable(False)
    out = {}
    for name in names:
        tot, cnt = _lib.prof_collect(name)
        if cnt:
            out[name] = {"avg_ms": tot / cnt, "launches": cnt}
    return out


def differentiation_bench(args, w, dev):
    """One force-matching step: E, F = model(coords, q, sigma, eps); L = mean((F - T)^2); dL/d(all).
    Inputs: the workload with charges perturbed by N(0, 0.02) (neutral per molecule), sigma and eps
    scaled by U(0.95, 1.05), target forces N(0, 50 kJ/mol/nm)."""
    from torchff.fused import for_workload
    from torchff.workloads import differentiation_inputs
    q, s, e, T = differentiation_inputs(w)
    Tt = torch.tensor(T, device=dev)
    res, grads = {}, {}
    for prec in ("mixed", "fp64"):
        m = for_workload(w, prec)
        c = torch.tensor(w["coords"], dtype=torch.float64, device=dev, requires_grad=True)
        box = torch.tensor(w["box"], dtype=torch.float64, device=dev)
        qt, st, et = (torch.tensor(a, dtype=torch.float64, device=dev, requires_grad=True) for a in (q, s, e))

        def step():
            _, F, _ = m(c, box, qt, st, et)
            L = torch.mean((F - Tt) ** 2)
            return torch.autograd.grad(L, [c, qt, st, et])

        steps = args.steps if prec == "mixed" else max(3, args.steps // 4)
        sec, out = time_steps(step, steps, args.warmup)
        m.check()
        res[prec] = sec
        grads[prec] = [g.double().cpu().numpy() for g in out]
    names = ["coords", "charges", "sigma", "epsilon"]
    na = w["natoms"]
    print(json.dumps({
        "metric": f"force-matching steps/s (E, F, loss, gradients wrt coords, q, sigma, eps), {na // 1000}k-atom "
                  "PME water",
        "value": 1.0 / res["mixed"], "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": res["mixed"] * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+i64", "data": "synthetic",
        "config": {"workload": f"{args.workload} with perturbed q / sigma / eps and target forces N(0, 50)",
                   "path": "FusedNonbondedPME forward + hand-written force-matching backward (the forward's "
                           "cluster arrays, pair list and potential grid reused)"},
        "paths_steps_per_s": {k: 1.0 / v for k, v in res.items()},
        "mixed_vs_fp64_gradients_rms_rel": {k: rms_rel(a, b) for k, a, b in zip(names, grads["mixed"], grads["fp64"])},
    }))


def widened_bench(args, w, dev):
    """Reciprocal space for point dipoles and quadrupoles (PME ranks 1 and 2 beside rank 0) on the
    workload's box and grid, and the direct Ewald op on the 900-atom box: one call = energy and the
    gradients of every input (coords, q, p, theta).  Dipoles N(0, 0.02) e nm, quadrupoles
    symmetric traceless N(0, 0.01) e nm^2."""
    import torchff
    from torchff.workloads import make_workload

    def inputs(wk, dt, rank, seed):
        rng = np.random.default_rng(seed)
        n = wk["natoms"]
        t = lambda a: torch.tensor(a, dtype=dt, device=dev)
        x, box, q = t(wk["coords"]), t(wk["box"]), t(wk["charges"])
        p = t(rng.normal(0.0, 0.02, (n, 3))) if rank >= 1 else None
        th = None
        if rank >= 2:
            a = rng.normal(0.0, 0.01, (n, 3, 3))
            a = 0.5 * (a + a.transpose(0, 2, 1))
            a -= np.trace(a, axis1=1, axis2=2)[:, None, None] * np.eye(3)[None] / 3.0
            th = t(a)
        return x, box, q, p, th

    def run(mod, x, box, q, p, th, steps, names):
        leaves = [v.requires_grad_(True) for v in (x, q, p, th) if v is not None]

        def step():
            E = mod(x, box, q, p, th)
            return torch.autograd.grad(E, leaves)

        sec, _ = time_steps(step, steps, 2)
        kt = kernel_times(step, steps, names) if names else {}
        return sec, {k: v["avg_ms"] for k, v in kt.items()}

    res = {}
    for dt, tag in ((torch.float32, "f32"), (torch.float64, "f64")):
        for rank in (0, 1, 2):
            m = torchff.PME(w["alpha"], w["K"], rank, order=w["order"]).to(device=dev, dtype=dt)
            sec, kt = run(m, *inputs(w, dt, rank, rank), max(5, args.steps // 5),
                          ["pme_spread", "pme_fft", "pme_fft_conv", "pme_gather"])
            res[f"pme_rank{rank}_{tag}"] = {"ms": sec * 1e3, "evals_per_s": 1.0 / sec, "kernels_ms": kt}
    w9 = make_workload("water900")
    for rank in (0, 1, 2):
        m = torchff.Ewald(w9["alpha"], 10, rank, return_fields=False)
        sec, _ = run(m, *inputs(w9, torch.float64, rank, 10 + rank), 5, None)
        res[f"ewald_rank{rank}_f64_water900_kmax10"] = {"ms": sec * 1e3, "evals_per_s": 1.0 / sec}
    res.update(md_consumer_results(dev))
    head = res["pme_rank2_f32"]
    print(json.dumps({
        "metric": f"PME energy + gradient evals/s, point charges + dipoles + quadrupoles, {w['natoms'] // 1000}k "
                  "atoms",
        "value": head["evals_per_s"], "unit": "evals/s", "n_gpus": 1, "steps": max(5, args.steps // 5),
        "warmup": 2, "ms_per_step": head["ms"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic",
        "config": {"workload": f"{args.workload} coordinates and charges, random dipoles and traceless quadrupoles, "
                               f"PME {w['K']}^3 order {w['order']}; direct Ewald on water900 with kmax 10; "
                               "16 water boxes through the torch-sim style model",
                   "path": "torchff.PME(rank) / torchff.Ewald(rank) drop-in ops, autograd backward"},
        "results": res}))


def md_consumer_results(dev):
    """The torch-sim style consumer: 16 water boxes of 20-300 waters in one state per call (the
    batch resident on the device, positions written in place each step), against the same boxes
    one model call each (mixed precision serving engine)."""
    from types import SimpleNamespace
    from torchff.md import BatchedSimModel, NM_PER_ANGSTROM
    from torchff.serving import ServingEngine
    from torchff.workloads import box_length_for_density, water_box
    res = {}
    rng = np.random.default_rng(5)
    sizes = [int(x) for x in np.exp(rng.uniform(np.log(20), np.log(300), 16))]
    ws = [water_box(nw, box_length_for_density(nw), seed=100 + k) for k, nw in enumerate(sizes)]
    eng = ServingEngine(slots=2, max_systems=16, max_atoms=16384, max_grid_points=1 << 18, max_k=32,
                        precision="mixed", device=dev)

    def state_of(sel, engine=eng):
        P = np.concatenate([ws[k]["coords"] for k in sel]) / NM_PER_ANGSTROM
        idx = np.concatenate([np.full(ws[k]["natoms"], j) for j, k in enumerate(sel)])
        cells = np.stack([np.asarray(ws[k]["box"]).T / NM_PER_ANGSTROM for k in sel])
        par = [np.concatenate([np.asarray(ws[k][f]) for k in sel]) for f in ("charges", "sigma", "epsilon")]
        mol = np.concatenate([np.arange(ws[k]["natoms"]) // 3 for k in sel])
        st = SimpleNamespace(positions=torch.tensor(P, device=dev), cell=torch.tensor(cells, device=dev),
                             system_idx=torch.tensor(idx, device=dev))
        return st, BatchedSimModel(engine, *par, mol)

    st_all, m_all = state_of(list(range(16)))
    # one small engine per single-system model (each keeps its batch resident in its own slot)
    singles = [state_of([k], ServingEngine(slots=1, max_systems=1, max_atoms=1024, max_grid_points=1 << 15,
                                           max_k=32, precision="mixed", device=dev)) for k in range(16)]
    sec_all, _ = time_steps(lambda: m_all(st_all), 20, 3)
    sec_req, _ = time_steps(lambda: m_all.forward_via_requests(st_all), 10, 2)
    sec_one, _ = time_steps(lambda: [m(st) for st, m in singles], 5, 1)
    res["md_batched_16_systems"] = {"ms": sec_all * 1e3, "systems": 16, "waters": sizes,
                                    "request_path_ms": sec_req * 1e3, "one_call_per_system_ms": sec_one * 1e3}
    return res


def serving_results(args, dev):
    """Request serving: args.requests water boxes of 20-1000 waters (log-uniform, 1 g/cm^3, rc 0.9 nm,
    PME spacing <= 0.1 nm rounded to 2^a 3^b 5^c, order 5, mixed precision) arriving as a Poisson
    stream at args.rate per second.  The same stream is served by continuous batching (graph-replayed
    batches on two slots) and one system per launch (same engine, one system per batch); then both
    again with every request available at once (saturation throughput).  Latency = completion on
    the host (energies and forces copied back) minus arrival."""
    from torchff.serving import ServingEngine, latency_stats, request_stream

    modes = {"batched": dict(slots=2, max_systems=64, max_atoms=1 << 16, max_grid_points=1 << 21),
             "one_per_launch": dict(slots=2, max_systems=1, max_atoms=3072, max_grid_points=1 << 15)}
    # the stated rate, ten times that rate, and everything queued at t = 0 (saturation)
    loads = {"poisson": args.rate, "poisson_x10": 10.0 * args.rate, "saturated": None}
    res, results, hbm = {}, {}, {}
    for name, caps in modes.items():
        eng = ServingEngine(max_k=32, precision="mixed", device=dev, **caps)
        hbm[name] = eng.hbm_bytes
        eng.evaluate(request_stream(64, seed=3))  # warm-up
        for load, rate in loads.items():
            reqs = request_stream(args.requests, rate=rate or args.rate, seed=11)
            done, nb = eng.serve(reqs, realtime=rate is not None)
            st = latency_stats(done)
            st["offered_req_per_s"] = rate
            st["batches"] = nb
            st["systems_per_batch"] = len(done) / nb
            res.setdefault(load, {})[name] = st
            if load == "poisson":
                results[name] = {r.rid: (r.energies, r.forces) for r in done}
    same = all(np.array_equal(results["batched"][k][0], results["one_per_launch"][k][0]) and
               np.array_equal(results["batched"][k][1], results["one_per_launch"][k][1]) for k in results["batched"])
    return {
        "workload": f"{args.requests} water boxes, 20-1000 waters log-uniform at 1 g/cm^3, Poisson arrivals at "
                    f"{args.rate:.0f} req/s, rc 0.9 nm, PME spacing <= 0.1 nm (2^a 3^b 5^c), order 5, mixed "
                    "precision, all periodic images in range",
        "latency_is": "host-side completion (energies and forces copied back) minus arrival",
        **res,
        "p99_ratio_one_per_launch_over_batched": {
            k: res[k]["one_per_launch"]["p99_ms"] / res[k]["batched"]["p99_ms"] for k in ("poisson", "poisson_x10")},
        "saturated_speedup_vs_one_per_launch": res["saturated"]["batched"]["evals_per_s"]
        / res["saturated"]["one_per_launch"]["evals_per_s"],
        "results_identical_batched_vs_alone": bool(same),
        "slot_hbm_bytes": hbm,
    }


def serving_sizes(args, dev):
    """Single-system batches of the listed water counts (20 each after 3 warm-up): device time of
    the upload + launches + read-back per batch from events on the slot's stream."""
    from torchff.serving import ServingEngine, water_request
    eng = ServingEngine(slots=1, max_systems=64, max_atoms=1 << 16, max_grid_points=1 << 21, max_k=32, device=dev)
    s = eng.slots[0]
    out = {}
    for nw in (int(x) for x in args.serving_sizes.split(",")):
        dev_ms = []
        for rep in range(23):
            r = water_request(0, nw, 1)
            assert s.try_add(r)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s.stream)
            s.launch()
            e1.record(s.stream)
            e1.synchronize()
            s.finish(0.0)
            if rep >= 3:
                dev_ms.append(e0.elapsed_time(e1))
        out[f"{nw}_waters"] = {"atoms": 3 * nw, "grid": r.grid[0], "device_ms_median": float(np.median(dev_ms))}
    print(json.dumps({"serving_single_system_batches": out}))


def serving_bench(args, dev):
    r = serving_results(args, dev)
    sat = r["saturated"]["batched"]["evals_per_s"]
    print(json.dumps({
        "metric": f"serving: p50/p99 latency and evals/s, {args.requests} Poisson requests at {args.rate:.0f} req/s, "
                  "20-1000 waters",
        "value": sat, "unit": "evals/s", "n_gpus": 1, "steps": args.requests, "warmup": 64,
        "ms_per_step": 1e3 / sat, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+i64", "data": "synthetic",
        "config": {"workload": r["workload"], "path": "tff_serve_energy_forces, continuous batching over 2 "
                   "graph-replayed slots", "value_is": "saturation throughput of the batched engine (all "
                   "requests queued at t=0)"},
        **{k: v for k, v in r.items() if k != "workload"},
    }))


def pairs_within(w, dev, r):
    """Number of non-excluded pairs within r (device neighbour list, float64 predicate)."""
    import torchff
    x = torch.tensor(w["coords"], dtype=torch.float64, device=dev)
    box = torch.tensor(w["box"], dtype=torch.float64, device=dev)
    nl = torchff.NeighborList(w["natoms"], exclusions=torch.tensor(w["exclusions"], device=dev),
                              algorithm="cell_list")
    cap = int(w["natoms"] * 4.0 / 3.0 * math.pi * r ** 3 * 100.0 / 2.0 * 1.3) + 1024
    _, n = nl(x, box, r, cap, padding=True)
    return int(n.item())


def cluster_pair_results(w, dev, reps=20):
    """The nonbonded forces op (shifted Coulomb + LJ, float32 as the reference's MD loop uses it) on a
    cluster-pair list against the same op on a (P, 2) int64 pair list from the cell list, both lists
    at the cutoff with the intramolecular exclusions: list build and force times, list bytes."""
    import torchff
    dt = torch.float32
    rc = w["cutoff"]
    x = torch.tensor(w["coords"], dtype=dt, device=dev)
    box = torch.tensor(w["box"], dtype=dt, device=dev)
    q, sig, eps = (torch.tensor(w[k], dtype=dt, device=dev) for k in ("charges", "sigma", "epsilon"))
    excl = torch.tensor(w["exclusions"], device=dev)
    k = 138.935456
    f = torch.zeros_like(x)
    nbl = None

    def build_cp():
        nonlocal nbl
        nbl = torchff.build_cluster_pairs(x, box, rc, excl, 0.6)
        return nbl

    t_cp_build, _ = time_steps(build_cp, 3, 1)
    t_cp_forces, _ = time_steps(
        lambda: torchff.compute_nonbonded_forces_from_cluster_pairs(x, box, sig, eps, q, k, rc, *nbl, f), reps, 2)
    nl = torchff.NeighborList(w["natoms"], exclusions=excl, algorithm="cell_list")
    cap = int(w["natoms"] * 4.0 / 3.0 * math.pi * rc ** 3 * 100.0 / 2.0 * 1.3) + 1024
    pairs = None

    def build_ap():
        nonlocal pairs
        pairs, npairs = nl(x, box, rc, cap, padding=True)
        return pairs

    t_ap_build, _ = time_steps(build_ap, 3, 1)
    t_ap_forces, _ = time_steps(
        lambda: torchff.compute_nonbonded_forces_from_atom_pairs(x, pairs, box, sig, eps, q, k, rc, f), reps, 2)
    E, M = nbl[1].shape[1], nbl[3].shape[0]
    cp_bytes = 4 * w["natoms"] + 8 * E + 128 * E + 4 * M + 128 * M
    npairs = pairs_within(w, dev, rc)
    return {"note": "float32, cutoff list without skin; build times include the host read-back of the sizes",
            "cluster_pairs": {"build_ms": t_cp_build * 1e3, "forces_ms": t_cp_forces * 1e3, "list_bytes": cp_bytes,
                              "exclusion_blocks": E, "rows": M},
            "atom_pairs": {"build_ms": t_ap_build * 1e3, "forces_ms": t_ap_forces * 1e3,
                           "list_bytes": 16 * npairs, "pairs": npairs}}


def rms_rel(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    if args.calibrate_copy > 0:
        calibrate_copy(args.calibrate_copy, args.steps)
        return
    from torchff.workloads import make_workload
    dev = "cuda:0"
    if args.serving:
        serving_bench(args, dev)
        return
    if args.serving_sizes:
        serving_sizes(args, dev)
        return
    only = set(args.only.split(",")) if args.only else {"mixed", "fp64", "dropin"}
    w = make_workload(args.workload)
    if args.differentiation:
        differentiation_bench(args, w, dev)
        return
    if args.widened:
        widened_bench(args, w, dev)
        return
    if args.md:
        print(json.dumps({"md_consumer": md_consumer_results(dev)}))
        return
    res = {}
    outs = {}
    if "mixed" in only:
        f = Fused(w, dev, "mixed")
        sec, out = time_steps(f.step, args.steps, args.warmup)
        f.m.check()
        tiles = f.m.tiles()
        res["mixed"] = sec
        outs["mixed"] = (float(out[0]), out[1].double().cpu().numpy())
        kern = kernel_times(f.step, args.steps,
                            ["fused_cluster_pairs", "fused_tiles", "fused_spread", "fused_fft", "fused_gather"])
    if "fp64" in only:
        f64 = Fused(w, dev, "fp64")
        sec, out = time_steps(f64.step, max(3, args.steps // 2), 2)
        res["fp64"] = sec
        outs["fp64"] = (float(out[0]), out[1].double().cpu().numpy())
    if "dropin" in only:
        d = DropIn(w, dev)
        sec, out = time_steps(d.step, args.dropin_steps, 1)
        res["dropin"] = sec
        outs["dropin"] = (float(out[0]), out[1].double().cpu().numpy())
    head = "mixed" if "mixed" in res else sorted(res)[0]
    sec = res[head]

    p_in = pairs_within(w, dev, w["cutoff"])
    roof = None
    if head == "mixed":
        t = kern["fused_tiles"]["avg_ms"]
        achieved = p_in * PAIR_FLOPS / (t * 1e-3) / 1e12
        traffic, traffic_info = tile_traffic()
        roof = {"kernel": "fused_tiles", "bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS, "traffic": traffic,
                "traffic_source": traffic_info, "avg_ms": t,
                "units": f"{p_in} in-cutoff pairs x {PAIR_FLOPS} FLOP per launch",
                "tiles": tiles, "pair_slot_efficiency": p_in / (tiles * 1024.0)}
    na = w["natoms"]
    size = f"{na // 1000}k-atom" if na >= 10000 else f"{na}-atom"
    out = {
        "metric": f"energy+force evals/s, {size} PME water",
        "value": 1.0 / sec,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+i64" if head == "mixed" else ("f64" if head in ("fp64", "dropin") else head),
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {w['natoms']}-atom rigid TIP3P box L={w['L']} nm, "
                               f"rc={w['cutoff']} nm, list {w['cutoff'] + w['skin']} nm rebuilt per call, "
                               f"PME {w['K']}^3 order {w['order']}, alpha={w['alpha']:.5f}",
                   "path": {"mixed": "FusedNonbondedPME float32 arithmetic, int64 fixed-point accumulation",
                            "fp64": "FusedNonbondedPME float64",
                            "dropin": "drop-in modules float64"}[head],
                   "pairs_in_cutoff": p_in},
        "paths_evals_per_s": {k: 1.0 / v for k, v in res.items()},
    }
    if "mixed" in res and "dropin" in res:
        out["speedup_vs_fp64_dropin"] = res["dropin"] / res["mixed"]
    if "mixed" in res and "fp64" in res:
        out["speedup_mixed_vs_fused_fp64"] = res["fp64"] / res["mixed"]
        # the same seeded input through the float64 GPU path (sizes where the CPU oracle is slow)
        out["mixed_vs_fused_fp64"] = {"energy_abs": abs(outs["mixed"][0] - outs["fp64"][0]),
                                      "force_rms_rel": rms_rel(outs["mixed"][1], outs["fp64"][1])}
    if head == "mixed":
        # the reciprocal-space stages once more, alone on the caller's stream (same results)
        from torchff import _lib as _L
        _L.load().tff_fused_set_overlap(0)
        try:
            alone = kernel_times(f.step, max(5, args.steps // 5), ["fused_spread", "fused_fft", "fused_gather"])
        finally:
            _L.load().tff_fused_set_overlap(1)
        out["kernels_ms"] = {k: v["avg_ms"] for k, v in kern.items()}
        out["roofline"] = roof
        sb = pme_stage_bytes(w)
        out["pme_stages"] = {
            "note": "reciprocal space runs on a second stream concurrently with the pair search and pair "
                    "kernel (from 300 000 atoms on, the spread runs first on the caller's stream and only "
                    "the FFT and the gather run beside the pair kernel): stage times are wall times under "
                    "that sharing (bandwidths are lower bounds); 'alone': the same stages run on the "
                    "caller's stream, one after another",
            **{st: {"bytes": b, "ms": kern[f"fused_{st}"]["avg_ms"],
                    "GB/s": b / (kern[f"fused_{st}"]["avg_ms"] * 1e-3) / 1e9,
                    "frac_of_hbm_peak": b / (kern[f"fused_{st}"]["avg_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS}
               for st, b in sb.items()},
            "alone": {st: {"ms": alone[f"fused_{st}"]["avg_ms"],
                           "GB/s": b / (alone[f"fused_{st}"]["avg_ms"] * 1e-3) / 1e9,
                           "frac_of_hbm_peak": b / (alone[f"fused_{st}"]["avg_ms"] * 1e-3) / 1e9 / HBM_PEAK_GBS}
                      for st, b in sb.items()}}
    if head == "mixed" and not args.no_serving:
        out["serving"] = serving_results(args, dev)
    if head == "mixed":
        out["nonbonded_list_formats"] = cluster_pair_results(w, dev)
    if not args.no_cpu_baseline:
        from oracle import composite
        threads = os.cpu_count() or 1
        torch.set_num_threads(threads)
        # one warm-up evaluation (also the error reference), then the median of cpu_samples
        E_ref, F_ref, terms, _ = composite.energy_forces(w)
        times = []
        for _ in range(args.cpu_samples):
            t0 = time.perf_counter()
            composite.energy_forces(w)
            times.append(time.perf_counter() - t0)
        dt = float(np.median(times))
        out["cpu_baseline"] = {
            "value": 1.0 / dt, "unit": "evals/s", "cores": threads, "kind": "port",
            "samples": len(times), "sample_seconds": times,
            "sample": f"median of {len(times)} full float64 energy+force evaluations (after one warm-up) of the "
                      f"same {w['natoms']}-atom box by the CPU oracle (numpy cell search + torch autograd real "
                      f"space + PME), {dt:.1f} s each on {threads} threads of "
                      f"{platform.processor() or platform.machine()}"}
        scale = sum(abs(v) for v in terms.values())
        out["errors_vs_fp64_oracle"] = {
            k: {"energy_abs": abs(v[0] - E_ref), "energy_rel_to_sum_abs_terms": abs(v[0] - E_ref) / scale,
                "force_rms_rel": rms_rel(v[1], F_ref)} for k, v in outs.items()}
    print(json.dumps(out))


if __name__ == "__main__":
    main()
