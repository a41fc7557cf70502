"""Energy + force evaluations per second on the 96,000-atom PME water box (one MI355X).

One step = one full evaluation of the TIP3P nonbonded energy and forces of
32,000 rigid waters (L = 9.86 nm): cell-list neighbor list rebuilt at
rc + skin = 1.1 nm with the intramolecular exclusions, fused real-space
erfc-Coulomb + Lennard-Jones with the exclusion correction at rc = 1.0 nm,
order-5 PME on a 96^3 grid (spread, FFTs, influence function, gather), self
energy; forces = -dE/dr.  Inputs are resident on the device before timing.

Prints one JSON line (see the end of main()).  ``--gpus`` above 1 is refused:
this build runs on one GPU.
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
PAIR_FLOPS = 60  # per in-cutoff pair: fixed tally (min-image 9, r^2 5, rsqrt 1, erfc+exp 10, Coulomb 4, LJ 8,
#                 force scalar 8, force vector 3, accumulations 12)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="water96k")
    ap.add_argument("--dtype", default="float64", choices=["float64", "float32"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-atoms", type=int, default=0,
                    help="size of the CPU-baseline sample (0: the full workload)")
    return ap.parse_args()


class DropIn:
    """The reference-shaped module composition (NeighborList + Nonbonded(ewald) + PME), autograd forces."""

    def __init__(self, w, dtype, dev):
        import torchff
        from torchff.workloads import COULOMB_KJMOL
        self.k = COULOMB_KJMOL
        self.w = w
        n = w["natoms"]
        self.cut = w["cutoff"]
        self.rl = w["cutoff"] + w["skin"]
        self.x = torch.tensor(w["coords"], dtype=dtype, device=dev)
        self.box = torch.tensor(w["box"], dtype=dtype, device=dev)
        self.q = torch.tensor(w["charges"], dtype=dtype, device=dev)
        self.sig = torch.tensor(w["sigma"], dtype=dtype, device=dev)
        self.eps = torch.tensor(w["epsilon"], dtype=dtype, device=dev)
        self.excl = torch.tensor(w["pairs_excl"], device=dev)
        self.nl = torchff.NeighborList(n, exclusions=torch.tensor(w["exclusions"], device=dev),
                                       algorithm="cell_list")
        self.cap = int(n * 4.0 / 3.0 * math.pi * self.rl ** 3 * 100.0 / 2.0 * 1.25)
        self.nb = torchff.Nonbonded()
        self.pme = torchff.PME(w["alpha"], w["K"], 0, order=w["order"]).to(device=dev, dtype=dtype)
        self.alpha = w["alpha"]
        self.npairs = None

    def step(self):
        x = self.x.detach().requires_grad_(True)
        pairs, npairs = self.nl(x.detach(), self.box, self.rl, self.cap, padding=True)
        self.npairs = npairs
        e = self.nb(x, pairs, self.box, self.sig, self.eps, self.q, self.k, self.cut, ewald_alpha=self.alpha,
                    pairs_excl=self.excl).sum() + self.pme(x, self.box, self.q) * self.k
        e.backward()
        return e.detach(), -x.grad


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


def cpu_baseline(w, sample_atoms):
    """Time the float64 CPU oracle (the reference's pure-PyTorch algorithms restated) on the host cores."""
    from oracle import composite
    from torchff.workloads import make_workload
    threads = os.cpu_count() or 1
    torch.set_num_threads(threads)
    ws = w
    if sample_atoms and sample_atoms < w["natoms"]:
        ws = make_workload("water96k", n_waters=sample_atoms // 3)
    t0 = time.perf_counter()
    composite.energy_forces(ws)
    dt = time.perf_counter() - t0
    return {"value": 1.0 / dt, "unit": "evals/s", "cores": threads, "kind": "port",
            "sample": f"one float64 energy+force evaluation of {ws['natoms']} atoms "
                      f"(cell-list pairs + autograd real space + PME), {dt:.1f} s on {threads} threads "
                      f"of {platform.processor() or platform.machine()}"}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1"}))
        sys.exit(2)
    from torchff import _lib
    from torchff.workloads import make_workload
    dev = "cuda:0"
    dtype = torch.float64 if args.dtype == "float64" else torch.float32
    w = make_workload(args.workload)
    model = DropIn(w, dtype, dev)
    sec, (E, F) = time_steps(model.step, args.steps, args.warmup)
    npairs = int(model.npairs.item())

    # per-kernel durations over a timed region of the same steps (events on the launch stream)
    _lib.prof_enable(True)
    time_steps(model.step, args.steps, 0)
    _lib.prof_enable(False)
    kern = {}
    for name in ("nblist_cell_pairs", "pair_terms", "pme_spread", "pme_fft", "pme_fft_conv", "pme_gather"):
        tot, cnt = _lib.prof_collect(name)
        if cnt:
            kern[name] = tot / cnt
    # in-cutoff pairs (the algorithmic unit of the pair kernel)
    import torchff  # noqa: F401
    p_in = None
    with torch.no_grad():
        pairs, _ = model.nl(model.x, model.box, model.rl, model.cap, padding=True)
        v = pairs[:npairs]
        d = model.x[v[:, 0]] - model.x[v[:, 1]]
        L = model.box.diagonal()
        d = d - torch.round(d / L) * L
        p_in = int(((d * d).sum(1) <= model.cut ** 2).sum().item())
    dom = max(kern, key=kern.get)
    peak = FP64_PEAK_TFLOPS if dtype == torch.float64 else FP32_PEAK_TFLOPS
    if dom == "pair_terms":
        achieved = p_in * PAIR_FLOPS / (kern[dom] * 1e-3) / 1e12
        roof = {"kernel": dom, "bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                "frac": achieved / peak, "traffic": None, "avg_ms": kern[dom]}
    else:
        nbytes = npairs * 16 + w["natoms"] * 32
        achieved = nbytes / (kern[dom] * 1e-3) / 1e9
        roof = {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None, "avg_ms": kern[dom]}
    out = {
        "metric": "energy+force evals/s, 96k-atom PME water",
        "value": 1.0 / sec,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if dtype == torch.float64 else "f32",
        "data": "synthetic",
        "config": {"workload": f"{args.workload}: {w['natoms']} atoms TIP3P, L={w['L']} nm, rc={w['cutoff']} "
                               f"+ skin {w['skin']}, PME {w['K']}^3 order {w['order']}",
                   "path": "drop-in modules (NeighborList cell_list + Nonbonded ewald + PME), autograd forces",
                   "pairs_listed": npairs, "pairs_in_cutoff": p_in},
        "kernels_ms": kern,
        "roofline": roof,
    }
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(w, args.cpu_sample_atoms)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
