"""Seeded rigid-TIP3P water boxes: the synthetic workloads of the benchmarks and tests.

The generator the water configurations are defined by: waters on the first
n sites of an n_lat^3 cubic lattice of spacing L / n_lat, oxygen jittered
uniformly by +-0.02 nm, each molecule rotated by a uniform random rotation
(normalised Gaussian quaternion), O-H 0.09572 nm, H-O-H 104.52 degrees.
Charges O -0.834, H +0.417; LJ on O only (sigma 0.31507 nm,
eps 0.6364 kJ/mol), H carries sigma 1.0 nm, eps 0 as in OpenMM's tip3p.xml.
Exclusion sets follow examples/tip3p/md.py:141-158: all 9 ordered intra-water
pairs (including i == i) for the neighbor list, the 3 unique intra-water
pairs for the Ewald exclusion correction.  The same generator, called with
the same seed, feeds the device runs and the CPU checks.
"""

import math

import numpy as np

COULOMB_KJMOL = 138.935456  # examples/tip3p/model.py:23
Q_O, Q_H = -0.834, 0.417
SIGMA_O, EPS_O = 0.31507, 0.6364
SIGMA_H, EPS_H = 1.0, 0.0
R_OH = 0.09572
HOH_DEG = 104.52
WATER_PER_NM3 = 33.4273  # 1 g/cm^3 of H2O: 6.02214076e23 / 18.01528 / 1e21


def _rotations(rng, n):
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def water_box(n_waters, L, n_lat=None, seed=0, jitter=0.02):
    """Return a dict of float64 numpy arrays describing the box (coords (3n,3), box (3,3), params)."""
    if n_lat is None:
        n_lat = int(math.ceil(n_waters ** (1.0 / 3.0) - 1e-9))
    assert n_lat ** 3 >= n_waters
    rng = np.random.default_rng(seed)
    a = L / n_lat
    idx = np.arange(n_waters)
    site = np.stack([idx // (n_lat * n_lat), (idx // n_lat) % n_lat, idx % n_lat], 1).astype(np.float64)
    O = (site + 0.5) * a + rng.uniform(-jitter, jitter, size=(n_waters, 3))
    half = math.radians(HOH_DEG) / 2.0
    h1 = np.array([R_OH * math.cos(half), R_OH * math.sin(half), 0.0])
    h2 = np.array([R_OH * math.cos(half), -R_OH * math.sin(half), 0.0])
    R = _rotations(rng, n_waters)
    H1 = O + R @ h1
    H2 = O + R @ h2
    coords = np.stack([O, H1, H2], 1).reshape(-1, 3)
    n = 3 * n_waters
    q = np.tile([Q_O, Q_H, Q_H], n_waters).astype(np.float64)
    sigma = np.tile([SIGMA_O, SIGMA_H, SIGMA_H], n_waters).astype(np.float64)
    eps = np.tile([EPS_O, EPS_H, EPS_H], n_waters).astype(np.float64)
    types = np.tile([0, 1, 1], n_waters).astype(np.int64)
    base = 3 * np.arange(n_waters)
    ii, jj = np.meshgrid(np.arange(3), np.arange(3), indexing="ij")
    excl = (base[:, None, None, None] + np.stack([ii, jj], -1)[None]).reshape(-1, 2)
    intra = np.stack([np.stack([base, base + 1], 1), np.stack([base, base + 2], 1),
                      np.stack([base + 1, base + 2], 1)], 1).reshape(-1, 2)
    sig_t = np.array([[SIGMA_O, 0.5 * (SIGMA_O + SIGMA_H)], [0.5 * (SIGMA_O + SIGMA_H), SIGMA_H]])
    eps_t = np.array([[EPS_O, math.sqrt(EPS_O * EPS_H)], [math.sqrt(EPS_O * EPS_H), EPS_H]])
    return dict(coords=coords, box=np.eye(3) * L, charges=q, sigma=sigma, epsilon=eps,
                atom_types=types, exclusions=excl.astype(np.int64), pairs_excl=intra.astype(np.int64),
                sigma_table=sig_t, epsilon_table=eps_t, natoms=n, n_waters=n_waters, L=L)


def box_length_for_density(n_waters, per_nm3=WATER_PER_NM3):
    return (n_waters / per_nm3) ** (1.0 / 3.0)


# Named workloads (sizes, cutoffs, PME grids).  alpha follows the direct-space
# tolerance rule alpha = sqrt(-ln(2 tol)) / rc with tol = 5e-4.
WORKLOADS = {
    # 300 waters, 900 atoms, the reference's own order-6 PME
    "water900": dict(n_waters=300, L=2.08, n_lat=7, cutoff=0.9, skin=0.0, K=24, order=6, tol=5e-4, seed=1),
    # 32,000 waters, 96,000 atoms, neighbor list rebuilt each call with a 0.1 nm skin
    "water96k": dict(n_waters=32000, L=9.86, n_lat=32, cutoff=1.0, skin=0.1, K=96, order=5, tol=5e-4, seed=2),
    # 333,000 waters, 999,000 atoms
    "water999k": dict(n_waters=333000, L=21.5, n_lat=70, cutoff=1.2, skin=0.0, K=216, order=5, tol=5e-4, seed=5),
}


def ewald_alpha(cutoff, tol=5e-4):
    """Ewald splitting parameter sqrt(-ln(2 tol)) / rc (nm^-1)."""
    return math.sqrt(-math.log(2.0 * tol)) / cutoff


def make_workload(name, **overrides):
    """Build a named workload; returns the :func:`water_box` dict plus its run parameters."""
    cfg = dict(WORKLOADS[name])
    cfg.update(overrides)
    d = water_box(cfg["n_waters"], cfg["L"], cfg["n_lat"], seed=cfg["seed"])
    d.update(cfg)
    d["alpha"] = ewald_alpha(cfg["cutoff"], cfg["tol"])
    d["name"] = name
    return d


def differentiation_inputs(w, seed=7):
    """Force-matching inputs for a water workload: per-atom charges perturbed by N(0, 0.02) and
    kept neutral per molecule, sigma and epsilon scaled by U(0.95, 1.05), and a fixed target
    force field N(0, 50 kJ/mol/nm).  Returns (charges, sigma, epsilon, target_forces)."""
    rng = np.random.default_rng(seed)
    n = w["natoms"]
    dq = rng.normal(0.0, 0.02, n).reshape(-1, 3)
    dq -= dq.mean(axis=1, keepdims=True)
    q = w["charges"] + dq.reshape(-1)
    sig = w["sigma"] * rng.uniform(0.95, 1.05, n)
    eps = w["epsilon"] * rng.uniform(0.95, 1.05, n)
    target = rng.normal(0.0, 50.0, (n, 3))
    return q, sig, eps, target

