"""The oracle's force-matching gradients (float64 double backward) against central finite
differences of the loss itself (h = 1e-5), on the 900-atom box with the perturbed parameters.
This pins ``oracle.composite.force_matching``, the reference for the GPU backward."""

import numpy as np
import pytest

from oracle import composite
from torchff.workloads import make_workload, differentiation_inputs


@pytest.fixture(scope="module")
def fm():
    w = make_workload("water900")
    q, s, e, T = differentiation_inputs(w)
    pairs = composite.pair_list(w, w["cutoff"] + w.get("skin", 0.0))
    L, F, g = composite.force_matching(w, q, s, e, T, pairs=pairs)
    return w, q, s, e, T, pairs, L, g


def _loss(w, q, s, e, T, pairs, coords=None):
    ww = dict(w)
    if coords is not None:
        ww["coords"] = coords
    return composite.force_matching(ww, q, s, e, T, pairs=pairs)[0]


@pytest.mark.parametrize("which,idx", [("charges", 0), ("charges", 301), ("sigma", 3), ("sigma", 600),
                                       ("epsilon", 9), ("epsilon", 450)])
def test_param_gradients_match_finite_differences(fm, which, idx):
    w, q, s, e, T, pairs, L, g = fm
    h = 1e-5
    args = {"charges": q, "sigma": s, "epsilon": e}
    plus = {k: v.copy() for k, v in args.items()}
    minus = {k: v.copy() for k, v in args.items()}
    plus[which][idx] += h
    minus[which][idx] -= h
    Lp = _loss(w, plus["charges"], plus["sigma"], plus["epsilon"], T, pairs)
    Lm = _loss(w, minus["charges"], minus["sigma"], minus["epsilon"], T, pairs)
    fd = (Lp - Lm) / (2 * h)
    assert abs(fd - g[which][idx]) <= 1e-5 * max(1.0, abs(fd)), (which, idx, fd, g[which][idx])


@pytest.mark.parametrize("atom,axis", [(0, 0), (1, 2), (452, 1)])
def test_coordinate_gradients_match_finite_differences(fm, atom, axis):
    w, q, s, e, T, pairs, L, g = fm
    h = 1e-5
    cp = w["coords"].copy()
    cm = w["coords"].copy()
    cp[atom, axis] += h
    cm[atom, axis] -= h
    fd = (_loss(w, q, s, e, T, pairs, cp) - _loss(w, q, s, e, T, pairs, cm)) / (2 * h)
    assert abs(fd - g["coords"][atom, axis]) <= 1e-5 * max(1.0, abs(fd)), (fd, g["coords"][atom, axis])


def test_hydrogen_epsilon_gradient_is_zero_by_convention(fm):
    # TIP3P hydrogens have eps = 0: every pair with them has eps_i eps_j = 0
    w, q, s, e, T, pairs, L, g = fm
    assert np.all(g["epsilon"][1::3] == 0.0) and np.all(g["epsilon"][2::3] == 0.0)
