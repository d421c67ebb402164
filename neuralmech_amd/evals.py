"""``cuwave.evals`` names of the FWI drivers (``projects/18_fwi/2D_fwi.py:48-49``,
``9_structural_optimization/fwi_penalization.py:174-175``, ``16_neuralopt/neuralfwi.py:132-133``):
reconstruction quality of an indicator against the truth.  Definitions are this build's (cuwave's
are not available): the relative L2 error and the F1 score of the void class ``x < threshold``."""

from __future__ import annotations

import torch


def _t(a):
    return torch.as_tensor(a.get() if hasattr(a, "get") else a).double()


def l2_error(x, reference) -> float:
    """``|x - reference|_2 / |reference|_2``."""
    x, r = _t(x), _t(reference).to(_t(x).device)
    return float(torch.linalg.norm(x - r) / torch.linalg.norm(r))


def f1_score(x, reference, threshold: float = 0.5) -> float:
    """F1 score of the voids (``value < threshold``) of ``x`` against those of ``reference``."""
    x, r = _t(x), _t(reference).to(_t(x).device)
    p, t = x < threshold, r < threshold
    tp = float((p & t).sum())
    fp = float((p & ~t).sum())
    fn = float((~p & t).sum())
    return 2 * tp / (2 * tp + fp + fn) if tp + fp + fn > 0 else 1.0
