"""Manning-field calibration by gradient descent on a gauge misfit.

Each iteration runs the forward model with checkpoints and the adjoint sweep
(``hydro_misfit_forward`` / ``hydro_misfit_backward`` through
:mod:`hydro_amd.adjoint`), then moves the per-cell Manning n against the
gradient with a Barzilai–Borwein step, clamped to a physical range.  The bed
stays fixed.  Hypotheses (different starting fields) are independent
calibrations of the same problem.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List

import numpy as np

from .adjoint import MisfitProblem, run_backward, run_forward


def smooth_field(x, y, seed, mean, std, length, n_modes=24):
    """Seeded smooth random field: a sum of random Fourier modes with
    wavelengths around `length`, shifted and scaled to `mean` and `std`."""
    rng = np.random.default_rng(seed)
    f = np.zeros_like(np.asarray(x, dtype=np.float64))
    for _ in range(n_modes):
        k = rng.normal(0, 1.0 / length, 2)
        f += math.sqrt(2.0 / n_modes) * np.cos(k[0] * x + k[1] * y + rng.uniform(0, 2 * math.pi))
    f = (f - f.mean()) / max(f.std(), 1e-30)
    return mean + std * f


@dataclass
class Calibration:
    n: "object"                               # final Manning field (torch f64, caller order)
    losses: List[float] = field(default_factory=list)   # misfit before each update, then the final one


def misfit_and_gradient(prob: MisfitProblem, n, zbc):
    zg, buf = run_forward(prob, n, zbc)
    loss = float(((zg - prob.z_obs) ** 2).sum())
    g_n, _ = run_backward(prob, buf, zg, n, zbc)
    return loss, g_n


def calibrate(prob: MisfitProblem, n0, zbc, iters: int = 20, n_min: float = 0.005, n_max: float = 0.2,
              first_step: float = 2e-3) -> Calibration:
    """`iters` gradient steps from n0; the first moves the largest entry by
    `first_step`, later ones use the Barzilai-Borwein length
    (dn.dn)/(dn.dg) of the last two iterates (the previous length when that
    is not positive)."""
    n = n0.clone()
    out = Calibration(n=n)
    prev_n = prev_g = None
    alpha = None
    for _ in range(iters):
        loss, g = misfit_and_gradient(prob, n, zbc)
        out.losses.append(loss)
        if prev_g is None:
            alpha = first_step / max(float(g.abs().max()), 1e-300)
        else:
            dn = n - prev_n
            dg = g - prev_g
            den = float((dn * dg).sum())
            if den > 0:
                alpha = float((dn * dn).sum()) / den
        prev_n, prev_g = n.clone(), g
        n = (n - alpha * g).clamp(n_min, n_max)
    zg, _ = run_forward(prob, n, zbc, keep_checkpoints=False)
    out.losses.append(float(((zg - prob.z_obs) ** 2).sum()))
    out.n = n
    return out
