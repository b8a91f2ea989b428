"""Manning-field calibration by gradient descent on a gauge misfit.

Each iteration runs the forward model with checkpoints and the adjoint sweep
(``hydro_misfit_forward`` / ``hydro_misfit_backward`` through
:mod:`hydro_amd.adjoint`), then moves the per-cell Manning n against the
gradient with a Barzilai–Borwein step, clamped to a physical range.  The bed
stays fixed.  Hypotheses (different starting fields) are independent
calibrations of the same problem; :func:`calibrate_many` runs them together,
each iteration one scenario-batched forward and adjoint sweep for all
(``hydro_scen_misfit_forward`` / ``hydro_scen_misfit_backward``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field
from typing import List

import numpy as np

from .adjoint import MisfitProblem, run_backward, run_backward_scen, run_forward, run_forward_scen


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
    n: "object"                               # best Manning field seen (torch f64, caller order)
    losses: List[float] = field(default_factory=list)   # misfit of every evaluated field, in order
    best: float = float("inf")                # misfit of n


def misfit_and_gradient(prob: MisfitProblem, n, zbc):
    zg, buf = run_forward(prob, n, zbc)
    loss = float(((zg - prob.z_obs) ** 2).sum())
    g_n, _ = run_backward(prob, buf, zg, n, zbc)
    return loss, g_n


def calibrate(prob: MisfitProblem, n0, zbc, iters: int = 20, n_min: float = 0.005, n_max: float = 0.2,
              first_step: float = 2e-3, max_growth: float = 4.0) -> Calibration:
    """`iters` gradient steps from n0; the first moves the largest entry by
    `first_step`, later ones use the Barzilai-Borwein length
    (dn.dn)/(dn.dg) of the last two iterates (the previous length when that
    is not positive), at most `max_growth` times the previous length.  A
    step that raises the misfit is undone and the length quartered.  The
    result is the best field seen; `losses` lists the misfit of every
    evaluated field."""
    n = n0.clone()
    out = Calibration(n=n)
    prev_n = prev_g = None
    alpha = None
    best_loss, best_n, best_g = None, None, None
    for _ in range(iters):
        loss, g = misfit_and_gradient(prob, n, zbc)
        out.losses.append(loss)
        if best_loss is not None and loss > best_loss:
            # back to the best field with a shorter step along its gradient
            alpha *= 0.25
            n = (best_n - alpha * best_g).clamp(n_min, n_max)
            prev_n = prev_g = None
            continue
        best_loss, best_n, best_g = loss, n.clone(), g
        if alpha is None:
            alpha = first_step / max(float(g.abs().max()), 1e-300)
        elif prev_g is not None:
            dn = n - prev_n
            dg = g - prev_g
            den = float((dn * dg).sum())
            if den > 0:
                alpha = min(float((dn * dn).sum()) / den, max_growth * alpha)
        prev_n, prev_g = n.clone(), g
        n = (n - alpha * g).clamp(n_min, n_max)
    zg, _ = run_forward(prob, n, zbc, keep_checkpoints=False)
    last = float(((zg - prob.z_obs) ** 2).sum())
    out.losses.append(last)
    if last <= best_loss:
        best_loss, best_n = last, n
    out.n = best_n
    out.best = best_loss
    return out


class _Descent:
    """The step rule of :func:`calibrate` for one hypothesis, fed one
    (loss, gradient) at a time; proposes the next field."""

    def __init__(self, n0, n_min, n_max, first_step, max_growth):
        self.n = n0.clone()
        self.out = Calibration(n=self.n)
        self.prev_n = self.prev_g = None
        self.alpha = None
        self.best_loss = self.best_n = self.best_g = None
        self.lim = (n_min, n_max)
        self.first_step, self.max_growth = first_step, max_growth

    def feed(self, loss, g):
        self.out.losses.append(loss)
        if self.best_loss is not None and loss > self.best_loss:
            self.alpha *= 0.25
            self.n = (self.best_n - self.alpha * self.best_g).clamp(*self.lim)
            self.prev_n = self.prev_g = None
            return
        self.best_loss, self.best_n, self.best_g = loss, self.n.clone(), g
        if self.alpha is None:
            self.alpha = self.first_step / max(float(g.abs().max()), 1e-300)
        elif self.prev_g is not None:
            dn = self.n - self.prev_n
            dg = g - self.prev_g
            den = float((dn * dg).sum())
            if den > 0:
                self.alpha = min(float((dn * dn).sum()) / den, self.max_growth * self.alpha)
        self.prev_n, self.prev_g = self.n.clone(), g
        self.n = (self.n - self.alpha * g).clamp(*self.lim)

    def finish(self, last):
        self.out.losses.append(last)
        if last <= self.best_loss:
            self.best_loss, self.best_n = last, self.n
        self.out.n = self.best_n
        self.out.best = self.best_loss
        return self.out


def calibrate_many(prob: MisfitProblem, n0s, zbc, iters: int = 20, n_min: float = 0.005, n_max: float = 0.2,
                   first_step: float = 2e-3, max_growth: float = 4.0) -> List[Calibration]:
    """:func:`calibrate` for S starting fields ``n0s`` [S, C] at once: each
    iteration evaluates every hypothesis's current field in one scenario
    batch (one launch per step and per adjoint step for all of them), then
    each takes its own step by the same rule.  Returns one Calibration per
    hypothesis, the same as :func:`calibrate` on each alone."""
    import torch

    ds = [_Descent(n0s[k], n_min, n_max, first_step, max_growth) for k in range(n0s.shape[0])]
    for _ in range(iters):
        ns = torch.stack([d.n for d in ds]).contiguous()
        zg, buf = run_forward_scen(prob, ns, zbc)
        losses = _losses(prob, zg)
        g_n, _ = run_backward_scen(prob, buf, zg, ns, zbc)
        del buf
        for k, d in enumerate(ds):
            d.feed(losses[k], g_n[k].clone())
    ns = torch.stack([d.n for d in ds]).contiguous()
    zg, _ = run_forward_scen(prob, ns, zbc, keep_checkpoints=False)
    last = _losses(prob, zg)
    return [d.finish(last[k]) for k, d in enumerate(ds)]


def _losses(prob, zg):
    """Each scenario's misfit, summed over a [steps, G] block as the one-field
    loss is (same order, same bits)."""
    return [float(((zg[:, k, :].contiguous() - prob.z_obs) ** 2).sum()) for k in range(zg.shape[1])]
