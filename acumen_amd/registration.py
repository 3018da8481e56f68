"""Gradient-based centre-of-rotation registration.

The reference aligns projections by a pattern search on the mirror error
(preprocessing/projections_to_sinogram_hpc.F90:287-364).  Here the per-angle
detector shifts s_t are fitted by Gauss-Newton steps on the data term
L = 1/2 |shift_s(radon(x)) - b|^2: every angle is an independent 1-D problem
whose gradient g_t = dL/ds_t and curvature
H_t = sum_d (y(d+o) - y(d+o+1))^2 (y = radon(x), o = floor(-s_t)) come from
one batched evaluation on the GPU.  Between the kinks of the linear
interpolation (where o changes) L is exactly quadratic in s_t with second
derivative H_t, so a step -g_t / H_t lands on the minimum of the current
piece; steps are clamped so that an iterate never jumps over more than one
piece at a time.
"""
from __future__ import annotations

from typing import Callable, List, Sequence, Tuple

import numpy as np

EvalFn = Callable[[List[np.ndarray]], Tuple[List[np.ndarray], List[np.ndarray]]]


def gauss_newton_shifts(evaluate: EvalFn, shifts0: Sequence[np.ndarray], iters: int = 10,
                        max_step: float = 1.0, tol: float = 1e-4):
    """Fit per-angle shifts by Gauss-Newton steps.

    ``evaluate(shifts)`` returns ``(g, H)``: per slice, the (T,) gradient of
    the data term in each shift and its (T,) Gauss-Newton curvature.  Each
    angle steps by -g/H (clamped to ``max_step`` bins; angles with H = 0,
    i.e. a flat projection, stay put) and stops moving once its step is
    below ``tol`` bins.  Stops after ``iters`` evaluations or when no angle
    moves.  Returns the shifts (float32) and the number of evaluations."""
    s = [np.asarray(a, np.float32).copy() for a in shifts0]
    active = [np.ones(a.shape, bool) for a in s]
    evals = 0
    while evals < iters and any(a.any() for a in active):
        g, H = evaluate(s)
        evals += 1
        for i in range(len(s)):
            gi = np.asarray(g[i], np.float64)
            hi = np.asarray(H[i], np.float64)
            step = np.where(hi > 0, -gi / np.where(hi > 0, hi, 1.0), 0.0)
            step = np.clip(step, -max_step, max_step)
            step[~active[i]] = 0.0
            active[i] &= np.abs(step) >= tol
            s[i] = (s[i] + step).astype(np.float32)
    return s, evals
