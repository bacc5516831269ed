"""CPU restatement of voltax/measure.py:35-95 (oracle / test infrastructure):
linear-interpolated crossings, ``delay`` and ``period``, plus their
derivatives with respect to the sampled waveform (what ``jax.grad`` of the
reference measure returns)."""

from __future__ import annotations

import numpy as np


def segment_crossings(t, y, level, direction):
    s = np.asarray(y, float) - level
    a, b = s[:-1], s[1:]
    rise = (a < 0) & (b >= 0)
    fall = (a > 0) & (b <= 0)
    hit = {"rise": rise, "fall": fall, "both": rise | fall}[direction]
    denom = np.where(hit, a - b, 1.0)
    frac = np.where(hit, a / denom, 0.0)
    return t[:-1] + frac * (t[1:] - t[:-1]), hit


def crossing(t, y, level, direction="rise", k=0, after=None):
    """Time of the k-th crossing (and the interval index), nan if none."""
    tc, hit = segment_crossings(t, y, level, direction)
    if after is not None:
        hit = hit & (tc > after)
    idx = np.flatnonzero(hit)
    if len(idx) <= k:
        return np.nan, -1
    return tc[idx[k]], idx[k]


def crossing_grad(t, y, level, i):
    """d(crossing time in interval i)/dy: nonzero at samples i and i+1."""
    g = np.zeros(len(y))
    a, b = y[i] - level, y[i + 1] - level
    dt = t[i + 1] - t[i]
    # tc = t_i + a/(a-b) dt
    g[i] = dt * (-b) / (a - b) ** 2
    g[i + 1] = dt * a / (a - b) ** 2
    return g


def delay(t, trig, targ, trig_level, targ_level=None, trig_direction="rise",
          targ_direction="both", k=0):
    t0, _ = crossing(t, trig, trig_level, trig_direction, k=k)
    lvl = trig_level if targ_level is None else targ_level
    t1, _ = crossing(t, targ, lvl, targ_direction, after=t0)
    return t1 - t0


def delay_grads(t, trig, targ, trig_level, targ_level=None, trig_direction="rise",
                targ_direction="both", k=0):
    """(d delay / d trig, d delay / d targ)."""
    t0, i0 = crossing(t, trig, trig_level, trig_direction, k=k)
    lvl = trig_level if targ_level is None else targ_level
    _, i1 = crossing(t, targ, lvl, targ_direction, after=t0)
    return -crossing_grad(t, trig, trig_level, i0), crossing_grad(t, targ, lvl, i1)


def period(t, y, level=None, direction="rise", skip=1):
    y = np.asarray(y, float)
    level = 0.5 * (y.min() + y.max()) if level is None else level
    tc, hit = segment_crossings(t, y, level, direction)
    hit = hit & (np.cumsum(hit) > skip)
    n = int(hit.sum())
    if n < 2:
        return np.nan
    return (tc[hit].max() - tc[hit].min()) / max(n - 1, 1)
