"""Waveform measurements (voltax/measure.py:35-95): linear-interpolated
level crossings, propagation delay and oscillation period, plus the
derivatives of `delay` with respect to the sampled waveforms (used to seed
adjoint gradients).  Host-side post-processing of GPU results."""

from __future__ import annotations

import numpy as np


def _crossings(t, y, level, direction):
    d = np.asarray(y, float) - level
    lo, hi = d[:-1], d[1:]
    up = (lo < 0) & (hi >= 0)
    down = (lo > 0) & (hi <= 0)
    hit = up if direction == "rise" else down if direction == "fall" else (up | down)
    den = np.where(hit, lo - hi, 1.0)
    w = np.where(hit, lo / den, 0.0)
    t = np.asarray(t, float)
    return t[:-1] + w * np.diff(t), hit


def crossings(t, y, level, direction="both"):
    tc, hit = _crossings(t, y, level, direction)
    return np.where(hit, tc, np.nan)


def _kth(t, y, level, direction, k, after):
    tc, hit = _crossings(t, y, level, direction)
    if after is not None:
        hit &= tc > after
    idx = np.nonzero(hit)[0]
    return (tc[idx[k]], int(idx[k])) if len(idx) > k else (np.nan, -1)


def crossing(t, y, level, direction="rise", *, k=0, after=None):
    """Time of the k-th crossing of `level` (nan if none)."""
    return _kth(t, y, level, direction, k, after)[0]


def delay(t, trig, targ, trig_level, *, targ_level=None, trig_direction="rise",
          targ_direction="both", k=0):
    t0 = crossing(t, trig, trig_level, trig_direction, k=k)
    level = trig_level if targ_level is None else targ_level
    return crossing(t, targ, level, targ_direction, after=t0) - t0


def _dcross(t, y, level, i):
    g = np.zeros(len(y))
    if i < 0:
        return g
    a, b = y[i] - level, y[i + 1] - level
    span = t[i + 1] - t[i]
    g[i] = -span * b / (a - b) ** 2
    g[i + 1] = span * a / (a - b) ** 2
    return g


def delay_vjp(t, trig, targ, trig_level, *, targ_level=None, trig_direction="rise",
              targ_direction="both", k=0):
    """(d delay/d trig, d delay/d targ) at the sampled waveforms."""
    t = np.asarray(t, float)
    t0, i0 = _kth(t, trig, trig_level, trig_direction, k, None)
    level = trig_level if targ_level is None else targ_level
    _, i1 = _kth(t, targ, level, targ_direction, 0, t0)
    return -_dcross(t, np.asarray(trig, float), trig_level, i0), _dcross(
        t, np.asarray(targ, float), level, i1)


def period(t, y, level=None, direction="rise", skip=1):
    """Mean period from `direction` crossings after the first `skip`."""
    y = np.asarray(y, float)
    level = 0.5 * (y.min() + y.max()) if level is None else level
    tc, hit = _crossings(t, y, level, direction)
    hit &= np.cumsum(hit) > skip
    n = int(hit.sum())
    if n < 2:
        return np.nan
    sel = tc[hit]
    return (sel.max() - sel.min()) / (n - 1)


def period_vjp(t, y, level, direction="rise", skip=1):
    """d period / d y at the sampled waveform, the level held fixed (the
    period of `period` with an explicit level: the first and last counted
    crossings move with their two samples)."""
    t, y = np.asarray(t, float), np.asarray(y, float)
    _, hit = _crossings(t, y, level, direction)
    hit &= np.cumsum(hit) > skip
    idx = np.flatnonzero(hit)
    if len(idx) < 2:
        return np.zeros(len(y))
    return (_dcross(t, y, level, idx[-1]) - _dcross(t, y, level, idx[0])) / (len(idx) - 1)


def frequency(t, y, **kw):
    return 1.0 / period(t, y, **kw)
