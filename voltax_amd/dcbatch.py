"""DC operating points of many variants of one circuit pattern on the sparse
path: the reference's algorithm (direct Newton from the all-zero guess, gmin
stepping for the instances it fails), the plain method of the warm-start
study (source stepping from the standard all-zero guess) and
the learned warm start (network guess, plain Newton to the same tolerance,
fallback to the plain method for every instance whose Newton step fails to
contract).

All instances of a chunk run in lockstep through `SparseNewton` (batched
stamping, supernodal LU, update kernel); a chunk holds up to `chunk`
variants and the instances left to a fallback run in a chunk of their own.
The reference has neither batching over variants nor source stepping (its DC
homotopy is gmin stepping, voltax/analysis.py:236-266); both are additions on
top of the same damped Newton (voltax/analysis.py:208-233).
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np
import torch

from .sparse_newton import SparseNewton

# Source-stepping schedule: every independent source scaled by s, Newton to
# tolerance at each s, each stage starting from the previous stage's root.
SOURCE_STEPS = (0.25, 0.5, 0.75, 1.0)


@dataclass
class DCResult:
    z: torch.Tensor            # [B, S] FP64 operating points (device)
    converged: np.ndarray      # [B] bool
    iters: np.ndarray          # [B] Newton iterations (all stages, fallback included)
    fallback: np.ndarray       # [B] bool: the warm start was abandoned
    seconds: float             # wall time of the solve (device synchronised)
    stats: dict = field(default_factory=dict)


class DCBatch:
    """Solver state for chunks of up to `chunk` variants of `circuit`."""

    def __init__(self, circuit, chunk, options=None, use_graph=True, leaf="auto"):
        from .analysis import Options

        self.circuit = circuit
        self.options = options or Options()
        self.chunk = chunk
        self.use_graph = use_graph
        self.leaf = leaf  # "auto": measured per batch size (voltax_amd.autotune)
        self._sn = {}
        self.tuned = {}

    def _newton(self, B):
        sn = self._sn.pop(B, None)
        if sn is None:
            sn = SparseNewton(self.circuit, B, options=self.options, use_graph=self.use_graph,
                              leaf=self.leaf)
        tuning = getattr(sn, "tuning", None)
        self.tuned = {B: {"leaf": sn.leaf, "mid_blocked": getattr(sn.plan, "mid_blocked", False),
                          "ms": None if tuning is None else
                          {f"{l}/{'blk' if m else 'wave'}": round(t, 3)
                           for (l, m), t in tuning.items()}}}
        self._sn[B] = sn  # most recently used last; keep the main and a fallback size
        while len(self._sn) > 2:
            self._sn.pop(next(iter(self._sn)))
        return sn

    def _source_stepping(self, sn, steps):
        """Source stepping from z = 0 on all rows of sn.z; returns (ok, iters)."""
        o = self.options
        sn.z.zero_()
        its = np.zeros(sn.B, np.int64)
        ok = np.ones(sn.B, bool)
        for s in steps:
            sn.source_scale = float(s)
            ok_s, it_s = sn.newton(sn.z, 0.0, 1.0, 0.0, o.gmin, o.max_steps, active=ok)
            its += it_s
            ok &= ok_s
        sn.source_scale = 1.0
        return ok, its

    def plain(self, params, steps=SOURCE_STEPS):
        """Source-stepping DC of every row of params [N, P] (device FP64)."""
        N = params.shape[0]
        S = self.circuit.size
        z = torch.empty((N, S), dtype=torch.float64, device="cuda")
        conv = np.zeros(N, bool)
        iters = np.zeros(N, np.int64)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for a in range(0, N, self.chunk):
            b = min(N, a + self.chunk)
            sn = self._newton(b - a)
            sn.params.copy_(params[a:b].reshape(-1))
            ok, its = self._source_stepping(sn, steps)
            z[a:b] = sn.z
            conv[a:b], iters[a:b] = ok, its
        torch.cuda.synchronize()
        return DCResult(z, conv, iters, np.zeros(N, bool), time.perf_counter() - t0)

    def direct(self, params):
        """The reference's DC algorithm on every row of params [N, P] (device
        FP64): damped Newton from the all-zero guess for `direct_steps`
        iterations, and gmin stepping (restarting from the guess) only for the
        instances whose direct attempt failed (voltax/analysis.py:236-266,
        410-431; `SparseNewton.dc`)."""
        N = params.shape[0]
        S = self.circuit.size
        z = torch.empty((N, S), dtype=torch.float64, device="cuda")
        conv = np.zeros(N, bool)
        iters = np.zeros(N, np.int64)
        stepped = np.zeros(N, bool)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for a in range(0, N, self.chunk):
            b = min(N, a + self.chunk)
            sn = self._newton(b - a)
            sn.params.copy_(params[a:b].reshape(-1))
            sn.z.zero_()
            ok, its = sn.dc(sn.z, 0.0)
            z[a:b] = sn.z
            conv[a:b], iters[a:b] = ok, its
            stepped[a:b] = its > self.options.direct_steps
        torch.cuda.synchronize()
        return DCResult(z, conv, iters, stepped, time.perf_counter() - t0,
                        {"gmin_stepped": int(stepped.sum())})

    def warm(self, params, guess_fn, features, max_steps=10, contraction=0.5,
             steps=SOURCE_STEPS):
        """Warm-started DC: z0 = guess_fn(features[a:b]) per chunk, plain Newton
        with the contraction test, then source stepping from the standard guess
        for every failed instance.  The network's time is inside `seconds`."""
        o = self.options
        N = params.shape[0]
        S = self.circuit.size
        z = torch.empty((N, S), dtype=torch.float64, device="cuda")
        conv = np.zeros(N, bool)
        iters = np.zeros(N, np.int64)
        fell = np.zeros(N, bool)
        t_guess = 0.0
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for a in range(0, N, self.chunk):
            b = min(N, a + self.chunk)
            sn = self._newton(b - a)
            sn.params.copy_(params[a:b].reshape(-1))
            tg = time.perf_counter()
            guess_fn(features[a:b], sn.z)
            torch.cuda.synchronize()
            t_guess += time.perf_counter() - tg
            ok, its = sn.newton(sn.z, 0.0, 1.0, 0.0, o.gmin, max_steps, contraction=contraction)
            z[a:b] = sn.z
            conv[a:b], iters[a:b] = ok, its
            fell[a:b] = ~ok
        bad = np.flatnonzero(fell)
        for a in range(0, len(bad), self.chunk):
            idx = bad[a:a + self.chunk]
            it = torch.from_numpy(idx).cuda()
            sn = self._newton(len(idx))
            sn.params.copy_(params.index_select(0, it).reshape(-1))
            ok, its = self._source_stepping(sn, steps)
            z.index_copy_(0, it, sn.z)
            conv[idx] = ok
            iters[idx] += its
        torch.cuda.synchronize()
        return DCResult(z, conv, iters, fell, time.perf_counter() - t0,
                        {"guess_seconds": t_guess})
