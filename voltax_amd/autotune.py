"""Autotuner of the batched sparse path.

Two schedule choices of the supernodal LU are measured rather than fixed:

* the supernode leaf size of the symbolic analysis (largest nested-dissection
  leaf kept as one supernode), which sets the sizes of the frontal matrices,
  hence which factor kernel runs them, and the number of elimination levels
  (one launch each);
* which kernel factors fronts of 33..64 unknowns: the register-resident
  wavefront kernel (one front per wavefront, no barriers, one wave per SIMD)
  or the blocked 256-thread LDS kernel (rank-16 trailing updates).

The best pair depends on the pattern and on the batch size.  Each candidate
runs one factorization plus a forward and a backward solve of the real batch
on the GPU, timed with HIP events on the launch stream; the fastest is cached
per (pattern, batch, precision).

The reference fixes its leaf size (voltax/sparse.py:381, `leaf=16`).
"""

from __future__ import annotations

import numpy as np
import torch

from . import engine, splu
from ._lib import check, lib, ptr, stream_ptr

CANDIDATES = tuple((leaf, mid) for leaf in (8, 16, 32) for mid in (False, True))
MAX_TUNE_BATCH = 2048  # time at most this many instances (cost is linear beyond)
_cache: dict = {}


def _time_leaf(topo, circuit, cand, B, data, fp32, reps, compact):
    leaf, mid = cand
    plan = splu.plan_of(circuit, leaf, compact, topo=topo)
    plan.configure(mid)
    front = torch.empty(B * plan.front_total, dtype=torch.float32 if fp32 else torch.float64,
                        device="cuda")
    piv = torch.empty(max(B * plan.piv_total, 1), dtype=torch.int32, device="cuda")
    work = torch.empty(B * (plan.work_per + 2 * topo.S), dtype=torch.float64, device="cuda")
    x = torch.ones((B, topo.S), dtype=torch.float64, device="cuda")
    s = stream_ptr()

    def once():
        check(lib.vxm_splu_factor(plan.handle, B, ptr(data), ptr(front), ptr(piv), int(fp32), s),
              "vxm_splu_factor")
        check(lib.vxm_splu_solve(plan.handle, B, ptr(front), ptr(piv), ptr(x), ptr(work), 0,
                                 int(fp32), s), "vxm_splu_solve")

    once()  # first use uploads the plan
    start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(reps):
        once()
    end.record()
    end.synchronize()
    return start.elapsed_time(end) / reps


def tune_leaf(circuit, batch, precision="fp64", candidates=CANDIDATES, reps=3, compact=False,
              table=None):
    """((leaf, mid_blocked), {candidate: ms per factor+solve}) for `batch`
    instances of circuit; the winning kernel choice is left configured on the
    plan of the winning leaf size (of the compact grouping if `compact`).
    `table`: the circuit's (topology, parameter row), if the caller holds them."""
    topo, row = table if table is not None else engine.topology_of(circuit)
    key = (id(topo), int(batch), precision, tuple(candidates), bool(compact))
    hit = _cache.get(key)
    if hit is not None and hit[2] is topo:  # topologies are cached per pattern
        return hit[:2]
    B = int(min(batch, MAX_TUNE_BATCH))
    # the circuit's own Jacobian at z = 0 (gmin on the node diagonals keeps it regular)
    params = torch.as_tensor(np.tile(row, B) if len(row) else np.zeros(1),
                             dtype=torch.float64).cuda()
    z = torch.zeros((B, topo.S), dtype=torch.float64, device="cuda")
    data = torch.empty((B, topo.nnz), dtype=torch.float64, device="cuda")
    res = torch.empty((B, topo.S), dtype=torch.float64, device="cuda")
    work = torch.empty(max(B * topo.work, 1), dtype=torch.float64, device="cuda")
    check(lib.vxm_stamp(topo.handle, B, ptr(params), ptr(z), 0.0, 1.0, 0.0, 1e-3, None,
                        ptr(data), ptr(res), ptr(work), stream_ptr()), "vxm_stamp")
    del res, work
    times = {cand: _time_leaf(topo, circuit, cand, B, data, precision == "mixed", reps, compact)
             for cand in candidates}
    best = min(times, key=times.get)
    splu.plan_of(circuit, best[0], compact, topo=topo).configure(best[1])
    _cache[key] = (best, times, topo)
    return best, times
