"""Newton for systems too large for the fused dense kernel: GPU stamping,
supernodal LU and a GPU update kernel, driven iteration by iteration.

Semantics follow voltax/analysis.py:208-266 (damped Newton, gmin stepping)
and :434-480 (transient); a batch of B instances of one pattern runs in
lockstep, each with its own convergence flag.  All arithmetic is in HIP
kernels; the host only sequences launches and reads one flag per instance
per iteration.
"""

from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import engine, splu
from ._lib import check, lib, ptr, stream_ptr


class SparseNewton:
    """Reusable device buffers for B instances of one circuit pattern."""

    def __init__(self, circuit, batch=1, params=None, options=None, leaf=None,
                 use_graph=True):
        from .analysis import Options

        engine.require_gpu()
        self.circuit = circuit
        self.options = options or Options()
        self.fp32 = self.options.precision == "mixed"
        self.compact = self.options.grouping == "compact"
        self.topo, row = engine.topology_of(circuit)
        mid = None
        if leaf == "auto":
            from .autotune import tune_leaf

            (leaf, mid), self.tuning = tune_leaf(circuit, batch, self.options.precision,
                                                 compact=self.compact, table=(self.topo, row))
        self.leaf = leaf or splu.LEAF_SIZE
        self.plan = splu.plan_of(circuit, self.leaf, self.compact, topo=self.topo)
        if mid is not None:
            self.plan.configure(mid)
        self.B = batch
        t = self.topo
        if params is None:
            params = np.tile(row, batch) if len(row) else np.zeros(1)
        self.params = torch.as_tensor(params, dtype=torch.float64).cuda().contiguous()
        dev = dict(dtype=torch.float64, device="cuda")
        self.data = torch.empty((batch, t.nnz), **dev)
        self.res = torch.empty((batch, t.S), **dev)
        self.hist = torch.zeros((batch, t.S), **dev)
        self.dz = torch.empty((batch, t.S), **dev)
        self.work = torch.empty(max(batch * t.work, 1), **dev)
        self.swork = torch.empty(batch * (self.plan.work_per + 2 * t.S), **dev)
        self.front = torch.empty(batch * self.plan.front_total,
                                 dtype=torch.float32 if self.fp32 else torch.float64,
                                 device="cuda")
        self.piv = torch.empty(max(batch * self.plan.piv_total, 1), dtype=torch.int32,
                               device="cuda")
        self.done = torch.zeros(batch, dtype=torch.int8, device="cuda")
        self.step_norm = torch.zeros(batch, **dev)
        self.active = torch.ones(batch, dtype=torch.int8, device="cuda")
        self.opts = engine.options_struct(self.options)
        self.refine = 2 if self.fp32 else 0
        # [t, cf, cq, gmin] read by the stamping kernels at run time, so one
        # captured HIP graph of a Newton iteration serves every time step
        self.scal = torch.zeros(5, **dev)
        self.source_scale = 1.0
        self.z = torch.zeros((batch, t.S), **dev)
        self.graph = None
        self.use_graph = use_graph
        # reused LU factors within a time step (Options.reuse, FP64 only); a
        # linear circuit's factors stay valid while (cf, cq, gmin) and the
        # source scale do not change, so they are kept across time steps
        self.reuse = bool(getattr(self.options, "reuse", False)) and not self.fp32
        self.linear = circuit.is_linear
        self._fact_key = None   # (cf, cq, gmin, source scale, params version) of the factors
        self.wnorm = torch.zeros(batch, **dev)
        self.factorizations = 0
        self.refactor = torch.zeros(batch, dtype=torch.int8, device="cuda")
        self.graphs = {}
        # pinned host staging: the per-solve scalars and the running mask go up
        # without a host synchronization (the mask only when it changes), the
        # per-iteration flags come back in one copy
        self._h_scal = torch.empty(5, dtype=torch.float64, pin_memory=True)
        self._h_active = torch.empty(batch, dtype=torch.int8, pin_memory=True)
        self._h_flags = torch.empty((2, batch), dtype=torch.int8, pin_memory=True)
        self._ev_scal = torch.cuda.Event()
        self._ev_active = torch.cuda.Event()
        self._active_now = np.ones(batch, bool)  # the mask self.active holds

    # -- host <-> device flags and scalars
    def _set_scal(self, t, cf, cq, gmin):
        self._ev_scal.synchronize()  # the staging buffer's last upload has run
        self._h_scal.numpy()[:] = (t, cf, cq, gmin, self.source_scale)
        self.scal.copy_(self._h_scal, non_blocking=True)
        self._ev_scal.record()

    def _set_active(self, run):
        if np.array_equal(run, self._active_now):
            return
        self._ev_active.synchronize()
        self._h_active.numpy()[:] = run
        self.active.copy_(self._h_active, non_blocking=True)
        self._ev_active.record()
        self._active_now = run.copy()

    def _flags(self, refactor=False):
        """done (and refactor) of the last iteration, as host bools."""
        h = self._h_flags
        h[0].copy_(self.done, non_blocking=True)
        if refactor:
            h[1].copy_(self.refactor, non_blocking=True)
        torch.cuda.current_stream().synchronize()
        f = h.numpy().astype(bool)
        return (f[0], f[1]) if refactor else f[0]

    # -- single kernels
    def stamp(self, z, t, cf, cq, gmin, hist=None, data=True):
        check(lib.vxm_stamp(self.topo.handle, self.B, ptr(self.params), ptr(z), float(t),
                            float(cf), float(cq), float(gmin), ptr(hist),
                            ptr(self.data) if data else None, ptr(self.res), ptr(self.work),
                            stream_ptr()), "vxm_stamp")

    def linear_solve(self):
        """dz = J^-1 res (the update applies the minus sign)."""
        self._fact_key = None
        s = stream_ptr()
        check(lib.vxm_splu_factor(self.plan.handle, self.B, ptr(self.data), ptr(self.front),
                                  ptr(self.piv), int(self.fp32), s), "vxm_splu_factor")
        if self.refine:
            check(lib.vxm_splu_solve_refined(self.plan.handle, self.B, ptr(self.data),
                                             ptr(self.front), ptr(self.piv), ptr(self.res),
                                             ptr(self.dz), ptr(self.swork), self.refine, 0, 1,
                                             s), "vxm_splu_solve_refined")
        else:
            self.dz.copy_(self.res)
            check(lib.vxm_splu_solve(self.plan.handle, self.B, ptr(self.front), ptr(self.piv),
                                     ptr(self.dz), ptr(self.swork), 0, 0, s), "vxm_splu_solve")

    def _iteration(self, z, hist):
        """One Newton iteration on z [B, S] in place: stamp -> factor -> solve -> update.
        The residual buffer doubles as the solve's right-hand side / result."""
        s = stream_ptr()
        check(lib.vxm_stamp_dev(self.topo.handle, self.B, ptr(self.params), ptr(z),
                                ptr(self.scal), ptr(hist), ptr(self.data), ptr(self.res),
                                ptr(self.work), s), "vxm_stamp_dev")
        if self.refine:
            check(lib.vxm_splu_factor(self.plan.handle, self.B, ptr(self.data), ptr(self.front),
                                      ptr(self.piv), int(self.fp32), s), "vxm_splu_factor")
            check(lib.vxm_splu_solve_refined(self.plan.handle, self.B, ptr(self.data),
                                             ptr(self.front), ptr(self.piv), ptr(self.res),
                                             ptr(self.dz), ptr(self.swork), self.refine, 0, 1,
                                             s), "vxm_splu_solve_refined")
            dz = self.dz
        else:  # factor and solve in one pass of the levels
            check(lib.vxm_splu_factor_solve(self.plan.handle, self.B, ptr(self.data),
                                            ptr(self.front), ptr(self.piv), ptr(self.res),
                                            ptr(self.swork), s), "vxm_splu_factor_solve")
            dz = self.res
        check(lib.vxm_newton_update(self.B, self.topo.S, self.topo.nv, ptr(z), ptr(dz), -1.0,
                                    C.byref(self.opts), ptr(self.active), ptr(self.done),
                                    ptr(self.step_norm), s), "vxm_newton_update")

    def _iteration_reuse(self, z, hist, reused, exact=False):
        """One Newton iteration whose linear solve reuses the last factors
        (residual-only stamping, substitution) or factors afresh; the update
        records the step size and asks for fresh factors when the reused ones
        stop contracting (vxm_newton_update_reuse).  `exact`: the kept factors
        are those of the current Jacobian (a linear circuit), so the update is
        tested like a fresh one."""
        s = stream_ptr()
        check(lib.vxm_stamp_dev(self.topo.handle, self.B, ptr(self.params), ptr(z),
                                ptr(self.scal), ptr(hist), None if reused else ptr(self.data),
                                ptr(self.res), ptr(self.work), s), "vxm_stamp_dev")
        if reused:
            check(lib.vxm_splu_solve(self.plan.handle, self.B, ptr(self.front), ptr(self.piv),
                                     ptr(self.res), ptr(self.swork), 0, 0, s), "vxm_splu_solve")
        else:
            check(lib.vxm_splu_factor_solve(self.plan.handle, self.B, ptr(self.data),
                                            ptr(self.front), ptr(self.piv), ptr(self.res),
                                            ptr(self.swork), s), "vxm_splu_factor_solve")
        check(lib.vxm_newton_update_reuse(self.B, self.topo.S, self.topo.nv, ptr(z),
                                          ptr(self.res), -1.0, C.byref(self.opts),
                                          ptr(self.active), ptr(self.done), ptr(self.wnorm),
                                          ptr(self.refactor), int(reused and not exact), s),
              "vxm_newton_update_reuse")

    def _replay_reuse(self, reused, exact=False):
        """_iteration_reuse on self.z / self.hist through one of three graphs."""
        key = (reused, exact)
        g = self.graphs.get(key)
        if g is None:
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                self._iteration_reuse(self.z, self.hist, reused, exact)  # outside capture
            torch.cuda.current_stream().wait_stream(side)
            torch.cuda.synchronize()
            g = torch.cuda.CUDAGraph()
            with torch.cuda.graph(g):
                self._iteration_reuse(self.z, self.hist, reused, exact)
            self.graphs[key] = g
            return  # the warm-up already performed this iteration
        g.replay()

    def _replay(self):
        """One Newton iteration on self.z with self.hist, through a captured graph."""
        if self.graph is None:
            side = torch.cuda.Stream()
            side.wait_stream(torch.cuda.current_stream())
            with torch.cuda.stream(side):
                self._iteration(self.z, self.hist)  # warm-up launch outside capture
            torch.cuda.current_stream().wait_stream(side)
            torch.cuda.synchronize()
            self.graph = torch.cuda.CUDAGraph()
            with torch.cuda.graph(self.graph):
                self._iteration(self.z, self.hist)
            return  # the warm-up already performed this iteration
        self.graph.replay()

    def newton(self, z, t, cf, cq, gmin, max_steps, hist=None, active=None, contraction=None,
               reuse=False):
        """Runs damped Newton in place on z [B, S]; returns (done, iters) numpy.

        With `contraction` = c, an instance whose Newton step (max |dz| over the
        nodes) fails to shrink to c times the previous one, or turns non-finite,
        is stopped and flagged in `self.failed` (the warm-start fallback test).
        With `reuse`, iterations after the first reuse the LU factors while
        every running instance's updates contract (vxm_newton_update_reuse);
        `self.factorizations` counts the fresh ones."""
        if reuse:
            return self._newton_reuse(z, t, cf, cq, gmin, max_steps, hist, active)
        self._fact_key = None  # the plain iteration overwrites the factors
        B = self.B
        act = np.ones(B, bool) if active is None else np.asarray(active, bool).copy()
        done = np.zeros(B, bool)
        iters = np.zeros(B, np.int32)
        self.failed = np.zeros(B, bool)
        prev = np.full(B, np.inf)
        self._set_scal(t, cf, cq, gmin)
        graph = self.use_graph and z is self.z and (hist is None or hist is self.hist)
        if graph and hist is None:
            self.hist.zero_()
        for _ in range(max_steps):
            run = act & ~done
            if not run.any():
                break
            self._set_active(run)
            if graph:
                self._replay()
            else:
                self._iteration(z, hist)
            iters[run] += 1
            done |= run & self._flags()
            if contraction is not None:
                norm = self.step_norm.cpu().numpy()
                bad = run & ~done & ~(norm <= contraction * prev)
                self.failed |= bad
                act &= ~bad
                prev = np.where(run, norm, prev)
        return done, iters

    def _newton_reuse(self, z, t, cf, cq, gmin, max_steps, hist, active):
        B = self.B
        act = np.ones(B, bool) if active is None else np.asarray(active, bool).copy()
        done = np.zeros(B, bool)
        iters = np.zeros(B, np.int32)
        self.failed = np.zeros(B, bool)
        self._set_scal(t, cf, cq, gmin)
        graph = self.use_graph and z is self.z and (hist is None or hist is self.hist)
        if graph and hist is None:
            self.hist.zero_()
        # the kept factors are this Jacobian's if nothing else factored since,
        # the parameters were not written (torch's in-place version count) and
        # the step size is the same to 1e-12 (grids from linspace differ in
        # the last bits; a Jacobian that close leaves round-off-sized errors)
        key = (cf, cq, gmin, self.source_scale, self.params._version)
        fk = self._fact_key
        exact = (self.linear and fk is not None and fk[0] == cf and fk[2:] == key[2:]
                 and abs(fk[1] - cq) <= 1e-12 * abs(cq))
        factored = exact
        for _ in range(max_steps):
            run = act & ~done
            if not run.any():
                break
            self._set_active(run)
            reused = 1 if factored else 0
            if graph:
                self._replay_reuse(reused, exact)
            else:
                self._iteration_reuse(z, hist, reused, exact)
            self.factorizations += 1 - reused
            if not reused:
                self._fact_key = key
                exact = self.linear
            factored = True
            iters[run] += 1
            fdone, frefactor = self._flags(refactor=True)
            done |= run & fdone
            if (run & frefactor).any():
                factored = False
        return done, iters

    def dc(self, z, t=0.0):
        """gmin stepping (voltax/analysis.py:236-266) in place on z [B, S]."""
        o = self.options
        guess = z.clone()
        if o.gmin_steps == 0:
            return self.newton(z, t, 1.0, 0.0, o.gmin, o.max_steps)
        ok, its = self.newton(z, t, 1.0, 0.0, o.gmin, o.direct_steps)
        need = ~ok
        if need.any():
            idx = torch.from_numpy(np.flatnonzero(need)).cuda()
            z[idx] = guess[idx]
            gmins = np.logspace(np.log10(o.gmin_start), np.log10(o.gmin), o.gmin_steps)
            for g in gmins:
                ok_k, it_k = self.newton(z, t, 1.0, 0.0, g, o.max_steps, active=need)
                ok = np.where(need, ok_k, ok)
                its = its + np.where(need, it_k, 0)
        return ok, its

    def history(self, z_prev, t_prev, h, method):
        """hist = -q(z_prev)/h (+ f(z_prev, t_prev)/2 for trap)."""
        cfh = 0.0 if method == "be" else 0.5
        check(lib.vxm_stamp(self.topo.handle, self.B, ptr(self.params), ptr(z_prev),
                            float(t_prev), cfh, -1.0 / h, 0.0, None, None, ptr(self.hist),
                            ptr(self.work), stream_ptr()), "vxm_stamp")

    def step(self, z, t_prev, t, method, start=None):
        """One time step in place: z (= z_prev on entry) -> z_n.  `start`, if
        given, is the initial guess of the Newton iteration (the history term
        still comes from z_prev); instances that do not converge from it within
        a few iterations are redone from z_prev with the full budget."""
        h = t - t_prev
        self.history(z, t_prev, h, method)
        cf = 1.0 if method == "be" else 0.5
        o = self.options
        if start is None:
            return self.newton(z, t, cf, 1.0 / h, o.gmin, o.max_steps, hist=self.hist,
                               reuse=self.reuse)
        z_prev = z.clone()
        z.copy_(start)
        ok, its = self.newton(z, t, cf, 1.0 / h, o.gmin, min(o.max_steps, PREDICT_STEPS),
                              hist=self.hist, reuse=self.reuse)
        if not ok.all():
            idx = torch.from_numpy(np.flatnonzero(~ok)).cuda()
            z[idx] = z_prev[idx]
            ok2, its2 = self.newton(z, t, cf, 1.0 / h, o.gmin, o.max_steps, hist=self.hist,
                                    active=~ok, reuse=self.reuse)
            ok = ok | ok2
            its = its + its2
        return ok, its


PREDICT_STEPS = 8  # iteration budget of a predicted start


def extrapolate(out, ts, n, order=4):
    """Value at ts[n] of the polynomial through the last min(n, order)
    accepted points out[n-1], out[n-2], ... (Lagrange form): the cubic
    extrapolated Newton start of backward-Euler steps."""
    k = min(n, order)
    tn = ts[n]
    acc = None
    for i in range(k):
        ti = ts[n - 1 - i]
        w = 1.0
        for j in range(k):
            if j != i:
                tj = ts[n - 1 - j]
                w *= (tn - tj) / (ti - tj)
        term = out[n - 1 - i] * w
        acc = term if acc is None else acc + term
    return acc


def dc_sparse(circuit, t=0.0, guess=None, options=None, leaf=None):
    from .analysis import Solution

    sn = SparseNewton(circuit, 1, options=options, leaf=leaf)
    z = torch.as_tensor(circuit.zeros() if guess is None else guess,
                        dtype=torch.float64).cuda().reshape(1, -1).contiguous()
    ok, its = sn.dc(z, t)
    return Solution(np.asarray(t, float), z[0].cpu().numpy(), bool(ok[0]), circuit.layout,
                    its[0])


def transient_sparse(circuit, ts, ic="dc", method="be", options=None, store=True, leaf=None):
    """Transient of one large circuit on the sparse path; leaf="auto" lets the
    autotuner pick the supernode leaf size and front kernels."""
    from .analysis import Options, Solution, _initial

    options = options or Options()
    ts = np.asarray(ts, float)
    sn = SparseNewton(circuit, 1, options=options, leaf=leaf)
    z0, do_dc = _initial(circuit, ic)
    z = sn.z
    z.copy_(torch.as_tensor(z0, dtype=torch.float64).reshape(1, -1))
    conv = np.ones(len(ts), bool)
    iters = np.zeros(len(ts), np.int32)
    if do_dc:
        ok, its = sn.dc(z, ts[0])
        conv[0], iters[0] = ok[0], its[0]
    out = torch.empty((len(ts), circuit.size), dtype=torch.float64, device="cuda")
    out[0] = z[0]
    predict = getattr(options, "predict", False) and method == "be"
    for n in range(1, len(ts)):
        start = extrapolate(out, ts, n) if predict and n >= 2 else None
        ok, its = sn.step(z, ts[n - 1], ts[n], method, start=start)
        conv[n], iters[n] = ok[0], its[0]
        out[n] = z[0]
    sol = Solution(ts, out.cpu().numpy(), conv, circuit.layout, iters)
    sol.factorizations = sn.factorizations if sn.reuse else None  # time steps only
    return sol


def transient_sparse_batch(circuit, params, z0, ts, method="be", options=None, leaf=None):
    """Transients of B instances of one circuit pattern on the sparse path:
    `params` (B, P) parameter rows (per-instance variation), `z0` (B, S)
    initial states, one Newton iteration for all instances per replay (each
    instance converges on its own; converged ones idle).  Options.predict
    (backward Euler) starts each step from the cubic through the last four
    points, Options.reuse keeps a step's LU factors while the updates
    contract, as on the single-instance path.  Returns device tensors
    (z (T, B, S), converged (T, B) bool, iterations (T, B))."""
    from .analysis import Options

    options = options or Options()
    ts = np.asarray(ts, float)
    params = np.asarray(params, float)
    B = params.shape[0]
    sn = SparseNewton(circuit, B, params=params.reshape(-1), options=options, leaf=leaf)
    z = sn.z
    z.copy_(torch.as_tensor(np.asarray(z0, float).reshape(B, -1)))
    out = torch.empty((len(ts), B, circuit.size), dtype=torch.float64, device="cuda")
    conv = torch.ones((len(ts), B), dtype=torch.bool)
    iters = torch.zeros((len(ts), B), dtype=torch.int32)
    out[0] = z
    predict = getattr(options, "predict", False) and method == "be"
    for n in range(1, len(ts)):
        start = extrapolate(out, ts, n) if predict and n >= 2 else None
        ok, its = sn.step(z, ts[n - 1], ts[n], method, start=start)
        conv[n] = torch.as_tensor(np.asarray(ok))
        iters[n] = torch.as_tensor(np.asarray(its, np.int32))
        out[n] = z
    return out, conv, iters
