"""CPU restatement of voltax/analysis.py (oracle / test infrastructure).

* `Options`                        voltax/analysis.py:50-89
* `newton`                         voltax/analysis.py:208-233
* `gmin_stepping`                  voltax/analysis.py:236-266
* residuals / stamped Jacobians    voltax/analysis.py:325-382
* dense / sparse dispatch          voltax/analysis.py:385-402
* `dc`, `transient`, `ac`          voltax/analysis.py:410-524
* implicit-function adjoint        voltax/analysis.py:269-288 (`_root_bwd`),
  chained through the time steps the way JAX reverses ``lax.scan``
  (`transient_vjp`, `dc_vjp`).

The reference uses ``jax.jacfwd`` for systems below 16 unknowns and the stamped
Jacobian above; both are exact derivatives, so the oracle always stamps.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from . import dual as D
from . import sparse as osp
from .circuit import Circuit, charges, currents


@dataclass(frozen=True)
class Options:
    max_steps: int = 100
    rtol: float = 1e-6
    atol: float = 1e-9
    max_dv: float = 1.0
    gmin: float = 1e-12
    gmin_steps: int = 10
    gmin_start: float = 1e-2
    direct_steps: int = 25
    solver: str = "auto"
    sparse_threshold: int = 200


@dataclass
class Solution:
    t: np.ndarray
    z: np.ndarray
    converged: np.ndarray
    layout: object
    iters: np.ndarray | None = None

    def v(self, node=None):
        if node is None:
            return self.z[..., : self.layout.n_nodes]
        i = self.layout.node(node)
        return np.zeros(self.z.shape[:-1]) if i < 0 else self.z[..., i]

    def i(self, device, which=0):
        return self.z[..., self.layout.internal(device, which)]


class Linear:
    """Stamped Jacobian + dense LAPACK or restated supernodal solve."""

    def __init__(self, c: Circuit, opts: Options):
        self.c = c
        self.st = osp.structure(c)
        use = (c.size >= opts.sparse_threshold if opts.solver == "auto"
               else opts.solver == "sparse")
        self.plan = osp.make_plan(self.st) if use else None

    def matrix(self, z, t, cf, cq, gmin):
        data = osp.jacobian(self.c, z, t, cf, cq, self.st)
        data[self.st.diagonal[: self.c.n_nodes]] += gmin
        return data

    def solve(self, data, b, transpose=False):
        if self.plan is not None:
            return osp.solve(self.plan, osp.factor(self.plan, data), b, transpose)
        A = self.st.to_dense(data)
        return np.linalg.solve(A.T if transpose else A, b)


def _gmin_vec(c, z, gmin):
    out = np.zeros_like(z)
    out[: c.n_nodes] = gmin * z[: c.n_nodes]
    return out


class Problem:
    """One nonlinear root: DC at ``t`` or one BE/trap step from ``z_prev``."""

    def __init__(self, c, kind, t, z_prev=None, t_prev=None):
        self.c, self.kind, self.t = c, kind, t
        self.z_prev, self.t_prev = z_prev, t_prev
        if kind in ("be", "trap"):
            h = t - t_prev
            self.q_prev = c.q(z_prev)
            self.h = h
            self.f_prev = c.f(z_prev, t_prev) if kind == "trap" else None

    @property
    def coefs(self):
        if self.kind == "dc":
            return 1.0, 0
        return (1.0 if self.kind == "be" else 0.5), 1.0 / self.h

    def residual(self, z, gmin):
        c = self.c
        if self.kind == "dc":  # voltax/analysis.py:333-335
            return c.f(z, self.t) + _gmin_vec(c, z, gmin)
        dq = c.q(z) - self.q_prev
        if self.kind == "be":  # :338-342
            return dq / self.h + c.f(z, self.t) + _gmin_vec(c, z, gmin)
        f_mid = 0.5 * (c.f(z, self.t) + self.f_prev)  # :345-350
        return dq / self.h + f_mid + _gmin_vec(c, z, gmin)


def newton(prob, lin, z0, gmin, opts, max_steps=None):
    """Damped Newton (voltax/analysis.py:208-233). Returns (z, ok, iters)."""
    max_steps = opts.max_steps if max_steps is None else max_steps
    cf, cq = prob.coefs
    z, done, k = np.array(z0, float), False, 0
    nv = prob.c.n_nodes
    while k < max_steps and not done:
        r = prob.residual(z, gmin)
        J = lin.matrix(z, prob.t, cf, cq, gmin)
        dz = lin.solve(J, -r)
        dv = np.max(np.abs(dz[:nv]), initial=0.0)
        scale = min(1.0, opts.max_dv / (dv + 1e-300))
        zn = z + scale * dz
        small = np.all(np.abs(dz) <= opts.atol + opts.rtol * np.abs(zn))
        done = bool(small and scale == 1.0 and np.all(np.isfinite(zn)))
        z, k = zn, k + 1
    return z, done, k


def gmin_stepping(prob, lin, guess, opts):
    """voltax/analysis.py:236-266. Returns (z, ok, total_iters)."""
    if opts.gmin_steps == 0:
        return newton(prob, lin, guess, opts.gmin, opts)
    gmins = np.concatenate([[opts.gmin], np.logspace(np.log10(opts.gmin_start),
                                                     np.log10(opts.gmin),
                                                     opts.gmin_steps)])
    steps = [opts.direct_steps] + [opts.max_steps] * opts.gmin_steps
    z, ok, total = np.array(guess, float), False, 0
    for k in range(opts.gmin_steps + 1):
        if k == 1 and ok:
            break
        start = guess if k == 1 else z
        z, ok, it = newton(prob, lin, start, gmins[k], opts, steps[k])
        total += it
    return z, ok, total


def dc(c, t=0.0, guess=None, options=Options(), lin=None):
    lin = lin or Linear(c, options)
    guess = c.zeros() if guess is None else guess
    z, ok, it = gmin_stepping(Problem(c, "dc", t), lin, guess, options)
    return Solution(np.asarray(t), z, np.asarray(ok), c.layout, np.asarray(it))


def transient(c, ts, ic="dc", method="be", options=Options()):
    ts = np.asarray(ts, float)
    lin = Linear(c, options)
    if isinstance(ic, str):
        z0 = dc(c, ts[0], options=options, lin=lin).z if ic == "dc" else c.zeros()
    else:
        z0 = np.asarray(ic, float)
    zs, oks, its = [z0], [True], [0]
    for t_prev, t in zip(ts[:-1], ts[1:]):
        prob = Problem(c, method, t, zs[-1], t_prev)
        z, ok, it = newton(prob, lin, zs[-1], options.gmin, options)
        zs.append(z)
        oks.append(ok)
        its.append(it)
    return Solution(ts, np.array(zs), np.array(oks), c.layout, np.array(its))


def linearize(c, z, t=0.0):
    st = osp.structure(c)
    return (st.to_dense(osp.jacobian(c, z, t, 1.0, 0, st)),
            st.to_dense(osp.jacobian(c, z, t, 0, 1.0, st)))


def ac(c, freqs, op=None, options=Options()):
    """voltax/analysis.py:494-524: solve ``(G + j 2 pi f C) z = -b``."""
    op = dc(c, options=options) if op is None else op
    G, C = linearize(c, op.z, op.t)
    b = c.ac_stimulus()
    return np.stack([np.linalg.solve(G + 2j * np.pi * f * C, -b) for f in freqs])


# =============================================================================
# Adjoint (implicit function theorem, one transposed solve per root)
# =============================================================================


_PROCESS_FIELDS = ("kp", "vth", "n", "lam", "vt", "cox", "cov")
_PROCESS_LEAVES = (("log_kp", "kp", True), ("vth", "vth", False), ("n", "n", False),
                   ("lam", "lam", False), ("vt", "vt", False), ("log_cox", "cox", True),
                   ("log_cov", "cov", True))


def _param_tangent_groups(c):
    """Differentiable leaves of each group, as (name, where, width): the
    per-device parameters, an EKV group's MOSProcess fields (shared by its
    devices; voltax/elements/mosfet.py:26-87) and a source's waveform fields
    (voltax/signals.py:56-189; PWL times / values one tangent per point)."""
    out = {}
    for name, g in c.groups.items():
        leaves = [(pn, "param", 1) for pn in g.params]
        if g.kind == "EKVMOSFET":
            leaves += [(f, "process", 1) for f in _PROCESS_FIELDS]
        if g.kind in ("VoltageSource", "CurrentSource"):
            for f in g.signal.fields:
                arr = np.asarray(g.wave[f])
                leaves.append((f, "wave", arr.shape[1] if arr.ndim == 2 else 1))
        if leaves:
            out[name] = leaves
    return out


class _Process:
    """MOSProcess with Dual fields (the EKV equations read them as attributes)."""

    def __init__(self, base, fields):
        self.__dict__.update(base.__dict__)
        self.__dict__.update(fields)


def param_vjp(c, z, t, cf, cq, lam):
    """``lam^T d(cf*f + cq*q)(z, t)/dp`` per group and leaf: per-device
    parameters and waveform fields as (N,) (PWL times / values (N, n)), an
    EKV group's process leaves under "process" as scalars, in the reference's
    parametrisation (log_kp, log_cox, log_cov)."""
    zext = np.append(z, 0.0)
    lext = np.append(lam, 0.0)
    grads = {}
    for name, leaves in _param_tangent_groups(c).items():
        g = c.groups[name]
        u = c.unknowns(name)
        T, n = g.n_terminals, g.size
        P = sum(w for _, _, w in leaves)
        p, proc, wave = dict(g.params), {}, {}
        k0, cols = 0, {}
        for leaf, where, w in leaves:
            if where == "param":
                p[leaf] = D.Dual.seed(np.broadcast_to(g.params[leaf], (n,)), k0, P)
            elif where == "process":
                proc[leaf] = D.Dual.seed(np.full(n, getattr(g.process, leaf)), k0, P)
            else:
                base = np.asarray(g.wave[leaf], dtype=float)
                if w == 1:
                    wave[leaf] = D.Dual.seed(np.broadcast_to(base, (n,)), k0, P)
                else:
                    d = np.zeros(base.shape + (P,))
                    for j in range(w):
                        d[:, j, k0 + j] = 1.0
                    wave[leaf] = D.Dual(base, d)
            cols[(leaf, where)] = (k0, w)
            k0 += w
        saved = (g.signal, g.wave, g.process)
        if wave:  # the source evaluates its waveform on Dual fields
            sig = g.signal

            class _DualSignal:
                kind, fields = sig.kind, sig.fields
                eval = staticmethod(sig.dual_eval)
            g.signal, g.wave = _DualSignal, wave
        if proc:
            g.process = _Process(g.process, proc)
        v = [zext[u[:, j]] for j in range(T)]
        x = [zext[u[:, T + j]] for j in range(g.n_internal)]
        acc = np.zeros((n, P))

        def add(A, B, coef):
            for off, lst in ((0, A), (T, B)):
                for j, a in enumerate(lst or []):
                    acc[:] += coef * lext[u[:, off + j]][:, None] * D.tan(a, (n, P))

        try:
            if cf != 0:
                add(*currents(g, v, x, t, p), cf)
            if cq != 0 and g.has_charges:
                add(*charges(g, v, x, p), cq)
        finally:
            g.signal, g.wave, g.process = saved
        out = {}
        for leaf, where, w in leaves:
            k, w = cols[(leaf, where)]
            if where != "process":
                out[leaf] = acc[:, k] if w == 1 else acc[:, k:k + w]
        if proc:
            out["process"] = {"gamma": 0.0, "phi": 0.0}
            for lname, fname, logp in _PROCESS_LEAVES:
                k, _ = cols[(fname, "process")]
                gsum = float(acc[:, k].sum())
                out["process"][lname] = gsum * getattr(g.process, fname) if logp else gsum
        grads[name] = out
    return grads


def _axpy(acc, g, s=1.0):
    """acc += s * g over nested {group: {leaf: array | {leaf: scalar}}}."""
    for gn, d in g.items():
        node = acc.setdefault(gn, {})
        for pn, v in d.items():
            if isinstance(v, dict):
                sub = node.setdefault(pn, {})
                for ln, lv in v.items():
                    sub[ln] = sub.get(ln, 0.0) + s * lv
            else:
                node.setdefault(pn, np.zeros_like(v))
                node[pn] += s * v
    return acc


def dc_vjp(c, z, t, g_z, options=Options(), lin=None):
    """Parameter gradient of a loss with ``dL/dz = g_z`` at DC root ``z``."""
    lin = lin or Linear(c, options)
    J = lin.matrix(z, t, 1.0, 0, options.gmin)
    lam = lin.solve(J, g_z, transpose=True)
    return _axpy({}, param_vjp(c, z, t, 1.0, 0, lam), -1.0)


def time_vjp(c, t, cf, lam):
    """``-lam^T d(cf*f)/dt``: time enters the residual through the sources'
    waveforms only (voltax/signals.py), differentiated by a Dual time."""
    lext = np.append(lam, 0.0)
    tt = D.Dual(np.asarray(t, float), np.ones(1))
    out = 0.0
    for name, g in c.groups.items():
        if g.kind not in ("VoltageSource", "CurrentSource") or cf == 0:
            continue
        v = g.signal.dual_eval(g.wave, tt)
        dv = np.broadcast_to(D.tan(v, np.shape(D.val(v)) + (1,))[..., 0], (g.size,))
        u = c.unknowns(name)
        if g.kind == "VoltageSource":  # F = (v_p - v_n) - value
            out -= cf * float(np.sum(lext[u[:, 2]] * -dv))
        else:  # currents (value, -value)
            out -= cf * float(np.sum((lext[u[:, 0]] - lext[u[:, 1]]) * dv))
    return out


def transient_vjp(c, sol, g_z, ic="dc", method="be", options=Options(), wrt_time=False):
    """Parameter gradient for a loss with ``dL/dz_n = g_z[n]`` over a transient
    `sol` computed by `transient` (reverse sweep over the steps).  With
    ``wrt_time`` also the cotangent of the time grid (the step residuals take
    t_prev and t as parameters, voltax/analysis.py:338-350): ``(tree, dts)``."""
    ts, zs = sol.t, sol.z
    lin = Linear(c, options)
    gbar = np.array(g_z, float).copy()
    grads: dict = {}
    dts = np.zeros(len(ts))
    for n in range(len(ts) - 1, 0, -1):
        t, tp = ts[n], ts[n - 1]
        h = t - tp
        cf = 1.0 if method == "be" else 0.5
        J = lin.matrix(zs[n], t, cf, 1.0 / h, options.gmin)
        lam = lin.solve(J, gbar[n], transpose=True)
        # residual: (q(z) - q(zp))/h + cf f(z,t) [+ 0.5 f(zp,tp)] + gmin v
        _axpy(grads, param_vjp(c, zs[n], t, cf, 1.0 / h, lam), -1.0)
        cfp = 0.0 if method == "be" else 0.5
        _axpy(grads, param_vjp(c, zs[n - 1], tp, cfp, -1.0 / h, lam), -1.0)
        st = lin.st
        Jp = osp.jacobian(c, zs[n - 1], tp, cfp, -1.0 / h, st)
        gbar[n - 1] -= st.to_dense(Jp).T @ lam
        if wrt_time:  # d r_n / d t_n = -dq / h^2 + cf df/dt; d / d t_n-1 = dq / h^2 + cfp df/dt
            dq = float(lam @ (c.q(zs[n]) - c.q(zs[n - 1]))) / (h * h)
            dts[n] += dq + time_vjp(c, t, cf, lam)
            dts[n - 1] += -dq + time_vjp(c, tp, cfp, lam)
    if isinstance(ic, str) and ic == "dc":
        _axpy(grads, dc_vjp(c, zs[0], ts[0], gbar[0], options, lin))
        if wrt_time:
            J0 = lin.matrix(zs[0], ts[0], 1.0, 0.0, options.gmin)
            lam0 = lin.solve(J0, gbar[0], transpose=True)
            dts[0] += time_vjp(c, ts[0], 1.0, lam0)
    return (grads, dts) if wrt_time else grads
