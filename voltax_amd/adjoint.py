"""Adjoint gradients through DC and transient solutions, on the GPU.

The reference differentiates every Newton root by the implicit function
theorem (voltax/analysis.py:269-288): one transposed solve with the
Jacobian re-stamped at the root, then a VJP of the residual with respect to
the parameters; JAX chains the steps of the transient scan backwards.
Here the same chain is explicit:

    for n = T-1 .. 1:   lam_n = J_n^-T gbar_n
                        grad -= lam_n^T dr_n/dp      (at z_n and z_{n-1})
                        gbar_{n-1} -= (dr_n/dz_{n-1})^T lam_n
    DC start:           lam_0 = J_dc^-T gbar_0,  grad -= lam_0^T df/dp

with stamping, the supernodal LU's transposed solve, and the parameter /
Jacobian-transpose VJP kernels all in HIP.  `transient_vjp(circuit, sol, g)`
returns ``{group: {param: (N,) array}}`` like the circuit pytree gradient
``jax.grad`` gives for the reference.
"""

from __future__ import annotations

import numpy as np
import torch

from . import engine, splu
from ._lib import check, lib, ptr, stream_ptr
from .circuit import KIND_C, KIND_EKV, KIND_R
from .signals import WAVE_PWL

_EKV_PROCESS = (("log_kp", 4, "kp"), ("vth", 5, None), ("n", 6, None), ("lam", 7, None),
                ("vt", 8, None), ("log_cox", 9, "cox"), ("log_cov", 10, "cov"))


def _unpack(circuit, grad_row):
    """Flat gradient row (parameter layout) -> the circuit pytree's gradient:
    {group: {param: (N,)}}, an EKV group's shared MOSProcess leaves under
    "process" (summed over its devices; log_kp / log_cox / log_cov as in
    voltax/elements/mosfet.py:26-87), a source group's waveform leaves by
    field name (voltax/signals.py: value; v0 v1 delay rise; v1 v2 delay rise
    fall width period; offset amplitude freq delay damping phase; times
    values delay repeat)."""
    out = {}
    off = 0
    for name, el in circuit.elements.items():
        if el.size == 0:
            continue
        rows = el.param_rows()
        w = rows.shape[1]
        g = np.asarray(grad_row[off: off + w * el.size]).reshape(el.size, w)
        off += w * el.size
        if getattr(el, "va", None) is not None:  # Verilog-A: d/d(physical row entry)
            out[name] = el.grad_from_rows(g, rows)
        elif el.kind == KIND_R:
            out[name] = {"log_r": g[:, 0].copy()}
        elif el.kind == KIND_C:
            out[name] = {"log_c": g[:, 0].copy()}
        elif el.kind == KIND_EKV:
            proc = {}
            for leaf, col, phys in _EKV_PROCESS:
                scale = getattr(el.process, phys) if phys else 1.0
                proc[leaf] = float(np.sum(g[:, col]) * scale)
            proc["gamma"] = proc["phi"] = 0.0  # Level-1 parameters: not in the EKV model
            out[name] = {"log_w": g[:, 0].copy(), "log_l": g[:, 1].copy(),
                         "dvth": g[:, 2].copy(), "process": proc}
        else:  # sources: the leaves of the waveform
            wave = el.waves[0]
            if wave.code == WAVE_PWL:
                n = int(rows[0, 0])
                out[name] = {"delay": g[:, 1].copy(), "repeat": g[:, 2].copy(),
                             "times": g[:, 3:3 + n].copy(), "values": g[:, 3 + n:3 + 2 * n].copy()}
            else:
                out[name] = {f: g[:, k].copy() for k, f in enumerate(wave.fields)}
    return out


class Adjoint:
    def __init__(self, circuit, options):
        engine.require_gpu()
        self.c = circuit
        self.o = options
        self.topo, row = engine.topology_of(circuit)
        self.plan = splu.plan_of(circuit, compact=options.grouping == "compact")
        dev = dict(dtype=torch.float64, device="cuda")
        self.params = torch.as_tensor(row if len(row) else np.zeros(1), **dev)
        self.grad = torch.zeros_like(self.params)
        self.data = torch.empty(self.topo.nnz, **dev)
        self.work = torch.empty(max(self.topo.work, 1), **dev)
        self.tmp = torch.empty(self.topo.S, **dev)

    def jac(self, z, t, cf, cq, gmin):
        res = torch.empty(self.topo.S, dtype=torch.float64, device="cuda")
        check(lib.vxm_stamp(self.topo.handle, 1, ptr(self.params), ptr(z), float(t), float(cf),
                            float(cq), float(gmin), None, ptr(self.data), ptr(res),
                            ptr(self.work), stream_ptr()), "vxm_stamp")
        return self.data

    def solve_t(self, data, g):
        f = splu.factor(self.plan, data)
        return splu.solve(self.plan, f, g, transpose=True)

    def pvjp(self, z, t, cf, cq, lam, scale=-1.0):
        check(lib.vxm_param_vjp(self.topo.handle, 1, ptr(self.params), ptr(z), float(t),
                                float(cf), float(cq), ptr(lam), float(scale), ptr(self.grad),
                                stream_ptr()), "vxm_param_vjp")

    def residual(self, z, t, cf, cq):
        """cf*f(z, t) + cq*q(z) (no gmin)."""
        res = torch.empty(self.topo.S, dtype=torch.float64, device="cuda")
        check(lib.vxm_stamp(self.topo.handle, 1, ptr(self.params), ptr(z), float(t), float(cf),
                            float(cq), 0.0, None, None, ptr(res), ptr(self.work),
                            stream_ptr()), "vxm_stamp")
        return res

    def tvjp(self, t, cf, lam):
        """-lam^T d(cf*f)/dt at time t (the sources' waveforms)."""
        out = torch.zeros(max(self.topo.D, 1), dtype=torch.float64, device="cuda")
        check(lib.vxm_time_vjp(self.topo.handle, 1, ptr(self.params), float(t), float(cf),
                               ptr(lam), -1.0, ptr(out), stream_ptr()), "vxm_time_vjp")
        return out.sum()

    def jt_vec(self, z, t, cf, cq, lam):
        check(lib.vxm_jt_vec(self.topo.handle, 1, ptr(self.params), ptr(z), float(t), float(cf),
                             float(cq), 0.0, ptr(lam), ptr(self.tmp), ptr(self.work),
                             stream_ptr()), "vxm_jt_vec")
        return self.tmp

    def dc_step(self, z, t, gbar):
        lam = self.solve_t(self.jac(z, t, 1.0, 0.0, self.o.gmin).clone(), gbar)
        self.pvjp(z, t, 1.0, 0.0, lam)
        return lam

    def result(self):
        return _unpack(self.c, self.grad.cpu().numpy())


def dc_vjp(circuit, z, t, g_z, options=None):
    """Parameter gradient of a loss with dL/dz = g_z at the DC root z."""
    from .analysis import Options

    adj = Adjoint(circuit, options or Options())
    zt = torch.as_tensor(np.asarray(z, float)).cuda()
    adj.dc_step(zt, t, torch.as_tensor(np.asarray(g_z, float)).cuda())
    return adj.result()


def transient_vjp(circuit, sol, g_z, ic="dc", method="be", options=None, wrt_time=False):
    """Parameter gradient of a loss with dL/dz_n = g_z[n] over the transient
    `sol` (from `transient` with the same ic / method / options).  With
    ``wrt_time`` also the cotangent of the time grid (the reference's step
    residuals take t_prev and t as parameters, voltax/analysis.py:338-350):
    returns ``(tree, dL/dts)``."""
    from .analysis import Options

    o = options or Options()
    adj = Adjoint(circuit, o)
    ts = np.asarray(sol.t, float)
    zs = torch.as_tensor(np.asarray(sol.z, float)).cuda()
    gbar = torch.as_tensor(np.asarray(g_z, float)).cuda().clone()
    cf = 1.0 if method == "be" else 0.5
    cfp = 0.0 if method == "be" else 0.5
    dts = torch.zeros(len(ts), dtype=torch.float64, device="cuda")
    for n in range(len(ts) - 1, 0, -1):
        t, tp = ts[n], ts[n - 1]
        cq = 1.0 / (t - tp)
        data = adj.jac(zs[n], t, cf, cq, o.gmin).clone()
        lam = adj.solve_t(data, gbar[n])
        adj.pvjp(zs[n], t, cf, cq, lam)
        adj.pvjp(zs[n - 1], tp, cfp, -cq, lam)
        gbar[n - 1] -= adj.jt_vec(zs[n - 1], tp, cfp, -cq, lam)
        if wrt_time:
            # r_n = (q(z_n) - q(z_n-1)) / h + cf f(z_n, t_n) + cfp f(z_n-1, t_n-1)
            dq = torch.dot(lam, adj.residual(zs[n], t, 0.0, 1.0)
                           - adj.residual(zs[n - 1], tp, 0.0, 1.0)) * cq * cq
            dts[n] += dq + adj.tvjp(t, cf, lam)
            dts[n - 1] += -dq + (adj.tvjp(tp, cfp, lam) if cfp else 0.0)
    if isinstance(ic, str) and ic == "dc":
        lam0 = adj.dc_step(zs[0], ts[0], gbar[0])
        if wrt_time:
            dts[0] += adj.tvjp(ts[0], 1.0, lam0)
    if wrt_time:
        return adj.result(), dts.cpu().numpy()
    return adj.result()
