"""The oracle's parameter cotangents for the MOSProcess and waveform leaves
(oracle/analysis.py param_vjp, restating what jax.grad gives for the
reference's circuit pytree: voltax/elements/mosfet.py:26-87,
voltax/signals.py:56-189) pinned by central finite differences of the
oracle's own transient, leaf by leaf, away from waveform break points."""

import numpy as np
import pytest

import circuits as tc
from oracle import analysis as oa
from oracle import circuit as oc

TIGHT = oa.Options(rtol=1e-13, atol=1e-16)


def _chain(kp_n=400e-6, vth_p=0.4, n_n=1.3, lam_p=0.04, vt=oc.VT_300K, cox_n=8.6e-3,
           cov_p=0.3e-9, delay=21.3e-12, rise=29.1e-12):
    nm = oc.MOSProcess.nmos(kp=kp_n, n=n_n, vt=vt, cox=cox_n)
    pm = oc.MOSProcess.pmos(vth=vth_p, lam=lam_p, vt=vt, cov=cov_p)
    b = oc.CircuitBuilder(nmos=nm, pmos=pm)
    return tc.inverter_chain(b, oc.Pulse(0, 1.2, delay, rise, 20e-12, 60e-12, 400e-12))


def _loss(c, ts):
    sol = oa.transient(c, ts, options=TIGHT)
    out = c.node("out")
    w = np.linspace(0.0, 1.0, len(ts))
    return float(np.sum(w * sol.z[:, out])), sol


def _grad(c, ts):
    L, sol = _loss(c, ts)
    g = np.zeros_like(sol.z)
    g[:, c.node("out")] = np.linspace(0.0, 1.0, len(ts))
    return oa.transient_vjp(c, sol, g, options=TIGHT)


TS = np.linspace(0, 120e-12, 61)


@pytest.mark.parametrize("leaf, group, kw, log", [
    ("log_kp", "EKVMOSFET_n", "kp_n", True), ("n", "EKVMOSFET_n", "n_n", False),
    ("vth", "EKVMOSFET_p", "vth_p", False), ("lam", "EKVMOSFET_p", "lam_p", False),
    ("log_cox", "EKVMOSFET_n", "cox_n", True), ("log_cov", "EKVMOSFET_p", "cov_p", True),
])
def test_process_leaves_match_central_differences(leaf, group, kw, log):
    base = dict(kp_n=400e-6, vth_p=0.4, n_n=1.3, lam_p=0.04, cox_n=8.6e-3, cov_p=0.3e-9)
    gr = _grad(_chain(), TS)[group]["process"][leaf]
    x0 = base[kw]
    h = 1e-4 if log else 1e-4 * max(abs(x0), 1.0)

    def at(x):
        return _loss(_chain(**{**base, kw: x}), TS)[0]

    if log:  # d/d log x
        fd = (at(x0 * np.exp(h)) - at(x0 * np.exp(-h))) / (2 * h)
    else:
        fd = (at(x0 + h) - at(x0 - h)) / (2 * h)
    np.testing.assert_allclose(gr, fd, rtol=2e-5, atol=1e-9 * abs(fd) + 1e-15)


def test_vt_is_shared_by_both_polarities():
    """vt enters both groups; its total derivative is the sum of the two
    groups' leaves."""
    g = _grad(_chain(), TS)
    total = g["EKVMOSFET_n"]["process"]["vt"] + g["EKVMOSFET_p"]["process"]["vt"]
    h = 1e-7
    fd = (_loss(_chain(vt=oc.VT_300K + h), TS)[0] - _loss(_chain(vt=oc.VT_300K - h), TS)[0]) / (2 * h)
    np.testing.assert_allclose(total, fd, rtol=2e-5)


@pytest.mark.parametrize("field, kw", [("delay", "delay"), ("rise", "rise")])
def test_pulse_leaves_match_central_differences(field, kw):
    g = _grad(_chain(), TS)
    (name,) = [k for k in g if k.startswith("VoltageSource_Pulse")]
    h = 1e-15
    base = dict(delay=21.3e-12, rise=29.1e-12)  # corners off the time grid
    fd = (_loss(_chain(**{**base, kw: base[kw] + h}), TS)[0]
          - _loss(_chain(**{**base, kw: base[kw] - h}), TS)[0]) / (2 * h)
    np.testing.assert_allclose(g[name][field][0], fd, rtol=1e-5)


def _driven(b, kind, vals):
    """RC node driven by a current source of waveform `kind`."""
    mod = oc
    if kind == "pwl":
        w = mod.PWL(vals["times"], vals["values"], delay=vals["delay"])
    elif kind == "sine":
        w = mod.Sine(offset=1e-3, amplitude=vals["amplitude"], freq=vals["freq"], delay=5e-12,
                     damping=2e9, phase=vals["phase"])
    else:
        w = mod.Step(v0=0.0, v1=vals["v1"], delay=vals["delay"], rise=15e-12)
    b.isource("0", "a", w)
    b.resistor("a", "0", 1e3)
    b.capacitor("a", "0", 20e-15)
    b.resistor("a", "out", 200.0)
    b.capacitor("out", "0", 10e-15)
    return b.build()


@pytest.mark.parametrize("kind, field, index", [
    ("pwl", "values", 1), ("pwl", "times", 1), ("pwl", "delay", None),
    ("sine", "amplitude", None), ("sine", "freq", None), ("sine", "phase", None),
    ("step", "v1", None), ("step", "delay", None),
])
def test_waveform_leaves_match_central_differences(kind, field, index):
    vals = {"pwl": {"times": np.array([0.0, 30e-12, 70e-12, 100e-12]),
                    "values": np.array([0.0, 2e-3, -1e-3, 0.5e-3]), "delay": 3e-12},
            "sine": {"amplitude": 2e-3, "freq": 8e9, "phase": 20.0},
            "step": {"v1": 1e-3, "delay": 12.7e-12}}[kind]  # corners off the grid
    ts = np.linspace(0, 110e-12, 56)
    c = _driven(oc.CircuitBuilder(), kind, vals)
    g = _grad(c, ts)
    (name,) = [k for k in g if k.startswith("CurrentSource")]
    got = g[name][field][0] if index is None else g[name][field][0, index]
    x0 = vals[field] if index is None else vals[field][index]
    h = 1e-6 * max(abs(x0), 1e-12) if field not in ("times", "delay") else 1e-16

    def at(x):
        v = dict(vals)
        if index is None:
            v[field] = x
        else:
            arr = v[field].copy()
            arr[index] = x
            v[field] = arr
        return _loss(_driven(oc.CircuitBuilder(), kind, v), ts)[0]

    fd = (at(x0 + h) - at(x0 - h)) / (2 * h)
    np.testing.assert_allclose(got, fd, rtol=1e-4, atol=1e-6 * abs(fd) + 1e-30)


@pytest.mark.parametrize("method", ["be", "trap"])
def test_time_grid_cotangents_match_central_differences(method):
    """dL/dts of the oracle (charge term of the step residual plus the
    sources' waveform slopes) against central differences in each probed
    time point, on the pulsed chain with a PWL current into its output."""
    def build():
        b = oc.CircuitBuilder()
        tc.inverter_chain(b, oc.Pulse(0, 1.2, 21.3e-12, 29.1e-12, 20e-12, 60e-12, 400e-12))
        b.isource("0", "out", oc.PWL([0.0, 33e-12, 71e-12], [0.0, 2e-5, -1e-5], delay=1e-12))
        return b.build()

    c = build()
    ts = np.linspace(0, 90e-12, 31)
    opts = oa.Options(rtol=1e-13, atol=1e-16)
    w = np.linspace(0.0, 1.0, len(ts))

    def loss(tt):
        sol = oa.transient(c, tt, method=method, options=opts)
        return float(np.sum(w * sol.z[:, c.node("out")])), sol

    _, sol = loss(ts)
    g = np.zeros_like(sol.z)
    g[:, c.node("out")] = w
    _, dts = oa.transient_vjp(c, sol, g, method=method, options=opts, wrt_time=True)
    for k in (0, 5, 11, 30):
        h = 1e-17
        tp, tm = ts.copy(), ts.copy()
        tp[k] += h
        tm[k] -= h
        fd = (loss(tp)[0] - loss(tm)[0]) / (2 * h)
        np.testing.assert_allclose(dts[k], fd, rtol=2e-4, atol=1e-6 * np.abs(dts).max(),
                                   err_msg=f"t[{k}]")
