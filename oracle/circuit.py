"""CPU restatement of the reference's circuit model (oracle / test infrastructure).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline
leg may import this package; it is the checker, never the product.

Restates, in numpy FP64:

* waveforms ``Constant/Step/Pulse/Sine/PWL``      voltax/signals.py:56-189
* ``Resistor``, ``Capacitor``                      voltax/elements/passive.py:13-59
* ``VoltageSource``, ``CurrentSource``             voltax/elements/sources.py:58-112
* ``MOSProcess``, ``EKVMOSFET``                    voltax/elements/mosfet.py:26-205
* ``CircuitBuilder`` (node order, fusion, names)   voltax/builder.py:39-256
* ``Layout`` / ``Circuit.f`` / ``Circuit.q``       voltax/circuit.py:31-226

Device groups are "struct of arrays": ``N`` devices of one type with
``(N, T)`` terminal node indices (``-1`` = ground) and per-device parameter
arrays.  Physics functions accept numpy arrays or `Dual` numbers so the same
code yields values, Jacobian blocks and parameter derivatives.
"""

from __future__ import annotations

from collections import defaultdict
from dataclasses import dataclass, field

import numpy as np

from . import dual as D
from . import va as OVA

VT_300K = 0.025852  # voltax/elements/semiconductor.py:13
GROUND = frozenset({"0", "gnd", "GND", "ground"})  # voltax/circuit.py:110


# =============================================================================
# Waveforms (voltax/signals.py)
# =============================================================================


class Signal:
    kind = "signal"
    fields: tuple[str, ...] = ()

    def __init__(self, **p):
        self.p = {k: np.asarray(v, dtype=float) for k, v in p.items()}

    def shape_key(self):
        return tuple((k, self.p[k].shape) for k in self.fields)


class Constant(Signal):
    kind, fields = "Constant", ("value",)

    def __init__(self, value=0.0):
        super().__init__(value=value)

    @staticmethod
    def eval(p, t):  # voltax/signals.py:56-66
        return p["value"]

    dual_eval = eval


class Step(Signal):
    kind, fields = "Step", ("v0", "v1", "delay", "rise")

    def __init__(self, v0=0.0, v1=1.0, delay=0.0, rise=1e-12):
        super().__init__(v0=v0, v1=v1, delay=delay, rise=rise)

    @staticmethod
    def eval(p, t):  # voltax/signals.py:80-82
        frac = np.clip((t - p["delay"]) / p["rise"], 0.0, 1.0)
        return p["v0"] + (p["v1"] - p["v0"]) * frac

    @staticmethod
    def dual_eval(p, t):  # eval with Dual fields
        frac = D.clip01((t - p["delay"]) / p["rise"])
        return p["v0"] + (p["v1"] - p["v0"]) * frac


class Pulse(Signal):
    kind = "Pulse"
    fields = ("v1", "v2", "delay", "rise", "fall", "width", "period")

    def __init__(self, v1=0.0, v2=1.0, delay=0.0, rise=1e-9, fall=1e-9, width=1e-6,
                 period=2e-6):
        super().__init__(v1=v1, v2=v2, delay=delay, rise=rise, fall=fall,
                         width=width, period=period)

    @staticmethod
    def eval(p, t):  # voltax/signals.py:108-114
        tc = np.remainder(t - p["delay"], p["period"])  # floor-mod, like jnp `%`
        up = np.clip(tc / (p["rise"] + 1e-30), 0.0, 1.0)
        t_fall_end = p["rise"] + p["width"] + p["fall"]
        down = np.clip((t_fall_end - tc) / (p["fall"] + 1e-30), 0.0, 1.0)
        value = p["v1"] + (p["v2"] - p["v1"]) * np.minimum(up, down)
        return np.where(t < p["delay"], p["v1"], value)

    @staticmethod
    def dual_eval(p, t):  # eval with Dual fields
        tc = D.floor_mod(t - p["delay"], p["period"])
        up = D.clip01(tc / (p["rise"] + 1e-30))
        t_fall_end = p["rise"] + p["width"] + p["fall"]
        down = D.clip01((t_fall_end - tc) / (p["fall"] + 1e-30))
        value = p["v1"] + (p["v2"] - p["v1"]) * D.minimum(up, down)
        return D.select(D.val(t) < D.val(p["delay"]), p["v1"], value)


class Sine(Signal):
    kind = "Sine"
    fields = ("offset", "amplitude", "freq", "delay", "damping", "phase")

    def __init__(self, offset=0.0, amplitude=1.0, freq=1.0, delay=0.0, damping=0.0,
                 phase=0.0):
        super().__init__(offset=offset, amplitude=amplitude, freq=freq, delay=delay,
                         damping=damping, phase=phase)

    @staticmethod
    def eval(p, t):  # voltax/signals.py:143-147
        tau = t - p["delay"]
        wave = np.sin(2 * np.pi * p["freq"] * tau + np.deg2rad(p["phase"]))
        active = np.where(tau >= 0, np.exp(-p["damping"] * tau) * wave, 0.0)
        return p["offset"] + p["amplitude"] * active

    @staticmethod
    def dual_eval(p, t):  # eval with Dual fields
        tau = t - p["delay"]
        wave = D.sin(2 * np.pi * p["freq"] * tau + p["phase"] * (np.pi / 180.0))
        active = D.select(D.val(tau) >= 0, D.exp(-p["damping"] * tau) * wave, 0.0)
        return p["offset"] + p["amplitude"] * active


class PWL(Signal):
    kind, fields = "PWL", ("times", "values", "delay", "repeat")

    def __init__(self, times, values, delay=0.0, repeat=-1.0):
        super().__init__(times=times, values=values, delay=delay, repeat=repeat)

    @staticmethod
    def eval(p, t):  # voltax/signals.py:174-189
        times, values = np.atleast_2d(p["times"]), np.atleast_2d(p["values"])
        n = times.shape[0]
        delay = np.broadcast_to(p["delay"], (n,))
        repeat = np.broadcast_to(p["repeat"], (n,))
        out = np.empty(n)
        for i in range(n):
            tau = t - delay[i]
            t_last = times[i, -1]
            period = t_last - repeat[i]
            safe = period if period > 0 else 1.0
            k = max(np.ceil((tau - t_last) / safe - 1e-9), 0.0)
            if repeat[i] >= 0 and period > 0:
                tau = tau - k * safe
            out[i] = np.interp(tau, times[i], values[i])
        return out

    @staticmethod
    def dual_eval(p, t):
        """eval with Dual fields (times / values of shape (N, n), their
        tangents (N, n, P)): the interpolation restated so the derivative
        follows the segment numpy.interp uses."""
        times, values = p["times"], p["values"]
        n = D.val(times).shape[0]
        outs = []
        for i in range(n):
            ti, vi = D.item(times, i), D.item(values, i)
            delay = D.item(np.broadcast_to(D.val(p["delay"]), (n,)), i) \
                if not isinstance(p["delay"], D.Dual) else D.item(p["delay"], i)
            repeat = D.item(np.broadcast_to(D.val(p["repeat"]), (n,)), i) \
                if not isinstance(p["repeat"], D.Dual) else D.item(p["repeat"], i)
            tau = t - delay
            t_last = D.item(ti, -1)
            period = t_last - repeat
            safe = period if D.val(period) > 0 else 1.0
            k = max(np.ceil((D.val(tau) - D.val(t_last)) / D.val(safe) - 1e-9), 0.0)
            if D.val(repeat) >= 0 and D.val(period) > 0:
                tau = tau - k * safe
            tv, m = D.val(ti), len(D.val(ti))
            if D.val(tau) <= tv[0]:
                outs.append(D.item(vi, 0))
            elif D.val(tau) >= tv[-1]:
                outs.append(D.item(vi, m - 1))
            else:
                j = int(np.searchsorted(tv, D.val(tau), side="right")) - 1
                frac = (tau - D.item(ti, j)) / (D.item(ti, j + 1) - D.item(ti, j))
                outs.append(D.item(vi, j) + frac * (D.item(vi, j + 1) - D.item(vi, j)))
        duals = [o for o in outs if isinstance(o, D.Dual)]
        if not duals:  # every point on a constant end of its curve
            return np.array([D.val(o) for o in outs])
        P = duals[0].d.shape[-1]
        return D.Dual(np.array([D.val(o) for o in outs]),
                      np.stack([o.d if isinstance(o, D.Dual) else np.zeros(P) for o in outs]))


def as_signal(value):
    return value if isinstance(value, Signal) else Constant(value)


# =============================================================================
# Element groups
# =============================================================================


@dataclass
class MOSProcess:
    """voltax/elements/mosfet.py:26-75 (physical units; log-space on use)."""

    kp: float = 400e-6
    vth: float = 0.4
    n: float = 1.3
    lam: float = 0.04
    vt: float = VT_300K
    gamma: float = 0.0
    phi: float = 0.7
    cox: float = 8.6e-3
    cov: float = 0.3e-9

    @classmethod
    def nmos(cls, **kw):
        return cls(**{"kp": 400e-6, **kw})

    @classmethod
    def pmos(cls, **kw):
        return cls(**{"kp": 200e-6, **kw})


# per-kind static description: terminals, internals, per-device parameters
KINDS = {
    "Resistor": (("p", "n"), 0, ("log_r",)),
    "Capacitor": (("p", "n"), 0, ("log_c",)),
    "VoltageSource": (("p", "n"), 1, ()),
    "CurrentSource": (("p", "n"), 0, ()),
    "EKVMOSFET": (("d", "g", "s", "b"), 0, ("log_w", "log_l", "dvth")),
}


@dataclass
class Group:
    kind: str
    nodes: np.ndarray  # (N, T) int, -1 = ground
    params: dict  # name -> (N,) float
    signal: type | None = None  # waveform class for sources
    wave: dict = field(default_factory=dict)  # waveform params, (N,) or (N, P)
    polarity: str = "n"
    process: MOSProcess | None = None
    va: object = None  # oracle.va.Model of a Verilog-A group (kind "VA")

    @property
    def size(self):
        return self.nodes.shape[0]

    @property
    def n_terminals(self):
        if self.kind == "VA":
            return len(self.va.terminals)
        return len(KINDS[self.kind][0])

    @property
    def n_internal(self):
        if self.kind == "VA":
            return len(self.va.internal)
        return KINDS[self.kind][1]

    @property
    def has_charges(self):
        return self.kind in ("Capacitor", "EKVMOSFET", "VA")

    def value(self, t):
        return self.signal.eval(self.wave, t)


def _through(i):
    return [i, -i]


def _soft_abs(x, eps=1e-3):  # voltax/elements/mosfet.py:151-152
    return D.sqrt(x * x + eps * eps) - eps


def _ekv_inversion(g, p, vd, vg, vs, vb):  # voltax/elements/mosfet.py:177-182
    pr = g.process
    vp = (vg - vb - pr.vth - p["dvth"]) / pr.n
    i_f = D.softplus((vp - (vs - vb)) / (2 * pr.vt)) ** 2
    i_r = D.softplus((vp - (vd - vb)) / (2 * pr.vt)) ** 2
    return i_f, i_r


def currents(g: Group, v, x, t, p):
    """``(I, F)`` lists (per terminal / per internal) or None."""
    k = g.kind
    if k == "VA":
        F, _ = OVA.group_rows(g.va, v, x, p)
        return F[: len(v)], F[len(v):]
    if k == "Resistor":  # passive.py:37-38
        return _through((v[0] - v[1]) * D.exp(-p["log_r"])), None
    if k == "Capacitor":
        return None, None
    if k == "VoltageSource":  # sources.py:83-85
        return _through(x[0]), [(v[0] - v[1]) - g.value(t)]
    if k == "CurrentSource":  # sources.py:108-110
        val = g.value(t)
        if not isinstance(val, D.Dual):
            val = np.broadcast_to(val, (g.size,)).astype(float)
        return _through(val), None
    if k == "EKVMOSFET":  # mosfet.py:144-148, 170-175
        s = 1.0 if g.polarity == "n" else -1.0
        vd, vg, vs, vb = (s * vi for vi in v)
        pr = g.process
        i_f, i_r = _ekv_inversion(g, p, vd, vg, vs, vb)
        w, l_ = D.exp(p["log_w"]), D.exp(p["log_l"])
        i_spec = 2 * pr.n * pr.kp * (w / l_) * pr.vt ** 2
        ids = i_spec * (i_f - i_r) * (1.0 + pr.lam * _soft_abs(vd - vs))
        zero = 0.0 * ids
        return [s * ids, s * zero, -s * ids, s * zero], None
    raise KeyError(k)


def charges(g: Group, v, x, p):
    """``(Q, P)`` lists or None."""
    k = g.kind
    if k == "VA":
        _, Q = OVA.group_rows(g.va, v, x, p)
        return Q[: len(v)], Q[len(v):]
    if k == "Capacitor":  # passive.py:58-59
        return _through(D.exp(p["log_c"]) * (v[0] - v[1])), None
    if k == "EKVMOSFET":  # mosfet.py:184-205
        s = 1.0 if g.polarity == "n" else -1.0
        vd, vg, vs, vb = (s * vi for vi in v)
        pr = g.process
        i_f, i_r = _ekv_inversion(g, p, vd, vg, vs, vb)
        xf, xr = D.sqrt(0.25 + i_f), D.sqrt(0.25 + i_r)
        w, l_ = D.exp(p["log_w"]), D.exp(p["log_l"])
        c_ox = pr.cox * w * l_
        scale = c_ox * pr.n * pr.vt
        sm = xf + xr
        q_d = -scale * ((4 / 15) * (3 * xr ** 3 + 6 * xr ** 2 * xf + 4 * xr * xf ** 2
                                    + 2 * xf ** 3) / sm ** 2 - 0.5)
        q_s = -scale * ((4 / 15) * (3 * xf ** 3 + 6 * xf ** 2 * xr + 4 * xf * xr ** 2
                                    + 2 * xr ** 3) / sm ** 2 - 0.5)
        q_i = q_d + q_s
        q_b = -(pr.n - 1) / pr.n * (c_ox * (vg - vb) + q_i)
        q_g = -(q_i + q_b)
        cov = pr.cov * w  # overlap, mosfet.py:127-130
        q_gs, q_gd = cov * (vg - vs), cov * (vg - vd)
        qs = [q_d - q_gd, q_g + q_gs + q_gd, q_s - q_gs, q_b]
        return [s * q for q in qs], None
    return None, None


# =============================================================================
# Layout and circuit (voltax/circuit.py)
# =============================================================================


@dataclass
class Layout:
    node_names: tuple
    groups: tuple  # (name, n, k, offset)
    devices: dict  # device name -> (group, index)

    @property
    def n_nodes(self):
        return len(self.node_names)

    @property
    def size(self):
        return self.n_nodes + sum(n * k for _, n, k, _ in self.groups)

    def node(self, name):
        if name in GROUND:
            return -1
        return self.node_names.index(name)

    def internal(self, device, which=0):
        g, i = self.devices[device]
        for name, n, k, off in self.groups:
            if name == g:
                return off + i * k + which
        raise KeyError(device)


class Circuit:
    def __init__(self, groups: dict, node_names, device_names):
        self.groups = {k: groups[k] for k in sorted(groups)}  # circuit.py:139
        nv = len(node_names)
        off, lay = nv, []
        for name, g in self.groups.items():
            lay.append((name, g.size, g.n_internal, off))
            off += g.size * g.n_internal
        self.layout = Layout(tuple(node_names), tuple(lay), dict(device_names))

    @property
    def n_nodes(self):
        return self.layout.n_nodes

    @property
    def size(self):
        return self.layout.size

    def node(self, name):
        return self.layout.node(name)

    def zeros(self):
        return np.zeros(self.size)

    def state(self, v=None):
        z = self.zeros()
        for k, val in (v or {}).items():
            z[self.node(k)] = val
        return z

    def unknowns(self, name):
        """``(N, T+K)`` state indices of a group's devices (-1 = ground)."""
        g = self.groups[name]
        _, n, k, off = next(x for x in self.layout.groups if x[0] == name)
        internal = np.arange(off, off + n * k).reshape(n, k)
        return np.concatenate([g.nodes, internal], axis=1)

    def set_param(self, device, **values):
        """In-place per-device parameter update in physical units."""
        gname, i = self.layout.devices[device]
        g = self.groups[gname]
        for key, v in values.items():
            if "log_" + key in g.params:
                g.params["log_" + key][i] = np.log(v)
            elif key in g.params:
                g.params[key][i] = v
            elif key == "value" and g.signal is Constant:
                g.wave["value"][i] = v
            else:
                raise AttributeError(key)

    def _assemble(self, z, fn):
        nv, S = self.n_nodes, self.size
        out = np.zeros(S + 1)  # last slot = ground, dropped
        zext = np.append(z, 0.0)
        for name, g in self.groups.items():
            u = self.unknowns(name)
            T = g.n_terminals
            v = [zext[u[:, j]] for j in range(T)]
            x = [zext[u[:, T + j]] for j in range(g.n_internal)]
            A, B = fn(g, v, x)
            for j, val in enumerate(A or []):
                np.add.at(out, np.where(u[:, j] < 0, S, u[:, j]),
                          np.broadcast_to(val, (g.size,)))
            for j, val in enumerate(B or []):
                np.add.at(out, u[:, T + j], np.broadcast_to(val, (g.size,)))
        return out[:S]

    def f(self, z, t=0.0):
        return self._assemble(z, lambda g, v, x: currents(g, v, x, t, g.params))

    def q(self, z):
        return self._assemble(z, lambda g, v, x: charges(g, v, x, g.params))

    def ac_stimulus(self):
        """Complex small-signal excitation (voltax/elements/sources.py:87-112)."""
        S = self.size
        b = np.zeros(S + 1, complex)
        for name, g in self.groups.items():
            if g.kind not in ("VoltageSource", "CurrentSource"):
                continue
            ph = g.wave_ac
            u = self.unknowns(name)
            if g.kind == "VoltageSource":
                np.add.at(b, u[:, 2], -ph)
            else:
                np.add.at(b, np.where(u[:, 0] < 0, S, u[:, 0]), ph)
                np.add.at(b, np.where(u[:, 1] < 0, S, u[:, 1]), -ph)
        return b[:S]


# =============================================================================
# Builder (voltax/builder.py)
# =============================================================================


_PREFIX = {"Resistor": "R", "Capacitor": "C", "VoltageSource": "V",
           "CurrentSource": "I", "EKVMOSFET": "M"}


class CircuitBuilder:
    def __init__(self, nmos: MOSProcess | None = None, pmos: MOSProcess | None = None):
        self.nmos_process = nmos or MOSProcess.nmos()
        self.pmos_process = pmos or MOSProcess.pmos()
        self._devices = []  # (name, kind, terminals, params, extra)
        self._names = set()
        self._nodes: dict[str, None] = {}
        self._counters = defaultdict(int)

    def node(self, name=None):
        if name is None:
            self._counters["_node"] += 1
            name = f"_{self._counters['_node']}"
        self._nodes.setdefault(name, None)
        return name

    def _add(self, kind, terminals, params, name, extra):
        if name is None:
            pre = _PREFIX[kind]
            self._counters[pre] += 1
            name = f"{pre}{self._counters[pre]}"
        if name in self._names:
            raise ValueError(f"duplicate device name {name!r}")
        self._names.add(name)
        terms = tuple(str(n) for n in terminals)
        for n in terms:
            if n not in GROUND:
                self._nodes.setdefault(n, None)
        self._devices.append((name, kind, terms, params, extra))
        return name

    def resistor(self, p, n, r, name=None):
        return self._add("Resistor", (p, n), {"log_r": np.log(r)}, name, {})

    def va(self, model, terminals, name=None, m=1.0, **instance):
        """A Verilog-A device of `model` (oracle.va.Model) on `terminals`;
        real instance parameters are stored as ``log_<name>`` when positive."""
        if name is None:
            self._counters["X"] += 1
            name = f"X{self._counters['X']}"
        bad = set(instance) - set(model.inst)
        if bad:
            raise ValueError(f"not instance parameters: {sorted(bad)}")
        vals = {}
        for k, sk in zip(model.inst, model.stored):
            x = float(instance.get(k, model.card[k]))
            vals[sk] = np.log(x) if sk.startswith("log_") else x
        vals["m"] = float(m)
        return self._add("VA", terminals, vals, name, {"va": model})

    def capacitor(self, p, n, c, name=None):
        return self._add("Capacitor", (p, n), {"log_c": np.log(c)}, name, {})

    def vsource(self, p, n, value=0.0, ac=0.0, ac_phase=0.0, name=None):
        return self._add("VoltageSource", (p, n), {}, name,
                         {"signal": as_signal(value), "ac": (ac, ac_phase)})

    def isource(self, p, n, value=0.0, ac=0.0, ac_phase=0.0, name=None):
        return self._add("CurrentSource", (p, n), {}, name,
                         {"signal": as_signal(value), "ac": (ac, ac_phase)})

    def _mos(self, pol, d, g, s, b, w, l, name, process, dvth):
        proc = process or (self.nmos_process if pol == "n" else self.pmos_process)
        return self._add("EKVMOSFET", (d, g, s, s if b is None else b),
                         {"log_w": np.log(w), "log_l": np.log(l), "dvth": float(dvth)},
                         name, {"polarity": pol, "process": proc})

    def nmos(self, d, g, s, b=None, w=1e-6, l=0.13e-6, name=None, process=None,
             dvth=0.0):
        return self._mos("n", d, g, s, b, w, l, name, process, dvth)

    def pmos(self, d, g, s, b=None, w=1e-6, l=0.13e-6, name=None, process=None,
             dvth=0.0):
        return self._mos("p", d, g, s, b, w, l, name, process, dvth)

    def build(self) -> Circuit:
        index = {n: i for i, n in enumerate(self._nodes)}
        resolve = lambda n: -1 if n in GROUND else index[n]  # noqa: E731
        buckets: dict = {}
        for name, kind, terms, params, extra in self._devices:
            sig = extra.get("signal")
            key = (kind, extra.get("polarity"), id(extra.get("process")),
                   None if sig is None else (sig.kind, sig.shape_key()), id(extra.get("va")))
            buckets.setdefault(key, []).append((name, kind, terms, params, extra))
        groups, device_names, used = {}, {}, defaultdict(int)
        for key, members in buckets.items():
            kind, extra = members[0][1], members[0][4]
            base = kind if kind != "VA" else extra["va"].mod.name
            if kind == "EKVMOSFET":
                base = f"{kind}_{extra['polarity']}"
            elif "signal" in extra and extra["signal"].kind != "Constant":
                base = f"{kind}_{extra['signal'].kind}"
            gname = base
            while gname in groups:
                used[base] += 1
                gname = f"{base}_{used[base] + 1}"
            nodes = np.array([[resolve(n) for n in m[2]] for m in members], np.int64)
            pnames = KINDS[kind][2] if kind != "VA" else extra["va"].stored + ["m"]
            params = {p: np.array([m[3][p] for m in members], float) for p in pnames}
            g = Group(kind, nodes, params)
            if kind == "VA":
                g.va = extra["va"]
            if kind == "EKVMOSFET":
                g.polarity, g.process = extra["polarity"], extra["process"]
            elif "signal" in extra:
                sig = extra["signal"]
                g.signal = type(sig)
                g.wave = {f: np.stack([m[4]["signal"].p[f] for m in members])
                          for f in sig.fields}
                ac = np.array([m[4]["ac"] for m in members], float)
                g.wave_ac = ac[:, 0] * np.exp(1j * np.deg2rad(ac[:, 1]))
            groups[gname] = g
            for i, m in enumerate(members):
                device_names[m[0]] = (gname, i)
        return Circuit(groups, tuple(self._nodes), device_names)
