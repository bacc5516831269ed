"""Circuits: element groups, layout, builder, and their GPU device tables.

Host-side mirror of the reference's circuit model so that code written for
it reads the same here:

* `Layout`, `Circuit`            voltax/circuit.py:31-272 (groups in sorted
                                 name order, node voltages then internals)
* element groups                 voltax/elements/{passive,sources,mosfet}.py
* `CircuitBuilder`               voltax/builder.py:52-256 (named nodes,
                                 fusion into groups, default group names)

Physics is not evaluated here: `Circuit.device_table` packs every device
into the flat tables the HIP kernels consume (kind, waveform, local
unknowns, parameter offsets) plus one FP64 parameter row per instance.
"""

from __future__ import annotations

import contextlib
from collections import defaultdict
from dataclasses import dataclass
from functools import cached_property

import numpy as np

from . import signals as sg

VT_300K = 0.025852
_GROUND = frozenset({"0", "gnd", "GND", "ground"})

KIND_R, KIND_C, KIND_V, KIND_I, KIND_EKV = 0, 1, 2, 3, 4
KIND_VA_BASE = 5  # compiled Verilog-A devices: kind 3 + m for m = 2..16 unknowns


# =============================================================================
# Elements
# =============================================================================


@dataclass(frozen=True)
class MOSProcess:
    """Shared technology parameters (voltax/elements/mosfet.py:26-87)."""

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

    def replace(self, **kw):
        return MOSProcess(**{**self.__dict__, **kw})


class Element:
    """A group of N devices of one type (struct of arrays)."""

    kind = -1
    terminals: tuple[str, ...] = ()
    n_internal = 0
    internal_names: tuple[str, ...] = ()
    params: tuple[str, ...] = ()  # per-device, differentiable (log-space where log_)

    def __init__(self, nodes, **values):
        self.nodes = np.asarray(nodes, dtype=np.int64).reshape(-1, len(self.terminals))
        n = self.nodes.shape[0]
        for p in self.params:
            setattr(self, p, np.broadcast_to(np.asarray(values[p], float), (n,)).copy())

    @property
    def size(self) -> int:
        return self.nodes.shape[0]

    def param_rows(self) -> np.ndarray:
        """(N, width) parameter rows of the devices."""
        return np.stack([getattr(self, p) for p in self.params], axis=1)

    def get(self, field, index=slice(None)):
        if "log_" + field in self.params:
            return np.exp(getattr(self, "log_" + field)[index])
        return getattr(self, field)[index]

    def set(self, index, **values):
        for key, v in values.items():
            if "log_" + key in self.params:
                getattr(self, "log_" + key)[index] = np.log(v)
            elif key in self.params:
                getattr(self, key)[index] = v
            else:
                raise AttributeError(f"{type(self).__name__} has no per-device parameter "
                                     f"{key!r}")


class Resistor(Element):
    kind, terminals, params = KIND_R, ("p", "n"), ("log_r",)

    def __init__(self, nodes, r=None, log_r=None):
        super().__init__(nodes, log_r=np.log(r) if log_r is None else log_r)


class Capacitor(Element):
    kind, terminals, params = KIND_C, ("p", "n"), ("log_c",)

    def __init__(self, nodes, c=None, log_c=None):
        super().__init__(nodes, log_c=np.log(c) if log_c is None else log_c)


class _Source(Element):
    terminals = ("p", "n")

    def __init__(self, nodes, waves, ac=0.0, ac_phase=0.0):
        super().__init__(nodes)
        waves = list(waves) if isinstance(waves, (list, tuple)) else [waves] * self.size
        self.waves = [sg.as_signal(w) for w in waves]
        self.ac = np.broadcast_to(np.asarray(ac, float), (self.size,)).copy()
        self.ac_phase = np.broadcast_to(np.asarray(ac_phase, float), (self.size,)).copy()

    @property
    def wave_code(self):
        return self.waves[0].code

    def param_rows(self):
        rows = [w.row() for w in self.waves]
        return np.stack(rows) if rows else np.zeros((0, 1))

    def get(self, field, index=slice(None)):
        """Waveform fields (``value`` of a DC source, ``v1``... of a pulse) or
        the element's own parameters."""
        if isinstance(index, (int, np.integer)) and field in self.waves[index].values:
            return float(np.asarray(self.waves[index].values[field]))
        return super().get(field, index)

    def set(self, index, **values):
        if "value" in values:
            w = self.waves[index]
            if not isinstance(w, sg.Constant):
                raise TypeError("set(value=...) on a non-DC source")
            self.waves[index] = sg.Constant(values.pop("value"))
        for key, v in list(values.items()):
            w = self.waves[index]
            if key in w.values:
                w.values[key] = np.asarray(v, float)
                values.pop(key)
        if values:
            super().set(index, **values)


class VoltageSource(_Source):
    kind, n_internal, internal_names = KIND_V, 1, ("i",)


class CurrentSource(_Source):
    kind = KIND_I


class EKVMOSFET(Element):
    """EKV 2.6-style MOSFET (voltax/elements/mosfet.py:155-205)."""

    kind, terminals = KIND_EKV, ("d", "g", "s", "b")
    params = ("log_w", "log_l", "dvth")

    def __init__(self, nodes, w=1e-6, l=0.13e-6, process=None, polarity="n", dvth=0.0):
        if polarity not in ("n", "p"):
            raise ValueError(f"polarity must be 'n' or 'p', got {polarity!r}")
        super().__init__(nodes, log_w=np.log(w), log_l=np.log(l), dvth=dvth)
        self.polarity = polarity
        self.process = process or (MOSProcess.nmos() if polarity == "n"
                                   else MOSProcess.pmos())

    def param_rows(self):
        p = self.process
        n = self.size
        sign = 1.0 if self.polarity == "n" else -1.0
        shared = np.array([sign, p.kp, p.vth, p.n, p.lam, p.vt, p.cox, p.cov])
        return np.concatenate([np.stack([self.log_w, self.log_l, self.dvth], 1),
                               np.broadcast_to(shared, (n, 8))], axis=1)


# =============================================================================
# Layout and circuit
# =============================================================================


@dataclass(frozen=True)
class Layout:
    node_names: tuple
    groups: tuple  # (group, n_devices, n_internal, offset)
    devices: tuple  # (device_name, group, index)
    internal_names: tuple

    @property
    def n_nodes(self):
        return len(self.node_names)

    @property
    def size(self):
        return self.n_nodes + sum(n * k for _, n, k, _ in self.groups)

    @cached_property
    def _node_index(self):
        return {n: i for i, n in enumerate(self.node_names)}

    @cached_property
    def _device_index(self):
        return {d: (g, i) for d, g, i in self.devices}

    def node(self, name) -> int:
        if name in _GROUND:
            return -1
        try:
            return self._node_index[name]
        except KeyError:
            raise KeyError(f"no node named {name!r}") from None

    def device(self, name):
        try:
            return self._device_index[name]
        except KeyError:
            raise KeyError(f"no device named {name!r}") from None

    def internal(self, device, which=None) -> int:
        g, idx = self.device(device)
        for name, n, k, off in self.groups:
            if name == g:
                names = dict(self.internal_names)[g]
                if k == 0:
                    raise ValueError(f"device {device!r} has no internal unknowns")
                w = 0 if which is None else names.index(which)
                return off + idx * k + w
        raise KeyError(device)


class Circuit:
    """Element groups plus a static layout (voltax/circuit.py:113-272)."""

    def __init__(self, elements: dict, node_names, device_names=None):
        self.elements = {k: elements[k] for k in sorted(elements)}
        nv = len(node_names)
        groups, internals, off = [], [], nv
        for name, el in self.elements.items():
            if el.size and (el.nodes.min() < -1 or el.nodes.max() >= nv):
                raise ValueError(f"group {name!r} references a node outside [-1, {nv})")
            groups.append((name, el.size, el.n_internal, off))
            internals.append((name, el.internal_names))
            off += el.size * el.n_internal
        if device_names is None:
            device_names = {f"{g}[{i}]": (g, i) for g, n, _, _ in groups for i in range(n)}
        devices = tuple((d, g, i) for d, (g, i) in device_names.items())
        self.layout = Layout(tuple(node_names), tuple(groups), devices, tuple(internals))

    # -- layout access
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

    def state(self, v=None, x=None, z=None):
        z = self.zeros() if z is None else np.array(z, float)
        for node, value in (v or {}).items():
            z[self.node(node)] = value
        for key, value in (x or {}).items():
            device, _, which = key.partition(".")
            z[self.layout.internal(device, which or None)] = value
        return z

    def device(self, name):
        g, i = self.layout.device(name)
        return self.elements[g], i

    def devices(self, group=None):
        found = [(g, i, d) for d, g, i in self.layout.devices if group is None or g == group]
        return [d for _, _, d in sorted(found)]

    def get(self, device, param):
        el, i = self.device(device)
        return el.get(param, i)

    def set(self, device, **values):
        """In-place per-device parameter update (physical units); returns self."""
        el, i = self.device(device)
        el.set(i, **values)
        self.__dict__.pop("_table", None)
        return self

    def ac_stimulus(self):
        """Complex small-signal excitation b added to f in AC analysis
        (voltax/circuit.py:204-207, voltax/elements/sources.py:87-116): minus
        the phasor ac * exp(j ac_phase) on each voltage source's branch row,
        the phasor through each current source (+ at p, - at n)."""
        S = self.size
        b = np.zeros(S + 1, complex)  # last slot = ground, dropped
        for name, el in self.elements.items():
            if not isinstance(el, _Source) or el.size == 0:
                continue
            ph = el.ac * np.exp(1j * np.deg2rad(el.ac_phase))
            u = self.unknowns(name)
            if isinstance(el, VoltageSource):
                np.add.at(b, u[:, 2], -ph)
            else:
                np.add.at(b, np.where(u[:, 0] < 0, S, u[:, 0]), ph)
                np.add.at(b, np.where(u[:, 1] < 0, S, u[:, 1]), -ph)
        return b[:S]

    def summary(self):
        lines = [f"Circuit: {self.n_nodes} nodes, state size {self.size}"]
        for name, n, k, _ in self.layout.groups:
            extra = f", {k} internal/device" if k else ""
            lines.append(f"  {name}: {n} x {type(self.elements[name]).__name__}{extra}")
        return "\n".join(lines)

    def unknowns(self, name):
        el = self.elements[name]
        _, n, k, off = next(g for g in self.layout.groups if g[0] == name)
        return np.concatenate([el.nodes, np.arange(off, off + n * k).reshape(n, k)], 1)

    # -- GPU device tables
    def topology_key(self):
        """Hashable description of everything the kernels treat as static."""
        parts = [self.layout.size, self.n_nodes]
        for name, el in self.elements.items():
            waves = tuple(w.key() for w in getattr(el, "waves", ()))
            va = getattr(el, "va", None)
            code = va.code.tobytes() if va is not None else b""
            parts.append((name, el.kind, el.nodes.tobytes(), waves, code))
        return tuple(parts)

    @property
    def is_linear(self):
        """Only resistors, capacitors and independent sources: the Jacobian
        depends on the integration coefficients and gmin alone."""
        return all(el.kind in (KIND_R, KIND_C, KIND_V, KIND_I) for el in self.elements.values()
                   if el.size)

    @property
    def has_va(self):
        return any(getattr(el, "va", None) is not None for el in self.elements.values())

    def va_code(self):
        """(int32 words of every Verilog-A group's program, {group: offset})."""
        words, offs = [], {}
        for name, el in self.elements.items():
            va = getattr(el, "va", None)
            if va is not None and el.size:
                offs[name] = len(words)
                words.extend(int(w) for w in va.code)
        return np.asarray(words, np.int32), offs

    def device_table(self):
        """(kind, wave, unknowns[D, W], param_off, param_row) in group order; W is
        4, or the most unknowns of a device (Verilog-A devices have up to 16)."""
        kinds, waves, unks, poffs, rows = [], [], [], [], []
        off = 0
        va_offs = self.va_code()[1]
        width = max([4] + [self.unknowns(n).shape[1] for n, el in self.elements.items()
                           if el.size])
        for name, el in self.elements.items():
            if el.size == 0:
                continue
            u = self.unknowns(name)
            pr = np.asarray(el.param_rows(), float).reshape(el.size, -1)
            wcode = el.wave_code if isinstance(el, _Source) else va_offs.get(name, 0)
            kinds.append(np.full(el.size, el.kind, np.int32))
            waves.append(np.full(el.size, wcode, np.int32))
            un = np.full((el.size, width), -1, np.int32)
            un[:, : u.shape[1]] = u
            unks.append(un)
            poffs.append(off + pr.shape[1] * np.arange(el.size, dtype=np.int64))
            rows.append(pr.reshape(-1))
            off += pr.size
        if not kinds:
            return (np.zeros(0, np.int32), np.zeros(0, np.int32), np.zeros((0, width), np.int32),
                    np.zeros(0, np.int32), np.zeros(0))
        return (np.concatenate(kinds), np.concatenate(waves), np.concatenate(unks),
                np.concatenate(poffs).astype(np.int32), np.concatenate(rows))


# =============================================================================
# Builder
# =============================================================================


_PREFIX = {Resistor: "R", Capacitor: "C", VoltageSource: "V", CurrentSource: "I",
           EKVMOSFET: "M"}


class CircuitBuilder:
    """Incrementally assemble a `Circuit` (voltax/builder.py:52-256)."""

    def __init__(self, nmos: MOSProcess | None = None, pmos: MOSProcess | None = None):
        self.nmos_process = nmos or MOSProcess.nmos()
        self.pmos_process = pmos or MOSProcess.pmos()
        self._devices = []  # (name, cls, terminals, kwargs)
        self._names = set()
        self._nodes: dict[str, None] = {}
        self._scope: list[str] = []
        self._counters = defaultdict(int)

    def node(self, name=None):
        if name is None:
            self._counters["_node"] += 1
            name = f"_{self._counters['_node']}"
        full = ".".join([*self._scope, name])
        self._nodes.setdefault(full, None)
        return full

    @contextlib.contextmanager
    def scope(self, prefix):
        self._scope.append(prefix)
        try:
            yield
        finally:
            self._scope.pop()

    @property
    def nodes(self):
        return list(self._nodes)

    def _add(self, cls, terminals, kwargs, name):
        if name is None:
            pre = _PREFIX[cls]
            self._counters[pre] += 1
            name = f"{pre}{self._counters[pre]}"
        name = ".".join([*self._scope, name])
        if name in self._names:
            raise ValueError(f"duplicate device name {name!r}")
        self._names.add(name)
        terms = tuple(str(n) for n in terminals)
        for n in terms:
            if n not in _GROUND:
                self._nodes.setdefault(n, None)
        self._devices.append((name, cls, terms, kwargs))
        return name

    def add(self, device, name=None):
        """Adds the devices of an element instance built on node names, e.g. a
        Verilog-A element ``R(("a", "b"), r=1e3)`` (voltax/builder.py ``add``)."""
        cls = type(device)
        if getattr(cls, "va", None) is None:
            raise TypeError("add() takes Verilog-A element instances (voltax_amd.va)")
        names = []
        for i in range(device.size):
            nm = name if (name is not None and device.size == 1) else (
                None if name is None else f"{name}[{i}]")
            if nm is None:
                self._counters["X"] += 1
                nm = f"X{self._counters['X']}"
            names.append(self._add(cls, tuple(device.nodes[i]),
                                   {"row": device.rows[i].copy()}, nm))
        return names[0] if len(names) == 1 else names

    def resistor(self, p, n, r, name=None):
        return self._add(Resistor, (p, n), {"r": r}, name)

    def capacitor(self, p, n, c, name=None):
        return self._add(Capacitor, (p, n), {"c": c}, name)

    def vsource(self, p, n, value=0.0, ac=0.0, ac_phase=0.0, name=None):
        return self._add(VoltageSource, (p, n),
                         {"waves": sg.as_signal(value), "ac": ac, "ac_phase": ac_phase}, name)

    def isource(self, p, n, value=0.0, ac=0.0, ac_phase=0.0, name=None):
        return self._add(CurrentSource, (p, n),
                         {"waves": sg.as_signal(value), "ac": ac, "ac_phase": ac_phase}, name)

    def nmos(self, d, g, s, b=None, w=1e-6, l=0.13e-6, name=None, process=None, dvth=0.0):
        return self._add(EKVMOSFET, (d, g, s, s if b is None else b),
                         {"w": w, "l": l, "process": process or self.nmos_process,
                          "polarity": "n", "dvth": dvth}, name)

    def pmos(self, d, g, s, b=None, w=1e-6, l=0.13e-6, name=None, process=None, dvth=0.0):
        return self._add(EKVMOSFET, (d, g, s, s if b is None else b),
                         {"w": w, "l": l, "process": process or self.pmos_process,
                          "polarity": "p", "dvth": dvth}, name)

    @staticmethod
    def _key(cls, kw):
        if cls is EKVMOSFET:
            return (cls, kw["polarity"], id(kw["process"]))
        if issubclass(cls, _Source):
            return (cls, kw["waves"].key())
        return (cls,)

    @staticmethod
    def _group_name(cls, kw):
        base = cls.__name__
        if getattr(cls, "polarity_label", None):
            return f"{base}_{cls.polarity_label}"
        if cls is EKVMOSFET:
            return f"{base}_{kw['polarity']}"
        if issubclass(cls, _Source) and not isinstance(kw["waves"], sg.Constant):
            return f"{base}_{type(kw['waves']).__name__}"
        return base

    def build(self) -> Circuit:
        index = {n: i for i, n in enumerate(self._nodes)}
        resolve = lambda n: -1 if n in _GROUND else index[n]  # noqa: E731
        buckets: dict = {}
        for dev in self._devices:
            buckets.setdefault(self._key(dev[1], dev[3]), []).append(dev)
        groups, device_names, used = {}, {}, defaultdict(int)
        for members in buckets.values():
            cls, kw0 = members[0][1], members[0][3]
            base = self._group_name(cls, kw0)
            gname = base
            while gname in groups:
                used[base] += 1
                gname = f"{base}_{used[base] + 1}"
            nodes = [[resolve(n) for n in m[2]] for m in members]
            if cls is EKVMOSFET:
                el = EKVMOSFET(nodes, w=[m[3]["w"] for m in members],
                               l=[m[3]["l"] for m in members], process=kw0["process"],
                               polarity=kw0["polarity"], dvth=[m[3]["dvth"] for m in members])
            elif issubclass(cls, _Source):
                el = cls(nodes, [m[3]["waves"] for m in members],
                         [m[3]["ac"] for m in members], [m[3]["ac_phase"] for m in members])
            elif getattr(cls, "va", None) is not None:
                el = cls.from_rows(nodes, np.stack([m[3]["row"] for m in members]))
            elif cls is Resistor:
                el = Resistor(nodes, r=[m[3]["r"] for m in members])
            else:
                el = Capacitor(nodes, c=[m[3]["c"] for m in members])
            groups[gname] = el
            for i, m in enumerate(members):
                device_names[m[0]] = (gname, i)
        return Circuit(groups, tuple(self._nodes), device_names)
