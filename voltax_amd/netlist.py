"""SPICE netlist front end for the devices this package simulates.

`parse_netlist(text)` turns a deck into a `Circuit` through `CircuitBuilder`,
with the reference's naming rules (voltax/netlist/__init__.py:55-100): node
and device names as written (SPICE names are case-insensitive, the first
spelling wins), subcircuit instances flattened with the instance name as a
prefix (``X1.M1``, internal node ``X1.mid``), the node ``0`` (and ``gnd``) is
ground.

Supported here: ``R C V I M X`` elements, ``.subckt/.ends``, ``.param``
(numbers and ``{expression}`` / ``'expression'`` with + - * / ** and
parentheses over parameters), ``.model <name> nmos|pmos (...)`` for the EKV
MOSFET's process parameters (kp, vto/vth, n, lambda, phi, gamma, cox,
cgso/cov) or, through the ``models=`` hook, for Verilog-A elements
(`voltax_amd.va.level_models`); source values ``DC v``, ``PULSE(...)``,
``SIN(...)``, ``PWL(...)`` and ``AC mag [phase]``; ``+`` continuation lines,
``*`` and ``;`` comments.  Analysis and output directives (``.tran .ac .dc
.op .print .plot .probe .options .temp .control ... .endc``) are skipped;
``.end`` stops the deck.  Anything else raises `NetlistError` (a
`ValueError`) whose message starts with ``<source>:<line>:``, as the
reference's does.  Not supported: the reference's other element letters
(``L K E F G H B D Q S``), ``.include/.lib``, ``.func``, binning.
"""

from __future__ import annotations

import math
import re
import warnings
from dataclasses import dataclass, field

from .circuit import CircuitBuilder, MOSProcess

__all__ = ["NetlistError", "parse_netlist", "parse_value"]


class NetlistError(ValueError):
    pass


_EXP10 = {"t": 12, "g": 9, "meg": 6, "k": 3, "m": -3, "u": -6, "n": -9, "p": -12, "f": -15,
          "a": -18}
_NUMBER = re.compile(r"^([+-]?(?:\d+\.?\d*|\.\d+))(?:[eE]([+-]?\d+))?(meg|mil|[tgkmunpfa])?"
                     r"[a-z]*$", re.I)


def parse_value(text: str) -> float:
    """A SPICE number: ``10f``, ``1.2u``, ``4.7k``, ``2meg``, ``10pF``, ``5mil``
    (trailing unit letters after the scale are ignored, as in SPICE).  Scale
    letters shift the decimal exponent, so ``0.13u`` is exactly the double
    nearest to 0.13e-6."""
    m = _NUMBER.match(text.strip())
    if not m:
        raise ValueError(f"cannot parse value {text!r}")
    exp = int(m.group(2) or 0)
    unit = (m.group(3) or "").lower()
    if unit == "mil":
        return float(f"{m.group(1)}e{exp}") * 25.4e-6
    return float(f"{m.group(1)}e{exp + _EXP10.get(unit, 0)}")


# ----------------------------------------------------------------- expressions
_EXPR_TOKEN = re.compile(r"\s*(?:(\d+\.?\d*(?:[eE][+-]?\d+)?[a-zA-Z]*|\.\d+(?:[eE][+-]?\d+)?"
                         r"[a-zA-Z]*)|([A-Za-z_][A-Za-z_0-9.]*)|(\*\*|[-+*/()]))")


def _eval_expr(text, scope, where):
    """Value of an arithmetic expression over parameters (recursive descent)."""
    toks, pos = [], 0
    text = text.strip()
    while pos < len(text):
        m = _EXPR_TOKEN.match(text, pos)
        if not m or m.end() == pos:
            raise NetlistError(f"{where}: cannot read expression {text!r}")
        pos = m.end()
        toks.append(m.groups())
    toks.append((None, None, None))
    i = 0

    def peek():
        return toks[i]

    def take():
        nonlocal i
        i += 1
        return toks[i - 1]

    def atom():
        num, name, op = take()
        if num is not None:
            return parse_value(num)
        if name is not None:
            key = name.lower()
            if key in scope:
                return scope[key]
            if key in _FUNCS and peek()[2] == "(":
                take()
                v = expr()
                if take()[2] != ")":
                    raise NetlistError(f"{where}: missing ')' in {text!r}")
                return _FUNCS[key](v)
            raise NetlistError(f"{where}: unknown parameter {name!r}")
        if op == "(":
            v = expr()
            if take()[2] != ")":
                raise NetlistError(f"{where}: missing ')' in {text!r}")
            return v
        if op == "-":
            return -power()
        if op == "+":
            return power()
        raise NetlistError(f"{where}: cannot read expression {text!r}")

    def power():
        v = atom()
        if peek()[2] == "**":
            take()
            return v ** power()
        return v

    def term():
        v = power()
        while peek()[2] in ("*", "/"):
            op = take()[2]
            w = power()
            v = v * w if op == "*" else v / w
        return v

    def expr():
        v = term()
        while peek()[2] in ("+", "-"):
            op = take()[2]
            w = term()
            v = v + w if op == "+" else v - w
        return v

    v = expr()
    if peek() != (None, None, None):
        raise NetlistError(f"{where}: cannot read expression {text!r}")
    return float(v)


_FUNCS = {"sqrt": math.sqrt, "exp": math.exp, "log": math.log, "ln": math.log,
          "abs": abs, "log10": math.log10}


def _value(text, scope, where):
    t = text.strip()
    if (t.startswith("{") and t.endswith("}")) or (t.startswith("'") and t.endswith("'")):
        return _eval_expr(t[1:-1], scope, where)
    try:
        return parse_value(t)
    except ValueError:
        if t.lower() in scope:
            return scope[t.lower()]
        raise NetlistError(f"{where}: cannot parse value {text!r}") from None


# ----------------------------------------------------------------- deck
@dataclass
class _Subckt:
    name: str
    ports: list
    params: dict
    lines: list = field(default_factory=list)


@dataclass
class _Model:
    name: str
    type: str
    params: dict


_GROUP_PAREN = re.compile(r"(\w+)\s*\(([^()]*)\)")


def _fields(text):
    """Split a logical line into fields; a function call ``PULSE(a b c)``
    stays one field, ``name = value`` becomes ``name=value``."""
    text = re.sub(r"\s*=\s*", "=", text)
    out, i, n = [], 0, len(text)
    while i < n:
        if text[i].isspace():
            i += 1
            continue
        j = i
        depth, quote = 0, None
        while j < n:
            ch = text[j]
            if quote:
                if ch == quote:
                    quote = None
            elif ch in "{'":
                quote = "}" if ch == "{" else "'"
            elif ch == "(":
                depth += 1
            elif ch == ")":
                depth -= 1
            elif ch.isspace() and depth == 0:
                break
            j += 1
        out.append(text[i:j])
        i = j
    # "name (args)" written with a space joins into one field
    merged = []
    for f in out:
        if f.startswith("(") and merged and re.match(r"^[A-Za-z_]\w*$", merged[-1]):
            merged[-1] += f
        else:
            merged.append(f)
    return merged


def _logical_lines(text, source, title):
    raw = text.splitlines()
    lines = []
    start = 1 if title else 0
    for no, line in enumerate(raw[start:], start=start + 1):
        s = line.split(";", 1)[0].rstrip()
        if not s.strip() or s.lstrip().startswith("*"):
            continue
        if s.lstrip().startswith("+"):
            if not lines:
                raise NetlistError(f"{source}:{no}: continuation line without a line")
            lines[-1][1] += " " + s.lstrip()[1:]
            continue
        lines.append([no, s.strip()])
    return [(f"{source}:{no}", s) for no, s in lines]


_SKIP = {".tran", ".ac", ".dc", ".op", ".print", ".plot", ".probe", ".options", ".option",
         ".temp", ".save", ".meas", ".measure", ".ic", ".nodeset", ".global", ".title"}

_MOS_KEYS = {"kp": "kp", "vto": "vth", "vth": "vth", "vth0": "vth", "n": "n", "lambda": "lam",
             "phi": "phi", "gamma": "gamma", "cox": "cox", "cgso": "cov", "cgdo": "cov"}


def _split_params(fields, where, scope):
    pos, kw = [], {}
    for f in fields:
        if "=" in f:
            k, v = f.split("=", 1)
            kw[k.lower()] = _value(v, scope, where)
        else:
            pos.append(f)
    return pos, kw


def _signal(fields, where, scope, mod):
    """(waveform, ac magnitude, ac phase) of a source's value fields."""
    wave, ac, phase = None, 0.0, 0.0
    i = 0
    while i < len(fields):
        f = fields[i]
        low = f.lower()
        m = _GROUP_PAREN.match(f)
        if m and m.end() == len(f):
            name = m.group(1).lower()
            args = [_value(a, scope, where) for a in m.group(2).replace(",", " ").split()]
            if name == "pulse":
                keys = ("v1", "v2", "delay", "rise", "fall", "width", "period")
                wave = mod.Pulse(**dict(zip(keys, args)))
            elif name in ("sin", "sine"):
                keys = ("offset", "amplitude", "freq", "delay", "damping", "phase")
                wave = mod.Sine(**dict(zip(keys, args)))
            elif name == "pwl":
                if len(args) % 2:
                    raise NetlistError(f"{where}: PWL needs time/value pairs")
                wave = mod.PWL(args[0::2], args[1::2])
            else:
                raise NetlistError(f"{where}: unsupported source function {m.group(1)!r}")
            i += 1
        elif low == "dc":
            wave = _value(fields[i + 1], scope, where)
            i += 2
        elif low == "ac":
            ac = _value(fields[i + 1], scope, where) if i + 1 < len(fields) else 1.0
            i += 2
            if i < len(fields) and _is_number(fields[i], scope):
                phase = _value(fields[i], scope, where)
                i += 1
        else:
            wave = _value(f, scope, where)
            i += 1
    return (0.0 if wave is None else wave), ac, phase


def _is_number(text, scope):
    try:
        _value(text, scope, "")
        return True
    except (NetlistError, ValueError):
        return False


class _Emitter:
    def __init__(self, builder, models_hook, source):
        self.b = builder
        self.hook = dict(models_hook or {})
        self.source = source
        self.devices = {}    # lower-case name -> first spelling
        self.nodes = {}

    def _name(self, name):
        return self.devices.setdefault(name.lower(), name)

    def _node(self, node, prefix, pins):
        low = node.lower()
        if low in ("0", "gnd"):
            return "0"
        if low in pins:
            return pins[low]
        full = prefix + node
        return self.nodes.setdefault(full.lower(), full)

    def _mos_process(self, model, where):
        kind = model.type
        base = MOSProcess.nmos() if kind == "nmos" else MOSProcess.pmos()
        vals = {}
        for k, v in model.params.items():
            if k in ("level",):
                continue
            if k in _MOS_KEYS:
                x = abs(v) if _MOS_KEYS[k] == "vth" else v
                vals[_MOS_KEYS[k]] = x
            else:
                warnings.warn(f"{where}: model {model.name}: parameter {k!r} is not used by the "
                              "EKV MOSFET; ignored", UserWarning, stacklevel=4)
        return base.replace(**vals)

    def run(self, lines, scope, subckts, models, prefix="", pins=None, depth=0):
        pins = pins or {}
        if depth > 50:
            raise NetlistError(f"{self.source}: subcircuits nested too deep")
        for where, fields in lines:
            head = fields[0]
            letter = head[0].lower()
            name = self._name(prefix + head)
            try:
                self._element(letter, name, fields, where, scope, subckts, models, prefix, pins,
                              depth)
            except NetlistError:
                raise
            except (ValueError, KeyError, TypeError) as err:
                raise NetlistError(f"{where}: {err}") from None

    def _element(self, letter, name, fields, where, scope, subckts, models, prefix, pins, depth):
        node = lambda n: self._node(n, prefix, pins)  # noqa: E731
        if letter in ("r", "c"):
            if len(fields) < 4:
                raise NetlistError(f"{where}: missing {'resistance' if letter == 'r' else 'capacitance'}")
            pos, kw = _split_params(fields[3:], where, scope)
            value = _value(pos[0], scope, where) if pos else kw.get(letter)
            if value is None:
                raise NetlistError(f"{where}: missing value")
            if letter == "r":
                self.b.resistor(node(fields[1]), node(fields[2]), value, name=name)
            else:
                self.b.capacitor(node(fields[1]), node(fields[2]), value, name=name)
        elif letter in ("v", "i"):
            if len(fields) < 3:
                raise NetlistError(f"{where}: a source needs two nodes")
            from . import signals as mod

            wave, ac, phase = _signal(fields[3:], where, scope, mod)
            add = self.b.vsource if letter == "v" else self.b.isource
            add(node(fields[1]), node(fields[2]), wave, ac=ac, ac_phase=phase, name=name)
        elif letter == "m":
            if len(fields) < 6:
                raise NetlistError(f"{where}: a MOSFET needs drain gate source bulk model")
            d, g, s, bulk = (node(x) for x in fields[1:5])
            mname = fields[5]
            _, kw = _split_params(fields[6:], where, scope)
            model = models.get(mname.lower())
            if model is None and mname.lower() in ("nmos", "pmos"):
                model = _Model(mname.lower(), mname.lower(), {})
            if model is None:
                raise NetlistError(f"{where}: unknown model {mname!r}")
            level = int(model.params.get("level", 0))
            hook = self.hook.get(f"{model.type}:{level}")
            if hook is not None:
                dev = hook(model.name, model.type, model.params, (d, g, s, bulk), kw)
                self.b.add(dev, name=name)
                return
            if level not in (0, 1, 26) and model.name not in ("nmos", "pmos"):
                warnings.warn(f"{where}: model {model.name} level={level}: simulated with the "
                              "EKV MOSFET", UserWarning, stacklevel=3)
            proc = self._mos_process(model, where)
            add = self.b.nmos if model.type == "nmos" else self.b.pmos
            add(d, g, s, bulk, w=kw.get("w", 1e-6), l=kw.get("l", 0.13e-6), name=name,
                process=proc if model.params else None, dvth=kw.get("dvth", 0.0))
        elif letter == "x":
            pos, kw = _split_params(fields[1:], where, scope)
            if not pos:
                raise NetlistError(f"{where}: a subcircuit instance needs a subcircuit name")
            sub = subckts.get(pos[-1].lower())
            if sub is None:
                raise NetlistError(f"{where}: unknown subcircuit {pos[-1]!r}")
            nodes = pos[:-1]
            if len(nodes) != len(sub.ports):
                raise NetlistError(f"{where}: subcircuit {sub.name} has {len(sub.ports)} ports, "
                                   f"got {len(nodes)} nodes")
            inner = dict(scope)
            inner.update(sub.params)
            inner.update(kw)
            inner_pins = {p.lower(): node(n) for p, n in zip(sub.ports, nodes)}
            self.run(sub.lines, inner, subckts, models, prefix=name + ".", pins=inner_pins,
                     depth=depth + 1)
        else:
            raise NetlistError(f"{where}: unsupported element {name!r}")


def parse_netlist(text: str, title: bool = False, *, models=None, source: str = "<netlist>",
                  **builder_kwargs):
    """Parse a SPICE netlist into a `Circuit` (voltax/netlist/__init__.py:55-100).

    Args:
        text: netlist source.
        title: the first line is a title (SPICE file convention).
        models: ``{"nmos:<level>": factory, "pmos:<level>": factory}`` mapping
            model cards to other elements (`voltax_amd.va.level_models`); a
            factory is called as ``factory(card_name, type, params, nodes,
            instance)`` and returns an element instance.
        source: name used in error messages.
        **builder_kwargs: passed to `CircuitBuilder` (default processes).
    """
    lines = _logical_lines(text, source, title)
    top, subckts, models_, scope = [], {}, {}, {}
    stack = []
    in_control = False
    for where, s in lines:
        fields = _fields(s)
        head = fields[0].lower()
        if in_control:  # ngspice control blocks are interpreter scripts: skipped
            in_control = head != ".endc"
            continue
        if head.startswith("."):
            if head == ".end":
                break
            if head == ".subckt":
                if len(fields) < 2:
                    raise NetlistError(f"{where}: .subckt needs a name")
                ports, params = [], {}
                for f in fields[2:]:
                    if "=" in f:
                        k, v = f.split("=", 1)
                        params[k.lower()] = _value(v, scope, where)
                    elif f.lower() != "params:":
                        ports.append(f)
                stack.append(_Subckt(fields[1], ports, params))
            elif head == ".ends":
                if not stack:
                    raise NetlistError(f"{where}: .ends without .subckt")
                sub = stack.pop()
                subckts[sub.name.lower()] = sub
            elif head == ".param":
                for f in fields[1:]:
                    if "=" not in f:
                        raise NetlistError(f"{where}: .param needs name=value")
                    k, v = f.split("=", 1)
                    scope[k.lower()] = _value(v, scope, where)
            elif head == ".model":
                if len(fields) < 3:
                    raise NetlistError(f"{where}: .model needs a name and a type")
                mtype = fields[2]
                body = []
                m = _GROUP_PAREN.match(mtype)
                if m:
                    mtype, body = m.group(1), m.group(2).split()
                rest = " ".join(body + fields[3:]).strip()
                if rest.startswith("(") and rest.endswith(")"):
                    rest = rest[1:-1]
                _, params = _split_params(_fields(rest), where, scope)
                mtype = mtype.lower()
                if mtype not in ("nmos", "pmos"):
                    raise NetlistError(f"{where}: unsupported model type {mtype!r}")
                models_[fields[1].lower()] = _Model(fields[1], mtype, params)
            elif head == ".control":
                in_control = True
            elif head in _SKIP:
                continue
            else:
                raise NetlistError(f"{where}: unsupported directive {fields[0]!r}")
        else:
            (stack[-1].lines if stack else top).append((where, fields))
    if stack:
        raise NetlistError(f"{source}: .subckt {stack[-1].name} without .ends")
    builder = CircuitBuilder(**builder_kwargs)
    _Emitter(builder, models, source).run(top, scope, subckts, models_)
    return builder.build()
