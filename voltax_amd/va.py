"""Verilog-A compact models on the GPU (subset).

`load(source)` parses one Verilog-A module; `VAModel.element(card)` returns an
element class whose devices stamp like the built-in ones.  The analog block is
compiled once, on the host, into a register bytecode over dual numbers (value
plus derivatives with respect to the device's unknowns); the stamping kernel
interprets it per device (`csrc/va_eval.hpp`) and writes residual and Jacobian
contributions exactly where a built-in device would.

Reference: the Verilog-A front end and its JAX code generator,
voltax/va/preprocess.py, voltax/va/parser.py, voltax/va/compiler.py:342-1525,
voltax/va/model.py:153-703, runtime helpers voltax/va/runtime.py:10-90.

Supported: the preprocessor (`define with arguments and continuations,
`undef, `ifdef/`ifndef/`elsif/`else/`endif, the standard disciplines and
constants headers, overrides); `parameter real|integer` with defaults, ranges
and excludes, ``(* type="instance" *)``, ``aliasparam``; `real`/`integer`
variables (module and block scope), named branches, `analog function`s;
statements: assignment (integer targets round), ``if``/``else``, ``case``,
``for`` and ``while`` loops (trip count bounded; data-dependent exits are
unrolled under per-device masks), system tasks such as ``$strobe`` (ignored);
contributions ``I(a, b) <+ e`` and ``V(a, b) <+ e`` (branch-current unknown
``i(a,b)``; a constant ``V(a, b) <+ 0`` collapses the two nodes), probes
``V(a, b)``, ``V(a)``, ``I(a, b)`` of voltage branches, ``ddt`` as a linear
term of a contribution; integer division and ``%`` truncate; the functions
exp, ln, log (base 10), log10, sqrt, abs, pow, min, max, limexp, tanh, sin,
cos and ``$temperature``, ``$vt``, ``$simparam``, ``$param_given``,
``$mfactor``.  As in the reference, both sides of a data-dependent branch are
evaluated and selected per device, and static conditions (parameters,
integer counters) are folded at compile time.  Gradients with respect to
real instance parameters, the multiplicity ``m`` and the model parameters
made ``differentiable`` (read from the device row like instance parameters)
come from seeded passes of the same program.
The smooth mode (``smooth=``: a width, or {source line: width}) replaces
min / max / abs and ``if (x < b) x = b`` clamps on device-dependent values
by softplus surrogates (voltax/va/runtime.py:76-90); ``equality="limit"``
keeps the values of ``if (a == b)`` special cases and differentiates the
general arm there (bytecode op ``limit``).
Programs of up to 4 unknowns and 64 registers are interpreted on the GPU;
larger ones (BSIM4-class: internal nodes, branch currents, up to 16 unknowns,
about a thousand statements) become straight-line HIP C++ that the library
compiles at run time (csrc/va_codegen.cpp).  Not supported: noise sources.
"""

from __future__ import annotations

import math
import re
import struct
from dataclasses import dataclass, field

import numpy as np

# opcodes of the bytecode (csrc/va_eval.hpp)
OPS = {name: i for i, name in enumerate(
    "const param unk add sub mul div neg exp log sqrt abs pow min max limexp tanh sin cos "
    "lt le gt ge eq ne and or not select trunc round limit".split())}
MAGIC = 0x5642
HEAD = 6  # words before the instructions
MAX_REGS = 1 << 16     # SSA registers of a program (above 64 it runs as generated code)
MAX_UNKNOWNS = 16      # unknowns per device (above 4 it runs as generated code)
MAX_UNROLL = 256


class VAError(ValueError):
    pass


_FATAL_TASKS = {"$finish", "$stop", "$fatal", "$error"}


# =============================================================================
# Lexer and parser
# =============================================================================

_SCALE = {"T": 1e12, "G": 1e9, "M": 1e6, "K": 1e3, "k": 1e3, "m": 1e-3, "u": 1e-6,
          "n": 1e-9, "p": 1e-12, "f": 1e-15, "a": 1e-18}
_TOKEN = re.compile(r"""
    (?P<ws>\s+|//[^\n]*|/\*.*?\*/)
  | (?P<attr>\(\*.*?\*\))
  | (?P<num>(?:\d+\.\d*|\.\d+|\d+)(?:[eE][+-]?\d+)?[TGMKkmunpfa]?)
  | (?P<str>"[^"]*")
  | (?P<id>[A-Za-z_$][A-Za-z_0-9$]*)
  | (?P<op><\+|<=|>=|==|!=|&&|\|\||[-+*/%()<>;,:=?!\[\]{}@`.])
""", re.S | re.X)


# Physical and mathematical constants of the standard `constants.vams` header
# (Verilog-AMS LRM annex; NIST 1998 values unless a PHYSICAL_CONSTANTS_*
# macro selects another set, as the header specifies).
_CONSTANTS = {
    "M_E": "2.7182818284590452354", "M_LOG2E": "1.4426950408889634074",
    "M_LOG10E": "0.43429448190325182765", "M_LN2": "0.69314718055994530942",
    "M_LN10": "2.30258509299404568402", "M_PI": "3.14159265358979323846",
    "M_TWO_PI": "6.28318530717958647693", "M_PI_2": "1.57079632679489661923",
    "M_PI_4": "0.78539816339744830962", "M_1_PI": "0.31830988618379067154",
    "M_2_PI": "0.63661977236758134308", "M_2_SQRTPI": "1.12837916709551257390",
    "M_SQRT2": "1.41421356237309504880", "M_SQRT1_2": "0.70710678118654752440",
    "P_C": "2.99792458e8", "P_U0": "(4.0e-7 * 3.14159265358979323846)", "P_CELSIUS0": "273.15",
}
_PHYS = {  # set -> (P_Q, P_K, P_H, P_EPS0)
    "SPICE": ("1.60219e-19", "1.38062e-23", "6.62620e-34", "8.854214871e-12"),
    "OLD": ("1.6021918e-19", "1.3806226e-23", "6.6260755e-34", "8.85418792394420013968e-12"),
    "NIST2010": ("1.602176565e-19", "1.3806488e-23", "6.62606957e-34", "8.854187817e-12"),
    "NIST1998": ("1.602176462e-19", "1.3806503e-23", "6.62606876e-34", "8.854187817e-12"),
}
_BUILTIN_INCLUDES = ("disciplines.vams", "discipline.h", "constants.vams", "constants.h")
_MACRO_USE = re.compile(r"`([A-Za-z_][A-Za-z_0-9]*)")


def _strip_comments(src):
    """`src` with // and /* */ comments blanked (newlines kept, so line
    numbers stay) and string literals left as they are."""
    out, i, n = [], 0, len(src)
    while i < n:
        ch = src[i]
        if ch == '"':
            j = i + 1
            while j < n and src[j] != '"':
                j += 2 if src[j] == "\\" else 1
            out.append(src[i:j + 1])
            i = j + 1
        elif src.startswith("//", i):
            j = src.find("\n", i)
            i = n if j < 0 else j
        elif src.startswith("/*", i):
            j = src.find("*/", i + 2)
            if j < 0:
                raise VAError("unterminated /* comment")
            out.append("".join(c if c == "\n" else " " for c in src[i:j + 2]))
            i = j + 2
        else:
            out.append(ch)
            i += 1
    return "".join(out)


def preprocess(src: str, defines=None, overrides=None, source="<string>") -> str:
    """Expands the compiler directives of a Verilog-A source: `define (with
    arguments, backslash continuations), `undef, `ifdef / `ifndef / `elsif /
    `else / `endif, and `include of the standard disciplines and constants
    headers (built in; other files are not read).  Comments are removed and
    line structure is kept; errors name ``<source>:<line>``.  `defines` are
    set before the text (a ``PHYSICAL_CONSTANTS_{SPICE,OLD,NIST2010}`` define
    picks that set of physical constants, NIST 1998 otherwise); `overrides`
    win over the text's own `define of the same name
    (voltax/va/preprocess.py)."""
    macros = {k: ([], v) for k, v in _CONSTANTS.items()}
    user = dict(defines or {})
    locked = {k: str(v) for k, v in (overrides or {}).items()}
    which = next((k for k in ("SPICE", "OLD", "NIST2010")
                  if f"PHYSICAL_CONSTANTS_{k}" in user or f"PHYSICAL_CONSTANTS_{k}" in locked),
                 "NIST1998")
    for name, val in zip(("P_Q", "P_K", "P_H", "P_EPS0"), _PHYS[which]):
        macros[name] = ([], val)
    for k, v in {**user, **locked}.items():
        macros[k] = ([], str(v))
    lines = _strip_comments(src).split("\n")
    out, stack, i = [], [], 0  # stack of [active, taken]
    active = lambda: all(a for a, _ in stack)  # noqa: E731
    try:
        while i < len(lines):
            i = _directive(lines, i, macros, locked, stack, active, out)
    except VAError as err:
        if not str(err).startswith(f"{source}:"):
            raise VAError(f"{source}:{i + 1}: {err}") from None
        raise
    if stack:
        raise VAError(f"{source}:{len(lines)}: `ifdef without `endif")
    return "\n".join(out)


def _directive(lines, i, macros, locked, stack, active, out):
    """Processes line `i` (with the continuation lines of a `define); returns
    the index of the next line."""
    line = lines[i]
    i += 1
    m = re.match(r"\s*`(define|undef|ifdef|ifndef|elsif|else|endif|include|timescale|"
                 r"default_nettype|resetall)\b(.*)", line)
    if not m:
        out.append(_expand(line, macros) if active() else "")
        return i
    kind, rest = m.group(1), m.group(2).strip()
    if kind == "define":
        body, extra = rest, 0
        while body.endswith("\\") and i < len(lines):
            body = body[:-1] + " " + lines[i].strip()
            i += 1
            extra += 1
        out.extend([""] * (extra + 1))
        if not active():
            return i
        mm = re.match(r"([A-Za-z_][A-Za-z_0-9]*)(\(([^)]*)\))?\s*(.*)$", body, re.S)
        if not mm:
            raise VAError(f"bad `define: {body!r}")
        name = mm.group(1)
        if name in locked:  # an override wins over the text's own definition
            return i
        args = [a.strip() for a in mm.group(3).split(",")] if mm.group(2) else []
        if mm.group(2) and not args[0]:
            args = []
        macros[name] = (args, mm.group(4).strip())
        return i
    out.append("")
    word = rest.split()[0] if rest.split() else ""
    if kind in ("undef", "ifdef", "ifndef", "elsif") and not word:
        raise VAError(f"`{kind} needs a macro name")
    if kind == "undef":
        if active() and word not in locked:
            macros.pop(word, None)
    elif kind in ("ifdef", "ifndef"):
        defined = word in macros
        cond = defined if kind == "ifdef" else not defined
        stack.append([cond, cond])
    elif kind == "elsif":
        if not stack:
            raise VAError("`elsif without `ifdef")
        cond = not stack[-1][1] and word in macros
        stack[-1] = [cond, stack[-1][1] or cond]
    elif kind == "else":
        if not stack:
            raise VAError("`else without `ifdef")
        stack[-1] = [not stack[-1][1], True]
    elif kind == "endif":
        if not stack:
            raise VAError("`endif without `ifdef")
        stack.pop()
    elif kind == "include" and active():
        name = rest.strip('"<> ')
        if name.split("/")[-1] not in _BUILTIN_INCLUDES:
            raise VAError(f"`include {name!r}: only the standard disciplines and "
                          "constants headers are available")
    return i


def _expand(line, macros, depth=0):
    """Replaces macro uses in one line (arguments of function-like macros are
    expanded first; nesting limited)."""
    if depth > 32:
        raise VAError("macro expansion too deep")
    res, pos = [], 0
    while True:
        m = _MACRO_USE.search(line, pos)
        if not m:
            res.append(line[pos:])
            break
        res.append(line[pos:m.start()])
        name = m.group(1)
        if name not in macros:
            raise VAError(f"undefined macro `{name}")
        params, body = macros[name]
        end = m.end()
        if params:
            # balanced argument list
            j = end
            while j < len(line) and line[j].isspace():
                j += 1
            if j >= len(line) or line[j] != "(":
                raise VAError(f"macro `{name} needs arguments")
            depth_p, k, args, cur = 0, j, [], ""
            while k < len(line):
                ch = line[k]
                if ch == "(":
                    depth_p += 1
                    if depth_p == 1:
                        k += 1
                        continue
                elif ch == ")":
                    depth_p -= 1
                    if depth_p == 0:
                        args.append(cur)
                        break
                elif ch == "," and depth_p == 1:
                    args.append(cur)
                    cur = ""
                    k += 1
                    continue
                cur += ch
                k += 1
            else:
                raise VAError(f"unterminated arguments of `{name}")
            if len(args) != len(params):
                raise VAError(f"macro `{name} takes {len(params)} arguments")
            vals = [_expand(a.strip(), macros, depth + 1) for a in args]
            text = body
            for p_, v in zip(params, vals):
                text = re.sub(rf"\b{re.escape(p_)}\b", lambda _m, v=v: v, text)
            end = k + 1
        else:
            text = body
        res.append(_expand(text, macros, depth + 1))
        pos = end
    return "".join(res)


def _tokens(src):
    src = preprocess(src)
    out, pos = [], 0
    while pos < len(src):
        m = _TOKEN.match(src, pos)
        if not m:
            raise VAError(f"cannot read Verilog-A near {src[pos:pos + 20]!r}")
        pos = m.end()
        kind = m.lastgroup
        if kind == "ws":
            continue
        text = m.group(kind)
        line = src.count("\n", 0, m.start()) + 1  # source line (preprocessing keeps lines)
        if kind == "num":
            unit = text[-1] if text[-1] in _SCALE else ""
            val = float(text[:-1] if unit else text) * (_SCALE[unit] if unit else 1.0)
            is_int = not unit and text.isdigit()
            out.append(("num", val, is_int, line))
        else:
            out.append((kind, text, False, line))
    out.append(("eof", "", False, src.count("\n") + 1))
    return out


@dataclass
class Param:
    name: str
    type: str
    default: object  # expression tree
    instance: bool = False
    ranges: list = field(default_factory=list)    # (lo_bracket, lo, hi, hi_bracket)
    excludes: list = field(default_factory=list)  # ranges or single expressions

    @property
    def positive(self):
        """The range excludes zero from below (``from (0:...``): such real
        instance parameters are kept in log space, as in the reference
        (voltax/va/model.py:473-477)."""
        return any(lo_b == "(" and lo[0] == "num" and lo[1] == 0.0
                   for lo_b, lo, _, _ in self.ranges)


@dataclass
class Module:
    name: str
    ports: list
    nodes: list
    params: dict
    variables: dict
    body: list = field(default_factory=list)
    grounds: list = field(default_factory=list)
    functions: dict = field(default_factory=dict)  # name -> Function
    branches: dict = field(default_factory=dict)   # named branch -> (a, b or None)
    aliases: dict = field(default_factory=dict)    # aliasparam name -> parameter


@dataclass
class Function:
    """An analog function: inputs in order, typed locals, body statements;
    the value is that of the variable named after the function."""
    name: str
    type: str
    inputs: list
    variables: dict
    body: list


class _Parser:
    def __init__(self, src):
        self.t = _tokens(src)
        self.i = 0
        self.module_vars = {}
        self.branches = {}

    def peek(self, k=0):
        return self.t[self.i + k]

    def next(self):
        tok = self.t[self.i]
        self.i += 1
        return tok

    def accept(self, text):
        if self.peek()[1] == text and self.peek()[0] in ("op", "id"):
            self.i += 1
            return True
        return False

    def expect(self, text):
        if not self.accept(text):
            raise VAError(f"expected {text!r}, found {self.peek()[1]!r}")

    def ident(self):
        kind, text = self.next()[:2]
        if kind != "id":
            raise VAError(f"expected a name, found {text!r}")
        return text

    # ---------------------------------------------------------------- module
    def module(self):
        while not (self.peek()[0] == "id" and self.peek()[1] == "module"):
            if self.peek()[0] == "eof":
                raise VAError("no module found")
            self.next()
        self.next()
        name = self.ident()
        ports = []
        if self.accept("("):
            if not self.accept(")"):
                ports.append(self.ident())
                while self.accept(","):
                    ports.append(self.ident())
                self.expect(")")
        self.expect(";")
        mod = Module(name, ports, [], {}, {})
        attr_instance = False
        while not self.accept("endmodule"):
            kind, text = self.peek()[:2]
            if kind == "attr":
                self.next()
                attr_instance = "instance" in text
                continue
            if kind == "eof":
                raise VAError("missing endmodule")
            if text in ("inout", "input", "output"):
                self.next()
                self._names()
            elif text == "ground":
                self.next()
                mod.grounds.extend(self._names())
            elif text == "electrical":
                self.next()
                for n in self._names():
                    if n not in mod.nodes:
                        mod.nodes.append(n)
            elif text == "parameter":
                self.next()
                ptype = "real"
                if self.peek()[1] in ("real", "integer"):
                    ptype = self.next()[1]
                while True:
                    pname = self.ident()
                    self.expect("=")
                    default = self.expr()
                    ranges, excludes = self._ranges()
                    mod.params[pname] = Param(pname, ptype, default, attr_instance, ranges,
                                              excludes)
                    if not self.accept(","):
                        break
                self.expect(";")
                attr_instance = False
            elif text in ("real", "integer"):
                vtype = self.next()[1]
                for n in self._names():
                    mod.variables[n] = vtype
            elif text == "analog" and self.peek(1)[1] == "function":
                self.next()
                self.next()
                f = self.function()
                mod.functions[f.name] = f
            elif text == "analog":
                self.next()
                self.module_vars = mod.variables
                mod.body.append(self.statement())
            elif text == "branch":  # branch (a, b) name, ...;
                self.next()
                self.expect("(")
                a = self.ident()
                b = self.ident() if self.accept(",") else None
                self.expect(")")
                for n in self._names():
                    mod.branches[n] = self.branches[n] = (a, b)
            elif text == "aliasparam":
                self.next()
                alias = self.ident()
                self.expect("=")
                mod.aliases[alias] = self.ident()
                self.expect(";")
            else:
                raise VAError(f"unsupported module item {text!r}")
        for p in mod.ports:
            if p not in mod.nodes:
                mod.nodes.insert(0, p)
        bad = [g for g in mod.grounds if g in mod.ports]
        if bad:
            raise VAError(f"port {bad[0]!r} declared as ground")
        mod.nodes = [n for n in mod.nodes if n not in mod.grounds]
        return mod

    def function(self):
        """``analog function real|integer name; input a, b; real a, b;
        statements endfunction`` (after ``analog function``)."""
        ftype = "real"
        if self.peek()[1] in ("real", "integer"):
            ftype = self.next()[1]
        name = self.ident()
        self.expect(";")
        inputs, variables, body = [], {}, []
        self.module_vars = variables
        while not self.accept("endfunction"):
            text = self.peek()[1]
            if text in ("input", "output", "inout"):
                self.next()
                inputs.extend(self._names())
            elif text in ("real", "integer"):
                vtype = self.next()[1]
                for n in self._names():
                    variables[n] = vtype
            elif self.peek()[0] == "eof":
                raise VAError("missing endfunction")
            else:
                body.append(self.statement())
        for n in inputs:
            variables.setdefault(n, "real")
        variables.setdefault(name, ftype)
        return Function(name, ftype, inputs, variables, body)

    def _names(self):
        names = [self.ident()]
        while self.accept(","):
            names.append(self.ident())
        self.expect(";")
        return names

    def _ranges(self):
        """``from <range>`` and ``exclude <range or constant>`` clauses."""
        ranges, excludes = [], []
        while self.peek()[1] in ("from", "exclude"):
            which = self.next()[1]
            if self.peek()[1] in ("(", "["):
                lo_b = self.next()[1]
                lo = self.expr()
                self.expect(":")
                hi = self.expr()
                hi_b = self.next()[1]
                if hi_b not in (")", "]"):
                    raise VAError(f"bad range end {hi_b!r}")
                (ranges if which == "from" else excludes).append((lo_b, lo, hi, hi_b))
            elif which == "exclude":
                excludes.append(self.unary())
            else:
                raise VAError("expected a range after 'from'")
        return ranges, excludes

    # ------------------------------------------------------------ statements
    def statement(self):
        kind, text = self.peek()[:2]
        if text == "begin":
            self.next()
            if self.accept(":"):
                self.ident()
            body = []
            while not self.accept("end"):
                if self.peek()[1] in ("real", "integer") and self.peek(1)[0] == "id":
                    vtype = self.next()[1]  # block-local declaration
                    for n in self._names():
                        self.module_vars.setdefault(n, vtype)
                    continue
                body.append(self.statement())
            return ("block", body)
        if text == "for":  # for (init; cond; step) body == init; while (cond) {body; step}
            self.next()
            self.expect("(")
            init = self._assignment(";")
            cond = self.expr()
            self.expect(";")
            step = self._assignment(")")
            body = self.statement()
            return ("block", [init, ("while", cond, ("block", [body, step]))])
        if text == "case":  # a chain of ifs on equality with the case expression
            self.next()
            self.expect("(")
            sel = self.expr()
            self.expect(")")
            items, default = [], ("block", [])
            while not self.accept("endcase"):
                if self.accept("default"):
                    self.accept(":")
                    default = self.statement()
                    continue
                labels = [self.expr()]
                while self.accept(","):
                    labels.append(self.expr())
                self.expect(":")
                items.append((labels, self.statement()))
            out = default
            for labels, stmt in reversed(items):
                cond = None
                for lab in labels:
                    c = ("==", sel, lab)
                    cond = c if cond is None else ("||", cond, c)
                out = ("if", cond, stmt, out)
            return out
        if kind == "id" and text.startswith("$"):  # system tasks: $strobe, $display, ...
            name = self.next()[1]
            if name in _FATAL_TASKS:
                raise VAError(f"{name} in the analog block is not supported")
            depth = 0
            while True:  # skip the argument list
                t = self.next()[1]
                if t == "(":
                    depth += 1
                elif t == ")":
                    depth -= 1
                elif t == ";" and depth == 0:
                    break
                elif t == "" and self.peek(-1)[0] == "eof":
                    raise VAError("unterminated system task")
            return ("block", [])
        if text == "if":
            line = self.next()[3]
            self.expect("(")
            cond = self.expr()
            self.expect(")")
            then = self.statement()
            other = self.statement() if self.accept("else") else ("block", [])
            return ("if", cond, then, other, line)
        if text == "while":
            self.next()
            self.expect("(")
            cond = self.expr()
            self.expect(")")
            return ("while", cond, self.statement())
        if text == ";":
            self.next()
            return ("block", [])
        if kind == "id" and text in ("I", "V") and self.peek(1)[1] == "(":
            save = self.i
            acc = self.next()[1]
            self.expect("(")
            a = self.ident()
            b = self.ident() if self.accept(",") else None
            self.expect(")")
            if b is None and a in self.branches:
                a, b = self.branches[a]
            if self.accept("<+"):
                e = self.expr()
                self.expect(";")
                return ("contrib", acc, a, b, e)
            self.i = save
        return self._assignment(";")

    def _assignment(self, end):
        name = self.ident()
        self.expect("=")
        e = self.expr()
        self.expect(end)
        return ("assign", name, e)

    # ----------------------------------------------------------- expressions
    _PREC = [("?",), ("||",), ("&&",), ("==", "!="), ("<", "<=", ">", ">="), ("+", "-"),
             ("*", "/", "%")]

    def expr(self, level=0):
        if level == 0:
            c = self.expr(1)
            if self.accept("?"):
                a = self.expr(0)
                self.expect(":")
                b = self.expr(0)
                return ("?", c, a, b)
            return c
        if level == len(self._PREC):
            return self.unary()
        lhs = self.expr(level + 1)
        while self.peek()[0] == "op" and self.peek()[1] in self._PREC[level]:
            op = self.next()[1]
            lhs = (op, lhs, self.expr(level + 1))
        return lhs

    def unary(self):
        if self.accept("-"):
            return ("neg", self.unary())
        if self.accept("+"):
            return self.unary()
        if self.accept("!"):
            return ("!", self.unary())
        return self.primary()

    def primary(self):
        kind, text, _, line = self.next()
        if kind == "num":
            return ("num", text, self.t[self.i - 1][2])
        if kind == "str":
            return ("str", text.strip('"'))
        if text == "(":
            e = self.expr()
            self.expect(")")
            return e
        if kind == "id":
            if self.accept("("):
                args = []
                if not self.accept(")"):
                    args.append(self.expr())
                    while self.accept(","):
                        args.append(self.expr())
                    self.expect(")")
                if (text in ("V", "I") and len(args) == 1 and args[0][0] == "var"
                        and args[0][1] in self.branches):
                    a, b = self.branches[args[0][1]]
                    args = [("var", a)] + ([("var", b)] if b is not None else [])
                return ("call", text, args, line)
            if text.startswith("$"):  # $temperature, $vt, ... without arguments
                return ("call", text, [])
            return ("var", text)
        raise VAError(f"unexpected {text!r} in an expression")


def parse(source: str) -> Module:
    return _Parser(source).module()


# =============================================================================
# Compiler: constant folding + SSA over duals -> register bytecode
# =============================================================================


def expr_is_int(e, types, mod, cmp_static=None):
    """Whether an expression has Verilog-A integer type: integer literals,
    variables and parameters, arithmetic on integers only, abs/min/max of
    integers, integer functions, and comparisons and logic on constants (on
    device-dependent values they are booleans, which arithmetic reads as
    real; compiler.py:600-662).  `cmp_static(e)` tells whether a comparison
    folds to a constant (None: assume it does)."""
    k = e[0]
    if k == "num":
        return bool(e[2])
    if k == "var":
        if e[1] in types:
            return types[e[1]] == "integer"
        p = mod.params.get(e[1])
        return p is not None and p.type == "integer"
    if k in ("+", "-", "*", "/", "%"):
        return (expr_is_int(e[1], types, mod, cmp_static)
                and expr_is_int(e[2], types, mod, cmp_static))
    if k == "neg":
        return expr_is_int(e[1], types, mod, cmp_static)
    if k in _CMP or k in ("&&", "||", "!"):
        return True if cmp_static is None else cmp_static(e)
    if k == "?":
        return (expr_is_int(e[2], types, mod, cmp_static)
                and expr_is_int(e[3], types, mod, cmp_static))
    if k == "call":
        base = e[1].lstrip("$")
        if base in ("abs", "min", "max"):
            return all(expr_is_int(a, types, mod, cmp_static) for a in e[2])
        f = mod.functions.get(e[1])
        return f is not None and f.type == "integer"
    return False


def strip_lines(e):
    """An expression without the source lines its calls carry."""
    if not isinstance(e, tuple):
        return e
    if e and e[0] == "call":
        return ("call", e[1], [strip_lines(a) for a in e[2]])
    return tuple(strip_lines(x) if isinstance(x, tuple) else x for x in e)


def clamp_pattern(s):
    """``if (x < b) x = b;`` (lower clamp: max) or ``if (x > b) x = b;`` (upper
    clamp: min), either operand order, no else branch: (x, "lower"|"upper",
    b) or None (voltax/va/compiler.py:1314-1336)."""
    if s[0] != "if" or s[3] != ("block", []):
        return None
    body = s[2]
    if body[0] == "block":
        if len(body[1]) != 1:
            return None
        body = body[1][0]
    cond = s[1]
    if body[0] != "assign" or cond[0] not in ("<", "<=", ">", ">="):
        return None
    x, rhs = body[1], strip_lines(body[2])
    op, a, b = cond
    if a == ("var", x) and strip_lines(b) == rhs:
        return x, ("lower" if op in ("<", "<=") else "upper"), b
    if b == ("var", x) and strip_lines(a) == rhs:
        return x, ("lower" if op in (">", ">=") else "upper"), a
    return None


def equality_side(cond):
    """"then" for ``if (a == b)`` (or a conjunction of equalities), "else"
    for ``if (a != b)`` (or a disjunction of inequalities): the arm taken on a
    set of measure zero (voltax/va/compiler.py:1297-1311); else None."""
    op = cond[0]
    if op == "==":
        return "then"
    if op == "!=":
        return "else"
    if op in ("&&", "||"):
        want = "then" if op == "&&" else "else"
        if equality_side(cond[1]) == equality_side(cond[2]) == want:
            return want
    return None


# $vt = k T / q with the constants the reference uses (compiler.py:781-783);
# default device temperature and simulator parameters (model.py:48, 208-221)
_KB_OVER_Q = 1.3806503e-23 / 1.602176462e-19
DEFAULT_TEMPERATURE = 300.15
SIMPARAMS = {"gmin": 1e-12}


class _Val:
    """A compiled value: a Python float (static) or an SSA index (dynamic)."""

    __slots__ = ("c", "r")

    def __init__(self, c=None, r=None):
        self.c, self.r = c, r

    @property
    def static(self):
        return self.r is None


@dataclass
class Compiled:
    module: str
    terminals: tuple
    internal: tuple     # internal nodes in use, then branch currents "i(a,b)"
    params: tuple       # parameter row: (name, default) of each real instance parameter,
                        # then the multiplicity ("m", 1.0)
    code: np.ndarray    # int32 words
    n_regs: int
    loop_sites: int     # unrolled loop iterations with a traced condition
    collapse: dict      # node -> representative ("gnd" for ground) of V(a,b) <+ 0


_GND = "<gnd>"


def _collapse_map(mod: Module, pairs):
    """Union of the node pairs shorted by static ``V(a, b) <+ 0``.  The
    representative of a set is ground if the set holds ground, else a port,
    else its first internal node (voltax/va/compiler.py:1388-1416)."""
    parent = {}

    def root(x):
        while parent.get(x, x) != x:
            x = parent[x]
        return x

    def weight(x):
        return 2 if (x == _GND or x in mod.grounds) else (1 if x in mod.ports else 0)

    for a, b in pairs:
        ra, rb = root(a), root(b)
        if ra == rb:
            continue
        wa, wb = weight(ra), weight(rb)
        if wa == wb and wa > 0:
            raise VAError(f"V({a},{b}) <+ 0 shorts two ports or ground (not supported)")
        if wb > wa:
            ra, rb = rb, ra
        parent[rb] = ra
    out = {}
    for n in mod.nodes:
        r = root(n)
        if r != n:
            out[n] = _GND if weight(r) == 2 else r
    return out


class _Compiler:
    """Partial evaluation of a module's analog block for one model card.

    Model-card parameters and integer parameters are constants folded into
    the program; real instance parameters are read per device (the reference
    traces the same set: voltax/va/model.py:208-221).  Statically decided
    conditions pick one side; data-dependent ones evaluate both sides and
    select per device (voltax/va/compiler.py:1004-1083)."""

    def __init__(self, mod: Module, card: dict, collapse=None, given=(),
                 temperature=None, simparams=None, traced=(), smooth=None,
                 equality="value"):
        self.mod = mod
        self.card = card             # physical value of every parameter
        self.given = set(given)      # parameters the model card sets ($param_given)
        self.temperature = DEFAULT_TEMPERATURE if temperature is None else float(temperature)
        self.simparams = {**SIMPARAMS, **(simparams or {})}
        self.smooth = smooth         # None, a width, or {source line: width}
        self.equality = equality     # "value" or "limit"
        self.types = mod.variables   # variable types of the current scope
        self.in_function = False
        self.collapse = dict(collapse or {})
        self.ins = []                # (op, a, b, c) with SSA operands
        self.bias = []               # per SSA value: depends on an unknown
        self._consts = {}
        self._unk_cache = {}
        self.loop_sites = 0
        # parameters read from the device's row: real instance parameters,
        # then the model parameters made differentiable
        self.inst = [n for n, p in mod.params.items() if p.instance and p.type == "real"]
        self.inst += [n for n in mod.params if n in set(traced) and n not in self.inst]
        self.used = set()            # internal nodes that are probed or contributed to
        self.branches = {}           # key -> (a, b) of voltage branches, first-use order
        self.vterms = {}             # key -> ([f terms], [q terms])
        self._arms = 0               # depth of limit-mode arms being compiled
        self.zero_v = []             # static V(a, b) <+ 0 pairs
        self.f, self.q = {}, {}      # unknown name -> residual / charge row

    # --------------------------------------------------------------- emit
    def emit(self, op, a=None, b=None, c=None):
        self.ins.append((op, a, b, c))
        if op == "unk":
            dep = True
        elif op in ("const", "param"):
            dep = False
        else:
            dep = any(self.bias[x] for x in (a, b, c) if x is not None)
        self.bias.append(dep)
        return _Val(r=len(self.ins) - 1)

    def is_bias(self, v: _Val):
        return not v.static and self.bias[v.r]

    def reg(self, v: _Val):
        if v.static:  # one SSA value per distinct constant
            key = struct.pack("<d", float(v.c))
            if key not in self._consts:
                self._consts[key] = self.emit("const", float(v.c))
            return self._consts[key]
        return v

    def op2(self, op, x: _Val, y: _Val):
        if x.static and y.static:
            return _Val(_FOLD[op](x.c, y.c))
        if op == "add":
            if x.static and x.c == 0:
                return y
            if y.static and y.c == 0:
                return x
        if op == "sub" and y.static and y.c == 0:
            return x
        if op == "sub" and x.static and x.c == 0:
            return self.emit("neg", y.r)
        if op == "mul":
            if (x.static and x.c == 0) or (y.static and y.c == 0):
                return _Val(0.0)
            if x.static and x.c == 1:
                return y
            if y.static and y.c == 1:
                return x
        if op == "div" and y.static and y.c == 1:
            return x
        return self.emit(op, self.reg(x).r, self.reg(y).r)

    def op1(self, op, x: _Val):
        if x.static:
            return _Val(_FOLD[op](x.c))
        return self.emit(op, x.r)

    def select(self, c: _Val, a: _Val, b: _Val):
        if c.static:
            return a if c.c else b
        if a.static and b.static and a.c == b.c:
            return a
        if not a.static and not b.static and a.r == b.r:
            return a
        return self.emit("select", c.r, self.reg(a).r, self.reg(b).r)

    # ---------------------------------------------------------- expressions
    def expr(self, e, env):
        """(f part, q part) of a contribution expression; the q part collects
        the arguments of ddt, which must enter linearly with factors that do
        not depend on the unknowns (voltax/va/compiler.py:1207-1246)."""
        if not _has_ddt(e):
            return self.value(e, env), _Val(0.0)
        kind = e[0]
        if kind == "call" and e[1] == "ddt":
            if len(e[2]) != 1:
                raise VAError("ddt() takes one argument here")
            return _Val(0.0), self.value(e[2][0], env)
        if kind in ("+", "-"):
            fa, qa = self.expr(e[1], env)
            fb, qb = self.expr(e[2], env)
            op = "add" if kind == "+" else "sub"
            return self.op2(op, fa, fb), self.op2(op, qa, qb)
        if kind == "neg":
            f, q = self.expr(e[1], env)
            return self.op1("neg", f), self.op1("neg", q)
        if kind == "*":
            if _has_ddt(e[1]) and _has_ddt(e[2]):
                raise VAError("product of two ddt() terms")
            kn, dn = (e[2], e[1]) if _has_ddt(e[1]) else (e[1], e[2])
            k = self.value(kn, env)
            if self.is_bias(k):
                raise VAError("ddt() multiplied by a bias-dependent factor (write ddt(k*q))")
            f, q = self.expr(dn, env)
            return self.op2("mul", k, f), self.op2("mul", k, q)
        if kind == "/" and not _has_ddt(e[2]):
            k = self.value(e[2], env)
            if self.is_bias(k):
                raise VAError("ddt() divided by a bias-dependent factor")
            f, q = self.expr(e[1], env)
            return self.op2("div", f, k), self.op2("div", q, k)
        raise VAError("unsupported use of ddt() (only linear combinations of ddt() terms)")

    def value(self, e, env):
        kind = e[0]
        if kind == "num":
            return _Val(float(e[1]))
        if kind == "var":
            name = e[1]
            if name in env:
                return env[name]
            if name in self.mod.params:
                return self.param(name)
            raise VAError(f"unknown name {name!r}")
        if kind in _ARITH or kind == "%":
            a, b = self.value(e[1], env), self.value(e[2], env)
            static = lambda x: self.value(x, env).static  # noqa: E731
            ints = kind in ("/", "%") and expr_is_int(e[1], self.types, self.mod, static) and \
                expr_is_int(e[2], self.types, self.mod, static)
            if kind == "%":  # a - b * trunc(a / b), the C remainder (fmod on reals)
                q = self.op1("trunc", self.op2("div", a, b))
                return self.op2("sub", a, self.op2("mul", b, q))
            if ints:  # integer division truncates (runtime.py:31-33)
                return self.op1("trunc", self.op2("div", a, b))
            return self.op2(_ARITH[kind], a, b)
        if kind == "neg":
            return self.op1("neg", self.value(e[1], env))
        if kind in _CMP:
            return self.op2(_CMP[kind], self.value(e[1], env), self.value(e[2], env))
        if kind in ("&&", "||"):
            a = self.value(e[1], env)
            if a.static and bool(a.c) == (kind == "||"):
                return _Val(float(kind == "||"))  # short circuit
            b = self.value(e[2], env)
            if a.static:
                return _Val(float(bool(b.c))) if b.static else self.op1("not", self.op1("not", b))
            return self.op2("and" if kind == "&&" else "or", a, b)
        if kind == "!":
            return self.op1("not", self.value(e[1], env))
        if kind == "?":
            return self.select(self.value(e[1], env), self.value(e[2], env),
                               self.value(e[3], env))
        if kind == "call":
            if e[1] == "ddt":
                raise VAError("ddt() outside the linear terms of a contribution")
            return self.call(e[1], e[2], env, e[3] if len(e) > 3 else None)
        if kind == "str":
            raise VAError(f"string {e[1]!r} used as a value")
        raise VAError(f"unsupported expression {kind!r}")

    def param(self, name):
        if name in self.inst:
            return self.emit("param", self.inst.index(name))
        v = float(self.card[name])
        if self.mod.params[name].type == "integer":
            v = float(_to_int(v))
        return _Val(v)

    # ----------------------------------------------------------- nodes
    def rep(self, n):
        if n == _GND or n in self.mod.grounds:
            return _GND
        if n not in self.mod.nodes:
            raise VAError(f"unknown node {n!r}")
        return self.collapse.get(n, n)

    def _ends(self, names):
        """(a, b, key) of an access V(a, b) / V(a); key names the branch."""
        if len(names) == 1:
            return names[0], _GND, f"{names[0]},"
        return names[0], names[1], f"{names[0]},{names[1]}"

    def unknown(self, name):
        if name not in self._unk_cache:
            self._unk_cache[name] = self.emit("unk", name)
        return self._unk_cache[name]

    def vnode(self, n):
        r = self.rep(n)
        if r == _GND:
            return _Val(0.0)
        self.used.add(r)
        return self.unknown(r)

    def probe_v(self, names):
        a, b, _ = self._ends(names)
        if b != _GND and self.rep(a) == self.rep(b):
            return _Val(0.0)
        return self.op2("sub", self.vnode(a), self.vnode(b))

    def probe_i(self, names):
        _, _, key = self._ends(names)
        if key not in self.branches:
            raise VAError(f"I({','.join(names)}) probe: only branches with voltage "
                          "contributions can be probed")
        return self.unknown(f"i({key})")

    # ------------------------------------------------------- smooth mode
    def width(self, line):
        """Smooth-mode width at source line `line` (None: exact)."""
        if isinstance(self.smooth, dict):
            return self.smooth.get(line)
        return self.smooth

    def softplus(self, u):
        """log(1 + e^u) = max(u, 0) + log(1 + e^-|u|)."""
        tail = self.op1("log", self.op2("add", _Val(1.0),
                                        self.op1("exp", self.op1("neg", self.op1("abs", u)))))
        return self.op2("add", self.op2("max", u, _Val(0.0)), tail)

    def smooth_minmax(self, which, a, b, w):
        """smax(a, b) = b + w softplus((a - b)/w), smin(a, b) = a - w
        softplus((a - b)/w): errors at most w ln 2 (voltax/va/runtime.py:76-90)."""
        sp = self.op2("mul", _Val(w),
                      self.softplus(self.op2("div", self.op2("sub", a, b), _Val(w))))
        return self.op2("add", b, sp) if which == "max" else self.op2("sub", a, sp)

    def smooth_abs(self, x, w):
        """w log(2 cosh(x/w)) = |x| + w log(1 + e^(-2|x|/w))."""
        ax = self.op1("abs", x)
        e = self.op1("exp", self.op2("mul", _Val(-2.0 / w), ax))
        return self.op2("add", ax, self.op2("mul", _Val(w),
                                            self.op1("log", self.op2("add", _Val(1.0), e))))

    def call(self, fn, args, env, line=None):
        if fn in ("V", "I"):
            names = [a[1] for a in args]
            if any(a[0] != "var" for a in args) or not 1 <= len(names) <= 2:
                raise VAError(f"bad probe {fn}(...)")
            return self.probe_v(names) if fn == "V" else self.probe_i(names)
        if fn in self.mod.functions:
            return self.inline(self.mod.functions[fn], args, env)
        if fn == "$temperature":
            return _Val(self.temperature)
        if fn == "$vt":
            temp = self.value(args[0], env) if args else _Val(self.temperature)
            return self.op2("mul", _Val(_KB_OVER_Q), temp)
        if fn == "$simparam":
            if not args or args[0][0] != "str":
                raise VAError("$simparam needs a parameter name")
            if args[0][1] in self.simparams:
                return _Val(float(self.simparams[args[0][1]]))
            if len(args) > 1:
                return self.value(args[1], env)
            raise VAError(f"$simparam({args[0][1]!r}) has no value or default")
        if fn == "$param_given":
            name = args[0][1] if args and args[0][0] == "var" else None
            if name not in self.mod.params:
                raise VAError(f"$param_given({name}): not a parameter")
            return _Val(float(name in self.given))
        if fn in ("$mfactor", "$port_connected"):
            return _Val(1.0)
        vals = [self.value(a, env) for a in args]
        base = fn[1:] if fn.startswith("$") else fn
        w = self.width(line)
        if w and base == "abs" and len(vals) == 1 and not vals[0].static:
            return self.smooth_abs(vals[0], w)
        if w and base in ("min", "max") and len(vals) == 2 and not (
                vals[0].static and vals[1].static):
            return self.smooth_minmax(base, vals[0], vals[1], w)
        if base in _UNARY and len(vals) == 1:
            if base == "limexp" and vals[0].static:  # folded as exp (compiler.py:897-900)
                return _Val(math.exp(vals[0].c))
            return self.op1(_UNARY[base], vals[0])
        if base in ("log", "log10") and len(vals) == 1:  # Verilog-A log is base 10
            return self.op2("mul", self.op1("log", vals[0]), _Val(1.0 / math.log(10.0)))
        if base in ("pow", "min", "max") and len(vals) == 2:
            return self.op2(base, *vals)
        raise VAError(f"unsupported function {fn}()")

    def inline(self, f, args, env):
        """An analog function call: the body runs on its own variables (inputs
        bound to the argument values) and returns the variable named after the
        function (voltax/va/compiler.py:909-938)."""
        if len(args) != len(f.inputs):
            raise VAError(f"{f.name}() takes {len(f.inputs)} arguments")
        local = {n: _Val(0.0) for n in f.variables}
        for n, a in zip(f.inputs, args):
            v = self.value(a, env)
            if f.variables.get(n) == "integer":
                v = _Val(float(_to_int(v.c))) if v.static else self.op1("round", v)
            local[n] = v
        saved = self.types, self.in_function
        self.types, self.in_function = f.variables, True
        try:
            self.run(f.body, local, None)
        finally:
            self.types, self.in_function = saved
        return local[f.name]

    # ----------------------------------------------------------- statements
    def run(self, stmts, env, cond):
        """Execute statements; `cond` is the path condition of data-dependent
        branches (None on the unconditional path).  As in the reference, each
        arm of a data-dependent `if` runs on its own copy of the variables,
        which are merged by a select afterwards, so that values assigned
        inside an arm stay constants there (compiler.py:1041-1083)."""
        for s in stmts:
            kind = s[0]
            if kind == "block":
                self.run(s[1], env, cond)
            elif kind == "assign":
                name = s[1]
                if name not in self.types:
                    raise VAError(f"assignment to undeclared {name!r}")
                v = self.value(s[2], env)
                if self.types[name] == "integer":  # real -> integer rounds (runtime.py:36-38)
                    if v.static:
                        v = _Val(float(_to_int(v.c)))
                    elif not expr_is_int(s[2], self.types, self.mod):
                        v = self.op1("round", v)
                env[name] = v
            elif kind == "if":
                c = self.value(s[1], env)
                if c.static:
                    self.run([s[2] if c.c else s[3]], env, cond)
                    continue
                clamp = clamp_pattern(s)
                w = self.width(s[4] if len(s) > 4 else None)
                if clamp and w and self.types.get(clamp[0]) == "real" and clamp[0] in env:
                    var, side, bound = clamp  # smooth clamp (compiler.py:1013-1027)
                    env[var] = self.smooth_minmax("max" if side == "lower" else "min",
                                                  env[var], self.value(bound, env), w)
                    continue
                nc = self.op1("not", c)
                env_then, env_else = dict(env), dict(env)
                side = equality_side(s[1]) if self.equality == "limit" else None
                if side is None:
                    self.run([s[2]], env_then, c if cond is None else self.op2("and", cond, c))
                    self.run([s[3]], env_else, nc if cond is None else self.op2("and", cond, nc))
                    self._merge(env, c, env_then, env_else)
                    continue
                # "limit": the special arm's values, the general arm's
                # derivatives where the special arm is taken (runtime.py:59-73).
                # Contributions are merged like variables (the reference keeps
                # them in the environment, compiler.py:1057-1083): each arm
                # accumulates into its own copy, guarded by the enclosing path
                # only, and the copies are merged by the arm condition.
                self._arms += 1
                acc0 = (self.f, self.q, self.vterms)
                arms = []
                for body, env_arm in ((s[2], env_then), (s[3], env_else)):
                    self.f, self.q = dict(acc0[0]), dict(acc0[1])
                    self.vterms = {k: (list(a), list(b)) for k, (a, b) in acc0[2].items()}
                    self.run([body], env_arm, cond)
                    arms.append((self.f, self.q, self.vterms))
                self._arms -= 1
                self.f, self.q, self.vterms = acc0
                on = c if side == "then" else nc
                gi = 1 if side == "then" else 0  # the general arm
                for k in set(env_then) | set(env_else):
                    env[k] = self._limit_merge(c, on, env_then.get(k, _Val(0.0)),
                                               env_else.get(k, _Val(0.0)), gi)
                for which in (0, 1):  # currents, charges
                    cur = self.f if which == 0 else self.q
                    at, ae = arms[0][which], arms[1][which]
                    for k in set(at) | set(ae):
                        cur[k] = self._limit_merge(c, on, at.get(k, _Val(0.0)),
                                                   ae.get(k, _Val(0.0)), gi)
                for key, (fl, ql) in acc0[2].items():  # voltage-branch terms
                    for which, lst in ((0, fl), (1, ql)):
                        n0 = len(lst)
                        parts = []
                        for arm in arms:
                            tot = _Val(0.0)
                            for t_ in arm[2][key][which][n0:]:
                                tot = self.op2("add", tot, t_)
                            parts.append(tot)
                        if any(not (p.static and p.c == 0) for p in parts):
                            lst.append(self._limit_merge(c, on, parts[0], parts[1], gi))
            elif kind == "while":
                self.loop(s, env, cond)
            elif kind == "contrib":
                if self.in_function:
                    raise VAError("contributions inside an analog function")
                if s[1] == "I":
                    self.contrib_i(s, env, cond)
                else:
                    self.contrib_v(s, env, cond)
            else:
                raise VAError(f"unsupported statement {kind!r}")

    def _limit_merge(self, c, on, a, b, gi):
        """select(c, a, b), with the derivatives of the general arm (a if
        gi == 0 else b) where `on` (the special arm) holds."""
        merged = self.select(c, a, b)
        g = (a, b)[gi]
        same = (merged.static and g.static and merged.c == g.c) or (
            not merged.static and not g.static and merged.r == g.r)
        return merged if same else self.emit("limit", on.r, self.reg(merged).r, self.reg(g).r)

    def _merge(self, env, c, env_then, env_else):
        for k in set(env_then) | set(env_else):
            a = env_then.get(k, _Val(0.0))
            b = env_else.get(k, _Val(0.0))
            env[k] = self.select(c, a, b)

    def loop(self, s, env, cond):
        """``while``, unrolled.  An iteration under a data-dependent condition
        runs, with the rest of the loop, as the taken arm of an ``if``
        (compiler.py:1137-1153); the nested arms are merged innermost first."""
        stack, cur = [], env
        for _ in range(MAX_UNROLL + 1):
            c = self.value(s[1], cur)
            if c.static and not c.c:
                break
            if not c.static:
                self.loop_sites += 1
                stack.append((c, cur))
                cur = dict(cur)
                cond = c if cond is None else self.op2("and", cond, c)
            self.run([s[2]], cur, cond)
        else:
            raise VAError(f"while loop does not end after {MAX_UNROLL} unrolled iterations "
                          "(its condition needs a part that folds to a constant)")
        for c, parent in reversed(stack):
            self._merge(parent, c, cur, parent)
            cur = parent

    def _guard(self, cond, f, q):
        if cond is None:
            return f, q
        return self.select(cond, f, _Val(0.0)), self.select(cond, q, _Val(0.0))

    def _acc(self, node, f, q, sign):
        op = "add" if sign > 0 else "sub"
        self.f[node] = self.op2(op, self.f.get(node, _Val(0.0)), f)
        self.q[node] = self.op2(op, self.q.get(node, _Val(0.0)), q)

    def contrib_i(self, s, env, cond):
        _, _, a, b, e = s
        names = [a] if b is None else [a, b]
        a, b, _ = self._ends(names)
        ra, rb = self.rep(a), self.rep(b)
        f, q = self._guard(cond, *self.expr(e, env))
        if ra == rb:
            return
        for node, sign in ((ra, 1), (rb, -1)):
            if node != _GND:
                self.used.add(node)
                self._acc(node, f, q, sign)

    def contrib_v(self, s, env, cond):
        _, _, a, b, e = s
        names = [a] if b is None else [a, b]
        a, b, key = self._ends(names)
        known = key in self.branches
        if not known:  # registered first: the right side may probe I(a, b)
            self.branches[key] = (a, b)
            self.vterms[key] = ([], [])
        f, q = self.expr(e, env)
        conditional = cond is not None or self._arms > 0
        if f.static and q.static and f.c == 0 and q.c == 0 and not known:
            del self.branches[key], self.vterms[key]
            if conditional:
                raise VAError("V(a, b) <+ 0 under a data-dependent condition is not supported")
            self.zero_v.append((a, b))
            return
        if conditional and not known:
            raise VAError("the first voltage contribution to a branch must not be under "
                          "a data-dependent condition")
        f, q = self._guard(cond, f, q)
        self.vterms[key][0].append(f)
        self.vterms[key][1].append(q)

    # ----------------------------------------------------------- assembly
    def compile(self):
        env = {n: _Val(0.0) for n in self.mod.variables}
        self.run(self.mod.body, env, None)
        # voltage branches: the branch current enters KCL at a (+) and b (-);
        # the branch row is V(a) - V(b) - sum(f terms) - d/dt sum(q terms)
        for key, (a, b) in self.branches.items():
            i_br = self.unknown(f"i({key})")
            for node, sign in ((self.rep(a), 1), (self.rep(b), -1)):
                if node != _GND:
                    self._acc(node, i_br, _Val(0.0), sign)
            fl, ql = self.vterms[key]
            f, q = self.probe_v([a] if b == _GND else [a, b]), _Val(0.0)
            for t in fl:
                f = self.op2("sub", f, t)
            for t in ql:
                q = self.op2("sub", q, t)
            self.f[f"i({key})"], self.q[f"i({key})"] = f, q
        ports = list(self.mod.ports)
        internal = [n for n in self.mod.nodes if n not in ports and n in self.used
                    and self.rep(n) == n]
        order = ports + internal + [f"i({k})" for k in self.branches]
        if len(order) > MAX_UNKNOWNS:
            raise VAError(f"{self.mod.name}: {len(order)} unknowns ({', '.join(order)}); at "
                          f"most {MAX_UNKNOWNS} per device are supported")
        return order

    def assemble(self, order):
        index = {n: i for i, n in enumerate(order)}
        # multiplicity m (the parameter after the instance parameters) scales
        # the terminal and internal-node rows, not the branch rows
        # (voltax/va/model.py:648-660)
        mult = self.emit("param", len(self.inst))
        rows = []
        for part in (self.f, self.q):
            for n in order:
                v = part.get(n, _Val(0.0))
                rows.append(v if n.startswith("i(") else self.op2("mul", v, mult))
        out_ids = [self.reg(v).r for v in rows]
        # liveness: last use of each SSA value, outputs live to the end
        n = len(self.ins)
        last = list(range(n))
        for i, (op, a, b, c) in enumerate(self.ins):
            if op in ("const", "param", "unk"):
                continue
            for x in (a, b, c):
                if x is not None:
                    last[x] = max(last[x], i)
        for o in out_ids:
            last[o] = n
        expire = {}
        for i, l in enumerate(last):
            expire.setdefault(l, []).append(i)
        # linear-scan register allocation; an operand whose last use is this
        # instruction frees its register before the destination is chosen
        # (the interpreter reads every operand before it writes)
        free, phys, n_regs, words = [], {}, 0, []
        for i, (op, a, b, c) in enumerate(self.ins):
            if op == "const":
                lo, hi = struct.unpack("<ii", struct.pack("<d", a))
                srcs = (lo, hi, 0)
            elif op == "param":
                srcs = (a, 0, 0)
            elif op == "unk":
                srcs = (index[a], 0, 0)
            else:
                srcs = tuple(phys[x] if x is not None else 0 for x in (a, b, c))
            for x in expire.get(i, []):
                if x != i and x in phys:
                    free.append(phys[x])
            if free:
                dst = free.pop()
            else:
                dst = n_regs
                n_regs += 1
            phys[i] = dst
            words += [OPS[op], dst, *srcs]
            if last[i] == i:  # never read: its register is free again at once
                free.append(dst)
        if n_regs > MAX_REGS:
            raise VAError(f"model needs {n_regs} registers; at most {MAX_REGS}")
        head = [MAGIC, len(order), n_regs, len(self.ins), len(self.inst) + 1,
                len(self.mod.ports)]
        tail = [phys[o] for o in out_ids]
        return np.asarray(head + words + tail, np.int32), n_regs


def _smooth_arg(smooth):
    """None, a positive width, or {source line: positive width}."""
    if smooth is None:
        return None
    if isinstance(smooth, dict):
        out = {int(k): float(v) for k, v in smooth.items()}
        if any(not v > 0 for v in out.values()):
            raise ValueError("smooth widths must be positive")
        return out
    if not float(smooth) > 0:
        raise ValueError("smooth width must be positive")
    return float(smooth)


def compile_module(mod: Module, card: dict, given=(), temperature=None,
                   simparams=None, traced=(), smooth=None, equality="value") -> Compiled:
    """Compile `mod` for the physical parameter values `card` (`given`: the
    names ``$param_given`` is true for; `temperature` in kelvin; `simparams`:
    ``$simparam`` values over the default ``gmin``; `traced`: real model
    parameters read from the device row, like instance parameters, so that
    they are differentiable; `smooth`: width of the smooth min / max / abs /
    clamp surrogates, one for all sites or {source line: width}; `equality`:
    "value" differentiates the arm taken, "limit" gives the special arm of an
    ``if (a == b)`` the general arm's derivatives), collapsing the nodes of
    static ``V(a, b) <+ 0`` contributions until a fixed point
    (voltax/va/compiler.py:1419-1434)."""
    if equality not in ("value", "limit"):
        raise ValueError(f"equality must be 'value' or 'limit', got {equality!r}")
    collapse = {}
    smooth = _smooth_arg(smooth)
    for _ in range(6):
        comp = _Compiler(mod, card, collapse, given, temperature, simparams, traced, smooth,
                         equality)
        order = comp.compile()
        new = _collapse_map(mod, comp.zero_v)
        if new == collapse:
            break
        collapse = new
    else:
        raise VAError("node collapsing did not reach a fixed point")
    code, n_regs = comp.assemble(order)
    nt = len(mod.ports)
    return Compiled(mod.name, tuple(order[:nt]), tuple(order[nt:]),
                    tuple((n, float(card[n])) for n in comp.inst) + (("m", 1.0),), code, n_regs,
                    comp.loop_sites, {k: ("gnd" if v == _GND else v) for k, v in collapse.items()})


def _has_ddt(e):
    if not isinstance(e, tuple) or not e:
        return False
    if e[0] == "call":
        return e[1] == "ddt" or any(_has_ddt(a) for a in e[2])
    if e[0] in ("num", "var"):
        return False
    return any(_has_ddt(x) for x in e[1:])


_ARITH = {"+": "add", "-": "sub", "*": "mul", "/": "div"}
_UNARY = {"exp": "exp", "ln": "log", "sqrt": "sqrt", "abs": "abs",
          "limexp": "limexp", "tanh": "tanh", "sin": "sin", "cos": "cos"}


def _to_int(x):
    return int(math.copysign(math.floor(abs(x) + 0.5), x))


def _np(fn):
    """IEEE semantics for folded constants (inf / NaN instead of exceptions)."""
    def run(*args):
        with np.errstate(all="ignore"):
            return float(fn(*(np.float64(a) for a in args)))
    return run


_FOLD = {
    "add": _np(np.add), "sub": _np(np.subtract), "mul": _np(np.multiply),
    "div": _np(np.divide), "pow": _np(np.power), "min": _np(np.minimum),
    "max": _np(np.maximum),
    "lt": lambda a, b: float(a < b), "le": lambda a, b: float(a <= b),
    "gt": lambda a, b: float(a > b), "ge": lambda a, b: float(a >= b),
    "eq": lambda a, b: float(a == b), "ne": lambda a, b: float(a != b),
    "and": lambda a, b: float(bool(a) and bool(b)), "or": lambda a, b: float(bool(a) or bool(b)),
    "neg": lambda a: -a, "exp": _np(np.exp), "log": _np(np.log), "sqrt": _np(np.sqrt),
    "abs": abs, "limexp": _np(np.exp), "tanh": _np(np.tanh), "sin": _np(np.sin),
    "cos": _np(np.cos), "not": lambda a: float(not a),
    "trunc": _np(np.trunc), "round": lambda a: float(_to_int(a)) if math.isfinite(a) else a,
}
_CMP = {"<": "lt", "<=": "le", ">": "gt", ">=": "ge", "==": "eq", "!=": "ne"}


# =============================================================================
# Models and elements
# =============================================================================


class VAModel:
    """A parsed Verilog-A module (reference: voltax/va/model.py:153-395)."""

    def __init__(self, module: Module):
        self.module = module
        self._lower = {n.lower(): n for n in module.params}
        for alias, target in module.aliases.items():  # aliasparam names
            if target not in module.params:
                raise VAError(f"aliasparam {alias} = {target}: not a parameter")
            self._lower.setdefault(alias.lower(), target)

    @property
    def name(self):
        return self.module.name

    @property
    def params(self):
        return self.module.params

    @property
    def instance_params(self):
        """Real instance parameters: per device, in parameter-row order."""
        return [n for n, p in self.module.params.items() if p.instance and p.type == "real"]

    def _const(self, e, vals):
        c = _Compiler(self.module, vals)
        c.inst = []  # every parameter is a constant here
        try:
            v = c.value(e, {"inf": _Val(math.inf)})
        except KeyError as err:
            raise VAError(f"{err.args[0]!r} used before its declaration") from None
        return None if not v.static else v.c

    def check_range(self, name, x, vals, what="model"):
        """ValueError if `x` lies outside the declared ranges of `name` or in
        one of its excluded ranges (voltax/va/model.py:480-503)."""
        p = self.module.params[name]
        for lo_b, lo, hi, hi_b in p.ranges:
            a, b = self._const(lo, vals), self._const(hi, vals)
            if a is None or b is None:
                continue
            if not ((x > a if lo_b == "(" else x >= a) and (x < b if hi_b == ")" else x <= b)):
                raise ValueError(f"{self.name}: {what} parameter {name} = {x:g} is outside "
                                 f"its range {lo_b}{a:g}:{b:g}{hi_b}")
        for ex in p.excludes:
            if isinstance(ex, tuple) and len(ex) == 4 and ex[0] in ("(", "["):
                lo_b, lo, hi, hi_b = ex
                a, b = self._const(lo, vals), self._const(hi, vals)
                if a is None or b is None:
                    continue
                if (x > a if lo_b == "(" else x >= a) and (x < b if hi_b == ")" else x <= b):
                    raise ValueError(f"{self.name}: {what} parameter {name} = {x:g} is in "
                                     f"the excluded range {lo_b}{a:g}:{b:g}{hi_b}")
            elif self._const(ex, vals) == x:
                raise ValueError(f"{self.name}: {what} parameter {name} = {x:g} is excluded")

    def _given(self, card):
        return {self._lower[k.lower()] for k in (card or {}) if k.lower() in self._lower}

    def normalize(self, values):
        """Case-insensitive parameter names -> declared names; unknown names
        are dropped with a warning, as the reference does."""
        out, unknown = {}, []
        for k, v in (values or {}).items():
            if k.lower() in self._lower:
                out[self._lower[k.lower()]] = v
            else:
                unknown.append(k)
        if unknown:
            import warnings

            warnings.warn(f"{self.name}: unknown parameters {sorted(unknown)}", UserWarning,
                          stacklevel=3)
        return out

    def defaults(self, card=None, ranges="error"):
        """Physical value of every parameter: the card's or the default, which
        may use parameters declared before it.  Card values outside a declared
        range raise (``ranges="error"``), warn (``"warn"``) or pass
        (``"ignore"``), as in voltax/va/model.py:480-503."""
        if ranges not in ("error", "warn", "ignore"):
            raise ValueError(f"ranges must be 'error', 'warn' or 'ignore', got {ranges!r}")
        given = {k: float(v) for k, v in self.normalize(card).items()}
        out = {}
        for name, p in self.module.params.items():
            if name in given:
                out[name] = given[name]
                if p.type == "integer" and out[name] != _to_int(out[name]):
                    raise ValueError(f"{self.name}: integer parameter {name} = "
                                     f"{out[name]:g} is not integral")
                if ranges == "error":
                    self.check_range(name, out[name], out)
                elif ranges == "warn":
                    try:
                        self.check_range(name, out[name], out)
                    except ValueError as err:
                        import warnings

                        warnings.warn(str(err), UserWarning, stacklevel=3)
                continue
            v = self._const(p.default, out)
            if v is None:
                raise VAError(f"default of {name!r} is not a constant")
            out[name] = v
        return out

    def differentiable(self, names):
        """Real model parameters named by `names` ("all": every real model
        parameter), validated (voltax/va/model.py:241-262)."""
        if isinstance(names, str):
            if names != "all":
                raise ValueError(f"differentiable must be 'all' or a list of names, got "
                                 f"{names!r}")
            return tuple(n for n, p in self.params.items()
                         if p.type == "real" and not p.instance)
        out = []
        for k in names:
            n = self._lower.get(k.lower())
            if n is None:
                raise ValueError(f"{self.name}: no parameter {k!r}")
            p = self.params[n]
            if p.type != "real" or p.instance:
                raise ValueError(f"{self.name}: {n} is not a real model parameter")
            out.append(n)
        return tuple(n for n in self.params if n in out)

    def compile(self, card=None, *, instance=None, temperature=None,
                simparams=None, differentiable=(), smooth=None,
                equality="value") -> Compiled:
        """The program for model card `card`; `instance` values of real
        instance parameters become the program's default parameter row
        (voltax/va/model.py:199-226); `differentiable` model parameters are
        read from the row after the instance parameters."""
        vals = self.defaults(card)
        inst = self.normalize(instance)
        for k, v in inst.items():
            if k not in self.instance_params:
                raise ValueError(f"{self.name}: not a real instance parameter: {k}")
            self.check_range(k, float(v), vals, "instance")
            vals[k] = float(v)
        return compile_module(self.module, vals, self._given(card) | set(inst),
                              temperature, simparams, self.differentiable(differentiable),
                              smooth, equality)

    def element(self, card=None, name=None, polarity=None, *, temperature=None,
                simparams=None, ranges="error", defaults=None, force=None, differentiable=(),
                smooth=None, equality="value"):
        """An element class for model card `card` (voltax/va/model.py:228-327).

        Devices are made as ``E(nodes, m=1, **instance)``; real instance
        parameters are per device (physical units in `get`/`set`; positive
        ones, ``from (0:...``, are differentiated in log space as
        ``log_<name>``), ``m`` is the multiplicity.  ``polarity`` "n" / "p"
        sets the module's ``type`` parameter to +1 / -1 unless the card gives
        it, and suffixes the group name (``<name>_n``).  `temperature` is the
        device temperature in kelvin (default 300.15), `simparams` the
        ``$simparam`` values (default ``gmin=1e-12``), `ranges` what a card
        value outside its declared range does ("error", "warn", "ignore"),
        `defaults` values for parameters the card does not set and `force`
        values that override the card.  `differentiable` real model
        parameters ("all" or names) join the device row after the instance
        parameters, at the card's value for every device, so the adjoint
        returns their gradients per device (``log_<name>`` for positive ones;
        the model gradient is the sum over the group's devices).  `smooth`
        replaces min / max / abs and ``if (x < b) x = b`` clamps on device-
        dependent values by smooth surrogates of that width (one width, or
        {source line: width}; errors at most width x ln 2).  `equality`
        "limit" keeps the values of ``if (a == b)`` special cases but
        differentiates the general arm there (SPICE-like Jacobians; "value",
        the default, differentiates the arm taken)."""
        from .circuit import Element, KIND_VA_BASE

        card = {**self.normalize(defaults), **self.normalize(card), **self.normalize(force)}
        if polarity is not None:
            if polarity not in ("n", "p"):
                raise ValueError(f"polarity must be 'n' or 'p', got {polarity!r}")
            tname = self._lower.get("type")
            if tname is not None and tname not in card:
                card[tname] = 1 if polarity == "n" else -1
        vals = self.defaults(card, ranges)
        compiled = compile_module(self.module, vals, self._given(card), temperature, simparams,
                                  self.differentiable(differentiable), smooth, equality)
        model = self
        insts = self.instance_params
        row_names = [k for k, _ in compiled.params]  # instance, traced model params, "m"
        grad_names = tuple(f"log_{k}" if self.module.params[k].positive else k
                           for k in row_names[:-1]) + ("m",)

        class VAElement(Element):
            kind = KIND_VA_BASE + len(compiled.terminals) + len(compiled.internal) - 2
            terminals = compiled.terminals
            params = grad_names
            n_internal = len(compiled.internal)
            internal_names = compiled.internal
            va = compiled
            va_model = model
            card = vals
            polarity_label = polarity
            device_temperature = DEFAULT_TEMPERATURE if temperature is None else float(temperature)

            def __init__(self, nodes, m=1.0, **instance):
                self.nodes = np.asarray(nodes, dtype=object).reshape(-1, len(self.terminals))
                n = self.nodes.shape[0]
                given = model.normalize(instance)
                bad = set(given) - set(insts)
                if bad:
                    raise ValueError(f"{compiled.module}: not real instance parameters: "
                                     f"{sorted(bad)}")
                self.rows = np.tile(np.array([v for _, v in compiled.params], float), (n, 1))
                for k, v in given.items():
                    col = np.broadcast_to(np.asarray(v, float), (n,))
                    for x in np.unique(col):
                        model.check_range(k, float(x), vals, "instance")
                    self.rows[:, row_names.index(k)] = col
                self.rows[:, -1] = np.broadcast_to(np.asarray(m, float), (n,))

            @classmethod
            def from_rows(cls, nodes, rows):
                """A group on resolved node indices with given parameter rows."""
                el = cls.__new__(cls)
                el.nodes = np.asarray(nodes, np.int64).reshape(-1, len(cls.terminals))
                el.rows = np.asarray(rows, float).reshape(el.nodes.shape[0], -1)
                return el

            @property
            def size(self):
                return self.rows.shape[0]

            def param_rows(self):
                return self.rows

            def _col(self, field):
                key = field.lower()
                for k in row_names:
                    if k.lower() == key:
                        return row_names.index(k)
                raise AttributeError(f"{type(self).__name__} has no real instance "
                                     f"parameter {field!r}")

            def get(self, field, index=slice(None)):
                return self.rows[index, self._col(field)]

            def set(self, index, **values):
                for k, v in values.items():
                    self.rows[index, self._col(k)] = v

            @staticmethod
            def grad_from_rows(grad_rows, rows):
                """Gradient per physical row entry -> {grad name: (N,)} (log-space
                entries scaled by the physical value)."""
                out = {}
                for j, (gname, k) in enumerate(zip(grad_names, row_names)):
                    out[gname] = grad_rows[:, j] * (rows[:, j] if gname.startswith("log_")
                                                    else 1.0)
                return out

        VAElement.__name__ = name or compiled.module
        return VAElement


def level_models(model: VAModel, level: int = 54):
    """``{"nmos:<level>": f, "pmos:<level>": f}`` for `parse_netlist`'s
    ``models=`` hook (voltax/va/netlist.py:19-75): devices of one model card
    share one element class (one group, named ``<card>_<n|p>``); instance
    parameters the module does not declare are dropped with one warning
    each, ``m`` is the multiplicity."""
    classes, warned = {}, set()

    def factory(card_name, kind, params, nodes, instance):
        kind = kind.lower()
        if kind not in ("nmos", "pmos"):
            raise ValueError(f"model {card_name!r}: Verilog-A MOSFET cards must be nmos or "
                             f"pmos, got {kind!r}")
        key = (card_name.lower(), kind)
        if key not in classes:
            card = {k: v for k, v in params.items() if k != "level"}
            if "level" in model._lower:
                card["level"] = params.get("level")
            known = {k.lower() for k in model.params}
            unknown = [k for k in card if k.lower() not in known]
            if unknown:
                import warnings

                warnings.warn(f"{model.name}: model card {card_name}: parameters "
                              f"{sorted(unknown)} are not declared by the module; ignored",
                              UserWarning, stacklevel=3)
            card = {k: v for k, v in card.items() if k.lower() in known}
            classes[key] = model.element(card, name=card_name, polarity=kind[0])
        cls = classes[key]
        if len(nodes) != len(cls.terminals):
            raise ValueError(f"{model.name} has {len(cls.terminals)} terminals, got "
                             f"{len(nodes)} nodes")
        inst = {}
        declared = {k.lower() for k in model.instance_params}
        for k, v in instance.items():
            if k.lower() in declared or k.lower() == "m":
                inst[k] = v
            elif k not in warned:
                warned.add(k)
                import warnings

                warnings.warn(f"{model.name}: instance parameter {k!r} is not declared by the "
                              "module; ignored", UserWarning, stacklevel=3)
        return cls(tuple(nodes), **inst)

    return {f"nmos:{int(level)}": factory, f"pmos:{int(level)}": factory}


def load(source: str) -> VAModel:
    """Parse Verilog-A source text (one module)."""
    return VAModel(parse(source))
