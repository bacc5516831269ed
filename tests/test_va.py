"""Verilog-A subset on the host: layouts and compiled programs against the
oracle's direct interpreter, compile-time errors, circuit tables, and the
C-ABI's program checks (no GPU needed)."""

import ctypes as C
import math
import struct

import numpy as np
import pytest

import va_models as vm
from oracle import analysis as oa
from oracle import circuit as oc
from oracle import dual as D
from oracle import va as ova

CASES = [("res", None), ("diode", None), ("diode", {"rs": 10.0}), ("rl", None),
         ("sumk", None), ("heron", None), ("sqnmos", None), ("vcvs", None), ("mixfn", None),
         ("guarded", None), ("iter", None), ("lang", None), ("lang", {"mode": 0, "tc": 2e-3})]

EXACT = dict(rtol=1e-12, atol=1e-15, gmin=0.0, gmin_steps=0)  # closed forms have no gmin term

OPS = ("const param unk add sub mul div neg exp log sqrt abs pow min max limexp tanh sin cos "
       "lt le gt ge eq ne and or not select trunc round limit").split()


def run_program(code, p, u, cf=1.0, cq=0.0):
    """Host restatement of the GPU interpreter (csrc/va_eval.hpp) and of the
    generated code (csrc/va_codegen.cpp) for one device: returns (res[m],
    jac[m][m]) of cf*f + cq*q."""
    m, nregs, nins = int(code[1]), int(code[2]), int(code[3])
    R = np.zeros((nregs, 1 + m))
    for k in range(nins):
        op, dst, a, b, c = (int(x) for x in code[6 + 5 * k: 11 + 5 * k])
        name = OPS[op]
        g = np.zeros(m)
        with np.errstate(all="ignore"):
            if name == "const":
                v = struct.unpack("<d", struct.pack("<ii", a, b))[0]
            elif name == "param":
                v = p[a]
            elif name == "unk":
                v = u[a]
                g[a] = 1.0
            else:
                A, B = R[a], R[b]
                x, y = A[0], B[0]
                dA, dB = A[1:], B[1:]
                if name == "add":
                    v, g = x + y, dA + dB
                elif name == "sub":
                    v, g = x - y, dA - dB
                elif name == "mul":
                    v, g = x * y, dA * y + x * dB
                elif name == "div":
                    v = x / y
                    g = (dA - v * dB) / y
                elif name == "neg":
                    v, g = -x, -dA
                elif name == "exp":
                    v = math.exp(x)
                    g = v * dA
                elif name == "log":
                    v, g = np.log(x), dA / x
                elif name == "sqrt":
                    v = np.sqrt(x)
                    g = dA / (2 * v)
                elif name == "abs":
                    v, g = abs(x), (-dA if x < 0 else dA)
                elif name == "pow":
                    v = np.power(x, y)
                    g = y * np.power(x, y - 1) * dA + np.where(dB != 0, v * np.log(x) * dB, 0.0)
                elif name in ("min", "max"):
                    pick = x <= y if name == "min" else x >= y
                    v, g = (x, dA) if pick else (y, dB)
                elif name == "limexp":
                    e = math.exp(min(x, 80.0))
                    v = e * (1 + x - 80.0) if x > 80.0 else e
                    g = e * dA
                elif name == "tanh":
                    v = math.tanh(x)
                    g = (1 - v * v) * dA
                elif name == "sin":
                    v, g = math.sin(x), math.cos(x) * dA
                elif name == "cos":
                    v, g = math.cos(x), -math.sin(x) * dA
                elif name == "select":
                    Cr = R[c]
                    v, g = (y, dB) if x != 0 else (Cr[0], Cr[1:])
                elif name == "limit":
                    Cr = R[c]
                    v, g = y, (Cr[1:] if x != 0 and math.isfinite(Cr[0]) else dB)
                elif name == "trunc":
                    v = math.trunc(x) if math.isfinite(x) else x
                elif name == "round":
                    v = math.copysign(math.floor(abs(x) + 0.5), x) if math.isfinite(x) else x
                else:
                    v = float({"lt": x < y, "le": x <= y, "gt": x > y, "ge": x >= y,
                               "eq": x == y, "ne": x != y, "and": x != 0 and y != 0,
                               "or": x != 0 or y != 0, "not": x == 0}[name])
        R[dst, 0] = v
        R[dst, 1:] = g
    out = code[6 + 5 * nins:]
    res = np.array([cf * R[out[i], 0] + cq * R[out[m + i], 0] for i in range(m)])
    jac = np.array([[cf * R[out[i], 1 + j] + cq * R[out[m + i], 1 + j] for j in range(m)]
                    for i in range(m)])
    return res, jac


def _points(name, rng, n):
    """Random unknown values (n, m) away from region boundaries."""
    m = {"res": 2, "diode": 2, "rl": 4, "sumk": 2, "heron": 2, "sqnmos": 3, "vcvs": 3,
         "mixfn": 2, "guarded": 2, "iter": 2, "lang": 3}[name]
    if name == "diode":
        return rng.uniform(-0.3, 0.8, (n, 4))
    return rng.uniform(-1.5, 1.5, (n, m + 1))


@pytest.mark.parametrize("name,card", CASES)
def test_layout_matches_oracle(name, card):
    pair = vm.Pair(vm.ALL[name], card)
    va = pair.product.va
    assert va.terminals + va.internal == tuple(pair.oracle.order)
    assert pair.product.kind == 5 + len(pair.oracle.order) - 2


def test_collapse_and_branch_names():
    p0 = vm.Pair(vm.DIODE).product.va
    assert p0.internal == () and p0.collapse == {"ai": "a"}
    p1 = vm.Pair(vm.DIODE, {"RS": 10}).product.va
    assert p1.internal == ("ai",) and p1.collapse == {}
    assert vm.Pair(vm.RL).product.va.internal == ("mid", "i(mid,n)")
    assert vm.Pair(vm.VCVS).product.va.internal == ("i(o,)",)


@pytest.mark.parametrize("name,card", CASES)
def test_program_matches_oracle(name, card):
    """Values and Jacobians of the compiled program against the oracle's
    direct interpretation of the parse tree, for cf*f + cq*q."""
    pair = vm.Pair(vm.ALL[name], card)
    code = pair.product.va.code
    model = pair.oracle
    m = len(model.order)
    rng = np.random.default_rng(len(name))
    U = _points(name, rng, 24)[:, :m]
    T = len(model.terminals)
    seeds = [D.Dual.seed(U[:, j], j, m) for j in range(m)]
    params = np.array([v for _, v in pair.product.va.params])
    params[-1] = 1.5  # multiplicity
    stored = {sk: np.full(len(U), np.log(model.card[k]) if sk.startswith("log_") else
                          model.card[k]) for k, sk in zip(model.inst, model.stored)}
    F, Q = ova.group_rows(model, seeds[:T], seeds[T:], {**stored, "m": np.full(len(U), 1.5)})
    for cf, cq in ((1.0, 0.0), (0.0, 1.0), (1.0, 1e9)):
        for i in range(len(U)):
            res, jac = run_program(code, params, U[i], cf, cq)
            ref = np.array([cf * D.val(F[r])[i] + cq * D.val(Q[r])[i] for r in range(m)])
            rjac = np.array([cf * D.tan(F[r], (len(U), m))[i] + cq * D.tan(Q[r], (len(U), m))[i]
                             for r in range(m)])
            np.testing.assert_allclose(res, ref, rtol=1e-13, atol=1e-300)
            np.testing.assert_allclose(jac, rjac, rtol=1e-12, atol=1e-14 * max(1.0, np.abs(rjac).max()))


def test_constant_loop_is_folded():
    """sum_{k=1..6} 2^-k = 1 - 2^-6 folds to one constant."""
    va = vm.Pair(vm.SUMK).product.va
    assert va.loop_sites == 0
    res, jac = run_program(va.code, np.ones(1), np.array([1.0, 0.0]))
    assert res[0] == 1 - 2.0 ** -6 and jac[0, 0] == 1 - 2.0 ** -6


def test_data_dependent_loop():
    """Heron's iteration stops per device; the result is sqrt(1 + v^2)."""
    va = vm.Pair(vm.HERON).product.va
    assert va.loop_sites == 8
    for v in (0.0, 0.3, -1.2, 3.0):
        res, jac = run_program(va.code, np.ones(1), np.array([v, 0.0]))
        assert res[0] == pytest.approx(1e-3 * (math.sqrt(1 + v * v) - 1), rel=1e-13, abs=1e-18)
        assert jac[0, 0] == pytest.approx(1e-3 * v / math.sqrt(1 + v * v), rel=1e-9, abs=1e-15)


def test_reference_loop_values():
    """The reference's loop fixture: v = 1 runs all six iterations, v = 0.01
    exits after four; five traced loop tests."""
    va = vm.Pair(vm.ITER).product.va
    assert va.loop_sites == 5
    res, jac = run_program(va.code, np.ones(1), np.array([1.0, 0.0]))
    assert res[0] == pytest.approx(1 - 2.0 ** -6, rel=1e-15) and jac[0, 0] == pytest.approx(1 - 2.0 ** -6)
    res, _ = run_program(va.code, np.ones(1), np.array([0.01, 0.0]))
    assert res[0] == pytest.approx(0.01 * (1 - 2.0 ** -4), rel=1e-15)


def test_guarded_branch_derivatives_are_finite():
    va = vm.Pair(vm.GUARDED).product.va
    _, jac = run_program(va.code, np.ones(1), np.array([-0.5, 0.0]))
    assert jac[0, 0] == pytest.approx(1e-3)
    _, jac = run_program(va.code, np.ones(1), np.array([0.25, 0.0]))
    assert jac[0, 0] == pytest.approx(1e-3 * 0.5 / 0.5)


def test_oracle_instance_parameter_gradient():
    """d V(b) / d log r2 of the divider = 2 r1 r2 / (r1 + r2)^2
    (tests/test_va.py:148-162 of the reference)."""
    res = vm.Pair(vm.RES)
    b = oc.CircuitBuilder()
    b.vsource("a", "0", 2.0, name="V1")
    vm.place(b, res, ("a", "b"), name="R1", r=1e3)
    vm.place(b, res, ("b", "0"), name="R2", r=3e3)
    c = b.build()
    sol = oa.dc(c, options=oa.Options(**EXACT))
    g = np.zeros(c.size)
    g[c.node("b")] = 1.0
    gr = oa.dc_vjp(c, sol.z, 0.0, g, options=oa.Options(**EXACT))["res"]
    assert gr["log_r"][1] == pytest.approx(2.0 * 1e3 * 3e3 / 4e3 ** 2, rel=1e-12)
    assert gr["m"][0] == pytest.approx(-gr["log_r"][0], rel=1e-12)  # m scales 1/r


BAD = [
    ("analog I(p, n) <+ V(p, n) * ddt(V(p, n));", "bias-dependent"),
    ("analog I(p, n) <+ ddt(V(p, n)) * ddt(V(p, n));", "two ddt"),
    ("real x; analog begin x = ddt(V(p, n)); I(p, n) <+ x; end", "ddt"),
    ("analog I(p, n) <+ I(p, n);", "voltage contribution"),
    ("analog I(p, q) <+ V(p, n);", "unknown node"),
    ("real x; analog begin x = 0; while (V(p, n) > x) x = x + 1; I(p, n) <+ x; end",
     "does not end"),
    ("analog begin if (V(p, n) > 0) V(p, n) <+ 0; end", "data-dependent"),
    ("analog begin if (V(p, n) > 0) V(p, n) <+ 1; end", "data-dependent"),
    ("analog I(p, n) <+ foo(V(p, n));", "unsupported function"),
]


@pytest.mark.parametrize("body,msg", BAD)
def test_compile_errors(body, msg):
    from voltax_amd import va

    src = f"module bad(p, n); inout p, n; electrical p, n; {body} endmodule"
    with pytest.raises(va.VAError, match=msg):
        va.load(src).compile()


def test_too_many_unknowns_and_unknown_card_parameter():
    from voltax_amd import va

    src = ("module big(a, b, c, d); inout a, b, c, d; electrical a, b, c, d, x; "
           "analog begin I(a, x) <+ V(a, x); I(x, b) <+ V(x, b); I(c, d) <+ V(c, d); end "
           "endmodule")
    c = va.load(src).compile()  # 5 unknowns: generated code on the GPU
    assert c.terminals + c.internal == ("a", "b", "c", "d", "x")
    nodes = [f"x{k}" for k in range(15)]
    chain = ["a"] + nodes + ["b"]
    body = " ".join(f"I({p}, {q}) <+ V({p}, {q});" for p, q in zip(chain[:-1], chain[1:]))
    src17 = (f"module huge(a, b); inout a, b; electrical a, b, {', '.join(nodes)}; "
             f"analog begin {body} end endmodule")
    with pytest.raises(va.VAError, match="at most 16"):
        va.load(src17).compile()
    # as in the reference (tests/test_va.py:389-392): unknown names warn,
    # values outside a declared range raise
    with pytest.warns(UserWarning, match="unknown parameters"):
        va.load(vm.RES).element({"rr": 1.0})
    with pytest.raises(ValueError, match="outside its range"):
        va.load(vm.DIODE).element({"rs": -1.0})
    R = va.load(vm.RES).element()
    with pytest.raises(ValueError, match="outside its range"):
        R(("a", "b"), r=0.0)
    dev = R([("a", "b"), ("b", "c")], r=[1e3, 2e3], m=[1, 3])
    assert R.params == ("log_r", "m")
    assert list(dev.get("R")) == [1e3, 2e3] and list(dev.get("m")) == [1, 3]


def test_circuit_tables_and_builder():
    import voltax_amd as vx

    res = vm.Pair(vm.RES)
    rl = vm.Pair(vm.RL)
    b = vx.CircuitBuilder()
    b.vsource("a", "0", 1.0, name="V1")
    vm.place(b, res, ("a", "b"), r=2e3)
    vm.place(b, res, ("b", "0"), name="Rlow")
    vm.place(b, rl, ("b", "0"))
    c = b.build()
    assert c.has_va
    kind, wave, unk, poff, row = c.device_table()
    code, offs = c.va_code()
    assert sorted(offs) == ["res", "rl"]
    assert list(kind) == [2, 5, 5, 7]  # groups in sorted name order: V, res, rl
    assert wave[1] == wave[2] == offs["res"] and wave[3] == offs["rl"]
    assert code[offs["rl"]] == 0x5642 and code[offs["res"]] == 0x5642
    np.testing.assert_array_equal(row[poff[1:3]], [2e3, 1e3])
    assert c.get("Rlow", "r") == 1e3
    # the RL device: terminals b, ground, then its internal node and branch current
    assert list(unk[3][:2]) == [c.node("b"), -1]
    assert c.layout.internal("X2", "i(mid,n)") == unk[3][3]
    # a different card changes the program and therefore the topology key
    b2 = vx.CircuitBuilder()
    b2.vsource("a", "0", 1.0, name="V1")
    vm.place(b2, vm.Pair(vm.RES.replace("1k", "2k")), ("a", "0"))
    b3 = vx.CircuitBuilder()
    b3.vsource("a", "0", 1.0, name="V1")
    vm.place(b3, res, ("a", "0"))
    assert b2.build().topology_key() == b3.build().topology_key()  # instance default only
    rl2 = vm.Pair(vm.RL, {"r": 20.0})
    b4 = vx.CircuitBuilder()
    vm.place(b4, rl2, ("a", "0"))
    b5 = vx.CircuitBuilder()
    vm.place(b5, rl, ("a", "0"))
    assert b4.build().topology_key() != b5.build().topology_key()


def test_capi_checks_programs():
    """vxm_topology_set_code accepts the compiled program and rejects
    corrupted ones (host-side checks; no GPU involved)."""
    from voltax_amd import engine
    from voltax_amd._lib import lib

    pair = vm.Pair(vm.DIODE, {"rs": 10.0})
    code = pair.product.va.code
    m = code[1]
    unk = np.array([[0, -1, 1, -1]], np.int32)
    good = dict(size=2, n_nodes=1, kind=np.array([5 + m - 2], np.int32),
                wave=np.array([0], np.int32), unknowns=unk, poff=np.array([0], np.int32),
                n_params=2)  # area, m

    def attempt(words, **kw):
        args = {**good, **kw}
        t = engine.Topology(args["size"], args["n_nodes"], args["kind"], args["wave"],
                            args["unknowns"], args["poff"], args["n_params"])
        w = np.ascontiguousarray(words, np.int32)
        return lib.vxm_topology_set_code(t.handle, len(w), w.ctypes.data)

    assert attempt(code) == 0
    bad = code.copy()
    bad[0] = 0
    assert attempt(bad) != 0                       # magic
    assert attempt(code, kind=np.array([5], np.int32)) != 0  # m differs from the kind
    assert attempt(code, n_params=1) != 0          # parameter row too short
    assert attempt(code[:-1]) != 0                 # truncated
    nins = code[3]
    for field, value in ((0, 32), (1, code[2]), (2, 99)):
        bad = code.copy()
        k = next(i for i in range(nins) if code[6 + 5 * i] >= 3)  # first arithmetic op
        bad[6 + 5 * k + field] = value
        assert attempt(bad) != 0, field
    bad = code.copy()
    bad[-1] = code[2]
    assert attempt(bad) != 0                       # output register out of range


def test_oracle_closed_forms():
    """The oracle against hand solutions: divider, diode with series
    resistance (scalar root), VCVS gain, square-law inverter."""
    from scipy.optimize import brentq

    res = vm.Pair(vm.RES)
    b = oc.CircuitBuilder()
    b.vsource("a", "0", 3.0, name="V1")
    vm.place(b, res, ("a", "b"))
    vm.place(b, res, ("b", "0"), r=2e3)
    z = oa.dc(b.build(), options=oa.Options(**EXACT)).z
    assert z[1] == pytest.approx(2.0, rel=1e-14)

    for rs in (0.0, 10.0):
        d = vm.Pair(vm.DIODE, {"rs": rs})
        b = oc.CircuitBuilder()
        b.vsource("in", "0", 5.0, name="V1")
        b.resistor("in", "a", 1e3, name="R1")
        vm.place(b, d, ("a", "0"), area=2.0)
        c = b.build()
        z = oa.dc(c, options=oa.Options(**EXACT)).z

        def kcl(vj):
            i = 2e-14 * (math.exp(vj / 0.025852) - 1)
            return (5 - vj - rs * i) / 1e3 - i

        vj = brentq(kcl, 0.0, 1.0, xtol=1e-16)
        i = 2e-14 * (math.exp(vj / 0.025852) - 1)
        assert z[c.node("a")] == pytest.approx(vj + rs * i, rel=1e-11)

    b = oc.CircuitBuilder()
    b.vsource("i", "0", 0.3, name="V1")
    vm.place(b, vm.Pair(vm.VCVS), ("i", "o"))
    b.resistor("o", "0", 1e3, name="RL")
    c = b.build()
    z = oa.dc(c, options=oa.Options(**EXACT)).z
    assert z[c.node("o")] == pytest.approx(1.2, rel=1e-14)

    b = oc.CircuitBuilder()
    b.vsource("vdd", "0", 1.8, name="Vdd")
    b.vsource("in", "0", 1.2, name="Vin")
    b.resistor("vdd", "out", 10e3, name="RL")
    vm.place(b, vm.Pair(vm.SQNMOS), ("out", "in", "0"), w=2.0)
    c = b.build()
    z = oa.dc(c, options=oa.Options(**EXACT)).z

    def kcl(vd):  # triode region: vds < vov = 0.7
        return (1.8 - vd) / 10e3 - 2 * 2e-3 * (0.7 - 0.5 * vd) * vd * (1 + 0.02 * vd)

    vd = brentq(kcl, 0.0, 0.7, xtol=1e-16)
    assert z[c.node("out")] == pytest.approx(vd, rel=1e-11)


def test_oracle_rl_step():
    """Inductor current of the RL model under a voltage step, backward Euler:
    i_k = (i_{k-1} + h V / L) / (1 + h R / L), and close to the exact
    V/R (1 - exp(-t R/L))."""
    rl = vm.Pair(vm.RL)
    b = oc.CircuitBuilder()
    b.vsource("a", "0", 1.0, name="V1")
    vm.place(b, rl, ("a", "0"), name="L1")
    c = b.build()
    ts = np.linspace(0, 5e-4, 501)
    sol = oa.transient(c, ts, ic="zero", options=oa.Options(**EXACT))
    k = c.layout.internal("L1", 1)
    h, R, L = ts[1], 10.0, 1e-3
    i = 0.0
    for step in range(1, len(ts)):
        i = (i + h / L) / (1 + h * R / L)
        assert sol.z[step, k] == pytest.approx(i, rel=1e-10)
    exact = 0.1 * (1 - np.exp(-ts[-1] * R / L))
    assert sol.z[-1, k] == pytest.approx(exact, rel=1e-2)
