"""The Verilog-A language around the analog block: preprocessor, parser
items, analog functions, loops, case, integer semantics, system functions
and the element options (temperature, $simparam values, range handling,
default and forced card values).  Programs are checked against closed forms
and the oracle's direct interpretation (no GPU needed).  Cases follow the
reference's tests/test_va.py:32-146 and 368-390."""

import math
import warnings

import numpy as np
import pytest

import va_models as vm
from oracle import dual as D
from oracle import va as ova
from test_va import run_program
from voltax_amd import va

K_OVER_Q = 1.3806503e-23 / 1.602176462e-19   # NIST 1998 constants


def _text(src, **kw):
    return va.preprocess(src, **kw)


def test_macros_with_arguments_and_nesting():
    src = "`define SQ(x) ((x)*(x))\n`define TWO 2\ny = `SQ(`TWO + a);"
    assert "y = ((2 + a)*(2 + a));" in _text(src)


def test_multiline_define_keeps_line_numbers():
    lines = _text("`define F(a) \\\n  a + 1\nx = `F(2);\nz = 3;").split("\n")
    assert len(lines) == 4
    assert lines[2].strip() == "x = 2 + 1;" and lines[3] == "z = 3;"


def test_ifdef_elsif_else_and_defines():
    src = "`ifdef FAST\na=1;\n`elsif MEDIUM\na=2;\n`else\na=3;\n`endif"
    assert "a=1;" in _text(src, defines={"FAST": ""})
    assert "a=2;" in _text(src, defines={"MEDIUM": ""})
    assert "a=3;" in _text(src)
    nested = "`ifndef A\n`ifdef B\nx=1;\n`else\nx=2;\n`endif\n`endif"
    assert "x=1;" in _text(nested, defines={"B": ""}) and "x=2;" not in _text(
        nested, defines={"B": ""})
    assert "x" not in _text(nested, defines={"A": ""})
    with pytest.raises(va.VAError, match="without `endif"):
        _text("`ifdef A\nx=1;")


def test_builtin_constants_and_locked_overrides():
    src = '`include "constants.vams"\nx = `P_Q;'
    assert "1.602176462e-19" in _text(src)   # NIST 1998, the LRM default
    assert "1.6021918e-19" in _text(src, defines={"PHYSICAL_CONSTANTS_OLD": ""})
    assert "x = 2;" in _text("`define K 1\nx = `K;", overrides={"K": "2"})
    assert "x = 1;" in _text("`define K 1\nx = `K;", defines={"K": "2"})
    with pytest.raises(va.VAError, match="only the standard"):
        _text('`include "bsim4.va"')


def test_comments_strings_and_continued_strings():
    src = 'a = 1; // c1\n/* multi\nline */ b = "x // y";\n$strobe("p \\\n q");'
    lines = _text(src).split("\n")
    assert lines[0].strip() == "a = 1;"
    assert 'b = "x // y";' in lines[2]
    assert len(lines) == 5
    # a macro use inside a comment is not expanded
    assert _text("x = 1; // `NOPE\n/* `NOPE */").split("\n")[0].strip() == "x = 1;"


def test_undefined_macro_raises_with_location():
    with pytest.raises(va.VAError, match=r"<string>:2: undefined macro `NOPE"):
        _text("a=1;\nb=`NOPE;")


def test_parser_module_items():
    src = """
module m(p, n);
  inout p, n;
  electrical p, n, x;
  branch (p, x) br;
  (* type="instance" *) parameter real w = 1u from (0:inf);
  parameter integer mode = 0 from [0:2] exclude 1;
  parameter real k = 2.0;
  aliasparam kk = k;
  real a;
  analog function real twice;
    input u; real u;
    twice = 2 * u;
  endfunction
  analog begin : main
    real loc;
    a = k;
    loc = twice(a);
    I(br) <+ V(br) / 1k * loc;
    V(x, n) <+ 0;
  end
endmodule
"""
    m = va.parse(src)
    assert m.ports == ["p", "n"] and m.nodes == ["p", "n", "x"]
    assert m.branches == {"br": ("p", "x")}
    assert m.params["w"].instance and m.params["w"].ranges[0][0] == "("
    assert m.params["mode"].type == "integer" and len(m.params["mode"].excludes) == 1
    assert m.aliases == {"kk": "k"} and "twice" in m.functions
    model = va.load(src)
    with pytest.raises(ValueError, match="excluded"):
        model.element({"mode": 1})
    # x collapses onto n, so the branch conducts 4 mS between p and n
    for card, g in ((None, 4e-3), ({"KK": 3.0}, 6e-3)):
        c = model.compile(card)
        assert c.terminals == ("p", "n") and c.internal == ()
        res, jac = run_program(c.code, np.array([1e-6, 1.0]), np.array([0.5, 0.0]))
        assert res[0] == pytest.approx(0.5 * g, rel=1e-15)
        assert jac[0, 0] == pytest.approx(g, rel=1e-15) and jac[1, 0] == pytest.approx(-g)


def _compare(src, card=None, lo=-1.5, hi=1.5, n=24, **kw):
    """The compiled program against the oracle at random unknowns, for f and
    q; returns the compiled program."""
    c = va.load(src).compile(card, **kw)
    model = ova.Model(src, card, **kw)
    assert c.terminals + c.internal == tuple(model.order)
    m = len(model.order)
    U = np.random.default_rng(n).uniform(lo, hi, (n, m))
    T = len(model.terminals)
    seeds = [D.Dual.seed(U[:, j], j, m) for j in range(m)]
    params = np.array([v for _, v in c.params])
    stored = {sk: np.full(n, np.log(model.card[k]) if sk.startswith("log_") else model.card[k])
              for k, sk in zip(model.inst, model.stored)}
    F, Q = ova.group_rows(model, seeds[:T], seeds[T:], {**stored, "m": np.ones(n)})
    for cf, cq in ((1.0, 0.0), (0.0, 1.0)):
        for i in range(n):
            res, jac = run_program(c.code, params, U[i], cf, cq)
            ref = np.array([cf * D.val(F[r])[i] + cq * D.val(Q[r])[i] for r in range(m)])
            rjac = np.array([cf * D.tan(F[r], (n, m))[i] + cq * D.tan(Q[r], (n, m))[i]
                             for r in range(m)])
            np.testing.assert_allclose(res, ref, rtol=1e-13, atol=1e-300)
            np.testing.assert_allclose(jac, rjac, rtol=1e-12,
                                       atol=1e-14 * max(1.0, np.abs(rjac).max()))
    return c


def _lang_current(u, card=None, temperature=300.15, gmin=1e-12):
    """I(p, mid) of the LANG model written out by hand."""
    card = card or {}
    order, mode = card.get("order", 3), card.get("mode", 1)
    g0, tc = card.get("g0", card.get("gzero", 1e-3)), card.get("tc")
    v = min(max(u / (K_OVER_Q * temperature), -10.0), 10.0)

    def poly(x, n):
        return sum(x ** j / (j + 1) for j in range(n))

    acc = sum(poly(0.1 * v, order + k % 2) for k in range(5))
    s = {0: 1.0, 1: 1 + 0.5 * math.tanh(v), 2: 1 + 0.5 * math.tanh(v)}[mode]
    if tc is not None:
        s *= 1 + tc * (temperature - 300.15)
    return g0 * (order // 2 + order % 2 + 3 / 4.0) * acc * s * u + gmin * u


CARDS = [
    (None, {}),
    ({"tc": 1e-3, "mode": 0, "order": 5}, dict(temperature=350.0, simparams={"gmin": 1e-9})),
    ({"gzero": 2e-3, "mode": 2, "order": 4}, {}),
]


@pytest.mark.parametrize("card,kw", CARDS)
def test_language_model_matches_oracle_and_hand_formula(card, kw):
    c = _compare(vm.LANG, card, **kw)
    assert c.internal == ("mid",) and c.loop_sites == 0   # every loop is static
    temp = kw.get("temperature", 300.15)
    gmin = kw.get("simparams", {}).get("gmin", 1e-12)
    for u in (-0.4, 0.01, 0.07, 0.3):
        res, jac = run_program(c.code, np.array([1.0, 1.0]), np.array([u, 0.0, 0.0]))
        assert res[0] == pytest.approx(_lang_current(u, card, temp, gmin), rel=1e-13)
        assert res[2] == pytest.approx(-res[0], rel=1e-15)   # KCL at mid, V(mid, n) = 0
        h = 1e-7
        fd = (_lang_current(u + h, card, temp, gmin) - _lang_current(u - h, card, temp, gmin)) / (2 * h)
        assert jac[0, 0] == pytest.approx(fd, rel=1e-6)


def test_integer_semantics_on_device_values():
    """Assignment to an integer rounds half away from zero; / and % of
    integers truncate; % of reals is the C fmod; rounded values carry no
    derivative."""
    c = _compare(vm.INTOPS, lo=-1.0, hi=1.0)
    for v, want in ((0.5, 8e-3), (-0.5, -8e-3), (0.2, (0 + 1 + 0.125 + 1.4 + 1.4) * 1e-3),
                    (1.0, (3 + 1 + 0.875 + 7 + 2) * 1e-3)):
        res, jac = run_program(c.code, np.array([1.0]), np.array([v, 0.0]))
        assert res[0] == pytest.approx(want, rel=1e-14)
        assert jac[0, 0] == pytest.approx(14e-3, rel=1e-14)


def test_static_integer_arithmetic_folds():
    src = """
module ints(p, n);
  inout p, n; electrical p, n;
  integer a, b, c, d, e;
  real r;
  analog begin
    a = 7 / 2; b = -7 / 2; c = 7 % 3; d = -7 % 3; e = -2.5;
    r = 7.0 / 2;
    I(p, n) <+ (a + 10 * b + 100 * c + 1000 * d + 10000 * e + 100000 * r) * V(p, n);
  end
endmodule
"""
    c = _compare(src)
    res, _ = run_program(c.code, np.array([1.0]), np.array([1.0, 0.0]))
    assert res[0] == 3 - 30 + 100 - 1000 - 30000 + 350000


def test_param_given_simparam_ranges_and_unknown_params():
    src = """
module q(p, n);
  inout p, n; electrical p, n;
  parameter real a = 1.0 from (0:10];
  parameter real b = 0.0;
  real g;
  analog begin
    if ($param_given(b)) g = b; else g = a;
    I(p, n) <+ (g + $simparam("gmin", 1.0)) * V(p, n);
  end
endmodule
"""
    m = va.load(src)

    def f(c):
        return run_program(c.code, np.array([1.0]), np.array([1.0, 0.0]))[0][0]

    assert f(m.compile()) == 1.0 + 1e-12
    assert f(m.compile({"b": 3.0}, simparams={"gmin": 0.0})) == 3.0
    assert f(m.compile({"b": 0.0})) == 1e-12   # given, even at its default value
    _compare(src, {"b": 3.0}, simparams={"gmin": 0.0})
    with pytest.raises(ValueError, match="outside its range"):
        m.element({"a": 20.0})
    with pytest.warns(UserWarning, match="outside its range"):
        m.element({"a": 20.0}, ranges="warn")
    with warnings.catch_warnings():
        warnings.simplefilter("error")
        E = m.element({"a": 20.0}, ranges="ignore")
    assert E.card["a"] == 20.0
    with pytest.warns(UserWarning, match="unknown parameters"):
        m.element({"zz": 1.0})
    with pytest.raises(va.VAError, match="no value or default"):
        va.load(src.replace('$simparam("gmin", 1.0)', '$simparam("tnom")')).compile()


def test_defaults_force_and_integral_integer_parameters():
    src = """
module d(p, n);
  inout p, n; electrical p, n;
  parameter real a = 1.0;
  parameter real b = 2.0;
  parameter integer k = 1;
  analog I(p, n) <+ (a + 10 * b + 100 * k) * V(p, n);
endmodule
"""
    m = va.load(src)
    E = m.element({"a": 3.0}, defaults={"a": 5.0, "b": 4.0}, force={"k": 2})
    assert (E.card["a"], E.card["b"], E.card["k"]) == (3.0, 4.0, 2.0)
    res, _ = run_program(E.va.code, np.array([1.0]), np.array([1.0, 0.0]))
    assert res[0] == 3.0 + 40.0 + 200.0
    with pytest.raises(ValueError, match="not integral"):
        m.element({"k": 1.5})


def test_temperature_and_thermal_voltage():
    src = """
module t(p, n);
  inout p, n; electrical p, n;
  analog I(p, n) <+ ($vt + $vt(400) * 1k + $temperature * 1M) * V(p, n);
endmodule
"""
    m = va.load(src)
    for temp in (None, 250.0):
        E = m.element(temperature=temp)
        T = 300.15 if temp is None else temp
        res, _ = run_program(E.va.code, np.array([1.0]), np.array([1.0, 0.0]))
        assert res[0] == pytest.approx(K_OVER_Q * T + K_OVER_Q * 400 * 1e3 + T * 1e6, rel=1e-15)
        assert E.device_temperature == T
        _compare(src, temperature=temp)


def test_case_default_and_named_block_locals():
    src = """
module c(p, n);
  inout p, n; electrical p, n;
  parameter integer sel = 2;
  real y;
  analog begin : body
    integer i;
    y = 0;
    for (i = 1; i <= 3; i = i + 1) begin : step
      case (i * sel)
        2, 4: y = y + 1;
        6: y = y + 10;
        default: y = y + 100;
      endcase
    end
    I(p, n) <+ y * V(p, n);
  end
endmodule
"""
    m = va.load(src)
    for sel, want in ((1, 100 + 1 + 100), (2, 1 + 1 + 10), (3, 100 + 10 + 100)):
        c = m.compile({"sel": sel})
        res, _ = run_program(c.code, np.array([1.0]), np.array([1.0, 0.0]))
        assert res[0] == want
        _compare(src, {"sel": sel})


def test_data_dependent_case_selects_per_device():
    src = """
module s(p, n);
  inout p, n; electrical p, n;
  integer q;
  analog begin
    q = 2 * V(p, n);
    case (q)
      0: I(p, n) <+ 1m * V(p, n);
      1: I(p, n) <+ 2m * V(p, n);
      default: I(p, n) <+ 3m * V(p, n);
    endcase
  end
endmodule
"""
    c = _compare(src, lo=-1.0, hi=1.5)
    for v, g in ((0.1, 1e-3), (0.3, 2e-3), (0.8, 3e-3), (-0.6, 3e-3)):
        res, jac = run_program(c.code, np.array([1.0]), np.array([v, 0.0]))
        assert res[0] == pytest.approx(g * v, rel=1e-15) and jac[0, 0] == pytest.approx(g)


def test_fatal_system_tasks_and_bad_functions_raise():
    body = "analog begin $finish; I(p, n) <+ V(p, n); end"
    with pytest.raises(va.VAError, match="finish"):
        va.load(f"module b(p, n); inout p, n; electrical p, n; {body} endmodule")
    fn = """
module b(p, n); inout p, n; electrical p, n;
  analog function real h; input x; real x; h = 2 * x; endfunction
  analog I(p, n) <+ h(V(p, n), 1);
endmodule"""
    with pytest.raises(va.VAError, match="takes 1 arguments"):
        va.load(fn).compile()


def test_differentiable_model_parameters_become_device_rows():
    """``differentiable`` model parameters are read from the device row after
    the instance parameters (voltax/va/model.py:241-262): the program equals
    the oracle's, and the oracle's adjoint gradient with respect to them
    equals finite differences of its DC solution."""
    from oracle import analysis as oa
    from oracle import circuit as oc

    m = va.load(vm.DIODE)
    assert m.differentiable("all") == ("is", "rs", "cj", "vt")
    with pytest.raises(ValueError, match="not a real model parameter"):
        m.differentiable(["area"])
    E = m.element({"rs": 10.0}, differentiable=["IS", "vt"])
    assert E.params == ("area", "is", "vt", "m")
    assert [k for k, _ in E.va.params] == ["area", "is", "vt", "m"]
    dev = E(("a", "c"), area=2.0)
    assert list(dev.rows[0]) == [2.0, 1e-14, 0.025852, 1.0]
    _compare(vm.DIODE, {"rs": 10.0}, lo=-0.3, hi=0.8, differentiable=["is", "vt"])

    pair = vm.Pair(vm.DIODE, {"rs": 10.0}, differentiable=["is"])
    EXACT = dict(rtol=1e-12, atol=1e-15, gmin=0.0, gmin_steps=0)

    def solve(is_value):
        b = oc.CircuitBuilder()
        b.vsource("in", "0", 5.0, name="V1")
        b.resistor("in", "a", 1e3, name="R1")
        vm.place(b, pair, ("a", "0"), name="D1", **{"is": is_value})
        c = b.build()
        return c, oa.dc(c, options=oa.Options(**EXACT))

    c, sol = solve(1e-14)
    g = np.zeros(c.size)
    g[c.node("a")] = 1.0
    gr = oa.dc_vjp(c, sol.z, 0.0, g, options=oa.Options(**EXACT))["diode"]["is"][0]
    h = 1e-18
    fd = (solve(1e-14 + h)[1].z[c.node("a")] - solve(1e-14 - h)[1].z[c.node("a")]) / (2 * h)
    assert gr == pytest.approx(fd, rel=1e-6)


CLAMP = """
module c(p, n);
  inout p, n; electrical p, n;
  real x;
  analog begin
    x = V(p, n);
    if (x < 0.0) x = 0.0;
    I(p, n) <+ 1m * x;
  end
endmodule
"""

SMOOTHED = """
module s(p, n);
  inout p, n; electrical p, n;
  (* type="instance" *) parameter real k = 0.2;
  real v, y;
  analog begin
    v = V(p, n);
    y = max(v, k) + min(v, -0.3) + abs(v - 0.1) + max(k, 0.5);
    if (0.4 < y) y = 0.4;
    I(p, n) <+ 1m * y;
  end
endmodule
"""


def test_smooth_mode_replaces_clamps():
    """tests/test_va.py:341-366 of the reference: the clamp becomes a smooth
    max within w ln 2; per-line widths smooth only the listed line."""
    m = va.load(CLAMP)
    v = np.linspace(-0.2, 0.2, 9)

    def currents(c):
        return np.array([run_program(c.code, np.ones(1), np.array([u, 0.0]))[0][0] for u in v])

    exact = m.compile()
    assert currents(exact)[0] == 0.0
    smooth = m.compile(smooth=0.01)
    err = np.abs(currents(smooth) - 1e-3 * np.maximum(v, 0.0))
    assert err.max() <= 1e-3 * 0.01 * np.log(2) + 1e-15
    assert err.max() >= 1e-3 * 0.01 * np.log(2) * (1 - 1e-12)  # at v = 0
    assert len(m.compile(smooth={7: 0.01}).code) == len(smooth.code)
    assert np.array_equal(m.compile(smooth={6: 0.01}).code, exact.code)
    with pytest.raises(ValueError, match="positive"):
        m.compile(smooth=0.0)


@pytest.mark.parametrize("smooth", [None, 0.05, {8: 0.02}, {9: 0.05}])
def test_smooth_min_max_abs_match_oracle(smooth):
    """Smoothed min / max / abs on device-dependent values (a constant max
    stays exact) and an upper clamp: program = oracle, values and Jacobians."""
    c = _compare(SMOOTHED, lo=-1.0, hi=1.0, smooth=smooth)
    if smooth is None:
        return
    # smax(a, b) - max(a, b) lies in [0, w ln 2]; at v = k = 0.2 the max is smoothed
    res, _ = run_program(c.code, np.array([0.2, 1.0]), np.array([0.2, 0.0]))
    exact, _ = run_program(va.load(SMOOTHED).compile().code, np.array([0.2, 1.0]),
                           np.array([0.2, 0.0]))
    if isinstance(smooth, dict) and 8 not in smooth:
        assert res[0] != exact[0]  # only the clamp (line 9) is smoothed
    else:
        assert res[0] > exact[0] - 1e-3 * 0.05 * np.log(2) * 3


EQUALITY = """
module e(p, n);
  inout p, n; electrical p, n;
  real x;
  analog begin
    x = V(p, n);
    if (V(p, n) == 0.0) x = 0.0;   // a guard that patches one point
    I(p, n) <+ 1m * x;
  end
endmodule
"""

EQ2 = """
module e2(p, n);
  inout p, n; electrical p, n;
  real x, y;
  analog begin
    x = V(p, n); y = 2 * V(p, n);
    if ((x == 0.0) && (y == 0.0)) x = 0.0; else x = x;
    if (y != 0.0) y = y; else y = 0.0;
    if ((x == 0.0) || (y > 1.0)) y = 3 * y;
    I(p, n) <+ 1m * (x + y);
  end
endmodule
"""


def test_equality_branch_limit_mode():
    """tests/test_va.py:298-317 of the reference: at the patched point the
    value is the special case's, the slope is 0 ("value") or the general
    formula's ("limit"); elsewhere both modes agree."""
    m = va.load(EQUALITY)
    for mode, slope in (("value", 0.0), ("limit", 1e-3)):
        c = m.compile(equality=mode)
        res, jac = run_program(c.code, np.ones(1), np.array([0.0, 0.0]))
        assert res[0] == 0.0 and jac[0, 0] == slope
        res, jac = run_program(c.code, np.ones(1), np.array([0.1, 0.0]))
        assert jac[0, 0] == pytest.approx(1e-3)
        _compare(EQUALITY, lo=-1.0, hi=1.0, equality=mode)
    with pytest.raises(ValueError, match="equality"):
        m.compile(equality="exact")
    assert va.equality_side(va.parse(EQ2).body[0][1][2][1]) == "then"


def test_equality_detection_conjunctions_and_not_equal():
    """Conjunctions of equalities and != tests are special-case arms; a
    disjunction mixing an inequality is not (compiler.py:1297-1311): the
    limit-mode program against the oracle, at the special points too."""
    c = _compare(EQ2, lo=-1.0, hi=1.0, equality="limit")
    for v, slope in ((0.0, 1e-3 * (1 + 6)), (0.25, 1e-3 * (1 + 2)), (0.7, 1e-3 * (1 + 6))):
        res, jac = run_program(c.code, np.ones(1), np.array([v, 0.0]))
        assert jac[0, 0] == pytest.approx(slope, rel=1e-14)
    value = va.load(EQ2).compile()
    _, jac = run_program(value.code, np.ones(1), np.array([0.0, 0.0]))
    assert jac[0, 0] == 0.0  # "value": both patched arms carry slope 0 at v = 0


EQ_CONTRIB = """
module ec(p, n);
  inout p, n; electrical p, n, m;
  analog begin
    I(p, m) <+ V(p, m) / 10.0;
    V(m, n) <+ 0.5m * ddt(I(m, n));
    if (V(p, n) == 0.0) begin
      I(p, n) <+ 0.0;
      V(m, n) <+ 0.0;
    end else begin
      I(p, n) <+ 1m * V(p, n) + ddt(2p * V(p, n));
      V(m, n) <+ 3.0 * I(m, n);
    end
  end
endmodule
"""


def test_limit_mode_merges_contributions_made_inside_the_arms():
    """Contributions inside an ``if (a == b)`` arm are merged like variables
    (the reference keeps them in its environment, compiler.py:1057-1083):
    in "limit" mode the current, charge and voltage-branch terms keep the
    special arm's values at V(p, n) = 0 with the general arm's derivatives,
    here by hand: dI(p,n)/dV = 1e-3, dQ/dV = 2e-12, d(branch row)/d i(m,n) =
    -3 at the special point; "value" mode gives the special arm's zero
    slopes.  Elsewhere both modes agree with the oracle."""
    for mode in ("value", "limit"):
        c = _compare(EQ_CONTRIB, lo=-1.0, hi=1.0, equality=mode)
        order = list(c.terminals + c.internal)
        p, n, br = order.index("p"), order.index("n"), order.index("i(m,n)")
        u = np.zeros(len(order))
        u[br] = 0.2
        res, jac = run_program(c.code, np.ones(1), u, 1.0, 0.0)
        _, cjac = run_program(c.code, np.ones(1), u, 0.0, 1.0)
        lim = mode == "limit"
        assert res[p] == 0.0
        # row p holds I(p, m) = V(p, m) / 10 (slope 0.1) plus the arm's current
        assert jac[p, p] - 0.1 == pytest.approx(1e-3 if lim else 0.0, abs=1e-15)
        assert cjac[p, p] == pytest.approx(2e-12 if lim else 0.0, abs=1e-24)
        assert jac[br, br] == pytest.approx(-3.0 if lim else 0.0, abs=1e-15)
        assert res[br] == pytest.approx(u[order.index("m")] - u[n] - 0.0, abs=1e-15)


TRACED_BOOL = """
module tb(p, n);
  inout p, n; electrical p, n;
  real x, y;
  analog begin
    x = (V(p, n) > 0) / 2;     // a bool on device values: read as real, 0.5
    y = (3 > 2) / 2;           // constants: integer division, 0
    I(p, n) <+ 1m * (x + y + (V(p, n) > 0) % 2) * V(p, n);
  end
endmodule
"""


def test_comparison_on_device_values_is_a_real_in_arithmetic():
    """compiler.py:600-662: a comparison on constants is an integer (3 > 2
    gives 1, and 1 / 2 truncates to 0), on device values a boolean that
    arithmetic reads as real ((V > 0) / 2 = 0.5, and % on it is fmod)."""
    c = _compare(TRACED_BOOL, lo=-1.0, hi=1.0)
    for v, want in ((0.4, 1e-3 * (0.5 + 0 + 1.0) * 0.4), (-0.4, 0.0)):
        res, _ = run_program(c.code, np.ones(1), np.array([v, 0.0]))
        assert res[0] == pytest.approx(want, rel=1e-15, abs=1e-300)


def test_bsim_sized_model_compiles_and_matches_the_oracle():
    """A BSIM4-sized synthetic model (10 unknowns: four terminals, five
    internal nodes and a branch current; about 600 instructions after the
    binning loop unrolls): the compiled program against the oracle's direct
    interpretation, currents and charges with their 10 x 10 Jacobians."""
    c = _compare(vm.BSIMISH, lo=-0.3, hi=1.2, n=8)
    assert len(c.terminals + c.internal) == 10 and c.code[3] > 300
