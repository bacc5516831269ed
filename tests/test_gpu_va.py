"""Verilog-A devices on the GPU: the stamping kernel's interpreter against
the oracle (residuals and Jacobians), and DC / transient / adjoint solves
against the oracle and closed forms."""

import math

import numpy as np
import pytest

import va_models as vm
from oracle import analysis as oa
from oracle import circuit as oc
from oracle import sparse as osp

pytestmark = pytest.mark.gpu

TIGHT = dict(rtol=1e-12, atol=1e-15)
EXACT = dict(rtol=1e-12, atol=1e-15, gmin=0.0, gmin_steps=0)


def _both(build):
    import voltax_amd as vx

    return build(vx.CircuitBuilder()), build(oc.CircuitBuilder())


def _mixed(b, vin=0.9):
    """Every test model in one circuit, next to built-in devices."""
    b.vsource("vdd", "0", 1.8, name="Vdd")
    b.vsource("in", "0", vin, name="Vin")
    b.resistor("vdd", "out", 10e3, name="RL")
    vm.place(b, vm.Pair(vm.SQNMOS), ("out", "in", "0"), w=1.5)
    vm.place(b, vm.Pair(vm.RES), ("out", "x"), r=4.7e3)
    vm.place(b, vm.Pair(vm.DIODE, {"rs": 5.0}), ("x", "0"), area=3.0)
    vm.place(b, vm.Pair(vm.RL), ("in", "y"))
    b.resistor("y", "0", 100.0, name="Ry")
    vm.place(b, vm.Pair(vm.HERON), ("x", "y"))
    vm.place(b, vm.Pair(vm.SUMK), ("y", "0"))
    vm.place(b, vm.Pair(vm.VCVS), ("x", "w"))
    b.resistor("w", "0", 1e3, name="Rw")
    vm.place(b, vm.Pair(vm.MIXFN), ("w", "out"))
    b.capacitor("out", "0", 10e-15, name="Cout")
    return b.build()


def test_stamp_matches_oracle():
    """Residual and Jacobian of cf*f + cq*q at random states: the GPU
    interpreter against the oracle's direct interpretation."""
    from voltax_amd import sparse

    c, o = _both(_mixed)
    rng = np.random.default_rng(3)
    st = osp.structure(o)
    for cf, cq in ((1.0, 0.0), (0.0, 1.0), (1.0, 2e9)):
        z = rng.uniform(-0.2, 0.9, c.size)
        got = sparse.jacobian(c, z, 0.0, cf, cq)
        got = got.cpu().numpy() if hasattr(got, "cpu") else np.asarray(got)
        ref = osp.jacobian(o, z, 0.0, cf, cq, st)
        scale = np.abs(ref).max()
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-15 * scale)
        r = sparse.residual(c, z, 0.0, cf, cq)
        r = r.cpu().numpy() if hasattr(r, "cpu") else np.asarray(r)
        ref_r = cf * o.f(z) + cq * o.q(z)
        np.testing.assert_allclose(r, ref_r, rtol=1e-12, atol=1e-15 * np.abs(ref_r).max())


def test_dc_divider_exact():
    import voltax_amd as vx

    def build(b):
        res = vm.Pair(vm.RES)
        b.vsource("a", "0", 3.0, name="V1")
        vm.place(b, res, ("a", "b"))
        vm.place(b, res, ("b", "0"), r=2e3)
        return b.build()

    c = build(vx.CircuitBuilder())
    sol = vx.dc(c, options=vx.Options(**EXACT))
    assert sol.converged
    assert sol.z[c.node("b")] == pytest.approx(2.0, rel=1e-14)
    assert sol.z[-1] == pytest.approx(-1e-3, rel=1e-14)  # source branch current


def test_resistor_like_the_reference():
    """tests/test_va.py:148-162 of the reference: the divider through one
    fused group named by the element, V(b) = 1.5, per-device r, and
    d V(b) / d log r2 = 2 r1 r2 / (r1 + r2)^2."""
    import voltax_amd as vx
    from voltax_amd import adjoint, va

    R = va.load(vm.RES).element(name="rva")
    b = vx.CircuitBuilder()
    b.vsource("a", "0", 2.0, name="V1")
    b.add(R(("a", "b"), r=1e3), name="R1")
    b.add(R(("b", "0"), r=3e3), name="R2")
    c = b.build()
    assert set(c.elements) == {"VoltageSource", "rva"}
    sol = vx.dc(c)  # default options: the final gmin of 1e-12 S stays in the root
    assert sol.z[c.node("b")] == pytest.approx(1.5, rel=1e-12, abs=1e-8)  # jnp.isclose's atol
    assert c.get("R2", "r") == 3e3
    g = np.zeros(c.size)
    g[c.node("b")] = 1.0
    d = adjoint.dc_vjp(c, sol.z, 0.0, g)["rva"]["log_r"][1]
    assert d == pytest.approx(2.0 * 1e3 * 3e3 / 4e3 ** 2, rel=1e-8)


def test_guarded_branch_jacobian_is_finite():
    """The untaken sqrt branch (NaN slope at v < 0) never reaches the
    stamped Jacobian (tests/test_va.py:209-247 of the reference)."""
    import voltax_amd as vx
    from voltax_amd import sparse

    b = vx.CircuitBuilder()
    b.vsource("p", "0", 0.0, name="V1")
    vm.place(b, vm.Pair(vm.GUARDED), ("p", "0"))
    c = b.build()
    st = sparse.structure(c)
    for v, slope in ((-0.5, 1e-3), (0.25, 1e-3)):
        z = c.zeros()
        z[c.node("p")] = v
        data = sparse.jacobian(c, z).cpu().numpy()
        A = st.to_dense(data)
        assert np.all(np.isfinite(A))
        assert A[c.node("p"), c.node("p")] == pytest.approx(slope, rel=1e-14)


@pytest.mark.parametrize("rs", [0.0, 10.0])
def test_dc_diode_closed_form(rs):
    import voltax_amd as vx
    from scipy.optimize import brentq

    d = vm.Pair(vm.DIODE, {"rs": rs})

    def build(b):
        b.vsource("in", "0", 5.0, name="V1")
        b.resistor("in", "a", 1e3, name="R1")
        vm.place(b, d, ("a", "0"), area=2.0)
        return b.build()

    c, o = build(vx.CircuitBuilder()), build(oc.CircuitBuilder())
    sol = vx.dc(c, options=vx.Options(**EXACT))
    ref = oa.dc(o, options=oa.Options(**EXACT))
    assert sol.converged and ref.converged

    def kcl(vj):
        i = 2e-14 * (math.exp(vj / 0.025852) - 1)
        return (5 - vj - rs * i) / 1e3 - i

    vj = brentq(kcl, 0.0, 1.0, xtol=1e-16)
    i = 2e-14 * (math.exp(vj / 0.025852) - 1)
    assert sol.z[c.node("a")] == pytest.approx(vj + rs * i, rel=1e-11)
    np.testing.assert_allclose(sol.z, ref.z, rtol=1e-10, atol=1e-15)


def test_dc_mixed_circuit_matches_oracle():
    import voltax_amd as vx

    c, o = _both(_mixed)
    sol = vx.dc(c, options=vx.Options(**TIGHT))
    ref = oa.dc(o, options=oa.Options(**TIGHT))
    assert sol.converged and ref.converged
    np.testing.assert_allclose(sol.z, ref.z, rtol=1e-9, atol=1e-13)


def test_per_device_instance_parameters():
    """Forty diodes with their own area in one group, each behind its own
    resistor: every branch against the oracle."""
    import voltax_amd as vx

    areas = np.linspace(0.5, 20.0, 40)
    d = vm.Pair(vm.DIODE, {"rs": 2.0})

    def build(b):
        b.vsource("in", "0", 2.0, name="V1")
        for k, a in enumerate(areas):
            b.resistor("in", f"a{k}", 500.0 + 10 * k)
            vm.place(b, d, (f"a{k}", "0"), area=a)
        return b.build()

    c, o = build(vx.CircuitBuilder()), build(oc.CircuitBuilder())
    sol = vx.dc(c, options=vx.Options(**TIGHT))
    ref = oa.dc(o, options=oa.Options(**TIGHT))
    assert sol.converged and ref.converged
    np.testing.assert_allclose(sol.z, ref.z, rtol=1e-9, atol=1e-13)
    v = [sol.z[c.node(f"a{k}")] for k in range(len(areas))]
    assert len(set(np.round(v, 9))) == len(areas)  # the areas reached the devices
    from voltax_amd import adjoint

    g = np.zeros(c.size)
    g[[c.node(f"a{k}") for k in range(len(areas))]] = np.linspace(1.0, 2.0, len(areas))
    gr = adjoint.dc_vjp(c, sol.z, 0.0, g)["diode"]
    go = oa.dc_vjp(o, ref.z, 0.0, g)["diode"]
    for par in ("area", "m"):
        np.testing.assert_allclose(gr[par], go[par], rtol=1e-7, atol=1e-14)
        assert np.all(np.abs(go[par]) > 1e-6)


@pytest.mark.parametrize("method", ["be", "trap"])
def test_transient_rl_matches_oracle_and_recursion(method):
    """RL step response: against the oracle, and (backward Euler) against the
    one-line recursion of the discretized inductor current."""
    import voltax_amd as vx

    rl = vm.Pair(vm.RL)

    def build(b):
        b.vsource("a", "0", vm_step(b), name="V1")
        vm.place(b, rl, ("a", "0"), name="L1")
        return b.build()

    c, o = build(vx.CircuitBuilder()), build(oc.CircuitBuilder())
    ts = np.linspace(0, 5e-4, 251)
    sol = vx.transient(c, ts, ic="zero", method=method, options=vx.Options(**TIGHT))
    ref = oa.transient(o, ts, ic="zero", method=method, options=oa.Options(**TIGHT))
    assert np.all(sol.converged)
    np.testing.assert_allclose(sol.z, ref.z, rtol=1e-9, atol=1e-13)
    k = c.layout.internal("L1", "i(mid,n)")
    if method == "be":
        h, R, L, i = ts[1], 10.0, 1e-3, 0.0
        for step in range(1, len(ts)):
            v = min(max((ts[step] - 1e-5) / 1e-12, 0.0), 1.0)  # the Step waveform
            i = (i + h * v / L) / (1 + h * R / L)
            assert sol.z[step, k] == pytest.approx(i, rel=1e-9, abs=1e-15)


def vm_step(b):
    """A 1 V step at 10 us, from the builder's own signal module."""
    mod = oc if isinstance(b, oc.CircuitBuilder) else __import__("voltax_amd")
    return mod.Step(0.0, 1.0, 1e-5)


def test_transient_mixed_circuit_matches_oracle():
    import voltax_amd as vx

    def build(b):
        mod = oc if isinstance(b, oc.CircuitBuilder) else vx
        return _mixed(b, mod.Pulse(0.9, 1.4, 0.2e-9, 0.1e-9, 0.1e-9, 1e-9, 4e-9))

    c, o = _both(build)
    ts = np.linspace(0, 2e-9, 41)
    sol = vx.transient(c, ts, ic="dc", options=vx.Options(**TIGHT))
    ref = oa.transient(o, ts, ic="dc", options=oa.Options(**TIGHT))
    assert np.all(sol.converged)
    np.testing.assert_allclose(sol.z, ref.z, rtol=1e-8, atol=1e-12)


def _lang_circuit(b, drive):
    """Two LANG devices (different cards, temperature and $simparam values)
    in a resistive ladder with a capacitor."""
    mod = oc if isinstance(b, oc.CircuitBuilder) else __import__("voltax_amd")
    b.vsource("in", "0", drive(mod), name="Vin")
    b.resistor("in", "a", 200.0, name="R1")
    vm.place(b, vm.Pair(vm.LANG), ("a", "b"), name="L1", scale=2.0)
    vm.place(b, vm.Pair(vm.LANG, {"tc": 2e-3, "mode": 0, "order": 5}, temperature=340.0,
                        simparams={"gmin": 1e-9}), ("b", "0"), name="L2")
    b.resistor("b", "0", 1e3, name="R2")
    b.capacitor("a", "0", 1e-12, name="Ca")
    return b.build()


def test_language_features_stamp_and_solve_like_the_oracle():
    """A model using macros, named branches, an analog function, for / case,
    integer arithmetic and system functions: the GPU interpreter's residual
    and Jacobian, the DC point and a transient against the oracle."""
    import voltax_amd as vx
    from voltax_amd import sparse

    c, o = _both(lambda b: _lang_circuit(b, lambda m: 0.6))
    rng = np.random.default_rng(5)
    st = osp.structure(o)
    for cf, cq in ((1.0, 0.0), (0.0, 1.0), (1.0, 1e10)):
        z = rng.uniform(-0.3, 0.7, c.size)
        got = sparse.jacobian(c, z, 0.0, cf, cq)
        got = got.cpu().numpy() if hasattr(got, "cpu") else np.asarray(got)
        ref = osp.jacobian(o, z, 0.0, cf, cq, st)
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-15 * np.abs(ref).max())
    sol = vx.dc(c, options=vx.Options(**TIGHT))
    ref = oa.dc(o, options=oa.Options(**TIGHT))
    assert sol.converged and ref.converged
    np.testing.assert_allclose(sol.z, ref.z, rtol=1e-10, atol=1e-14)

    def sine(m):
        return m.Sine(0.3, 0.4, 2e8)

    c, o = _both(lambda b: _lang_circuit(b, sine))
    ts = np.linspace(0, 1e-8, 51)
    tr = vx.transient(c, ts, ic="dc", options=vx.Options(**TIGHT))
    tref = oa.transient(o, ts, ic="dc", options=oa.Options(**TIGHT))
    assert np.all(tr.converged)
    np.testing.assert_allclose(tr.z, tref.z, rtol=1e-8, atol=1e-12)


def test_adjoint_through_verilog_a_devices():
    """d V(x) / d parameters of a circuit mixing built-in and Verilog-A
    devices: the transposed Jacobian products and the parameter-seeded
    interpreter passes against the oracle's adjoint, every group and
    parameter (log r, area, w, multiplicities, source values)."""
    import voltax_amd as vx
    from voltax_amd import adjoint

    c, o = _both(_mixed)
    sol = vx.dc(c, options=vx.Options(**TIGHT))
    ref = oa.dc(o, options=oa.Options(**TIGHT))
    g = np.zeros(c.size)
    g[c.node("x")] = 1.0
    gr = adjoint.dc_vjp(c, sol.z, 0.0, g)
    go = oa.dc_vjp(o, ref.z, 0.0, g)
    scale = max(np.abs(v).max() for d in go.values() for v in d.values())
    assert set(go) == set(gr)
    for grp, d in go.items():
        assert set(d) == set(gr[grp]), grp
        for par, v in d.items():
            np.testing.assert_allclose(gr[grp][par], v, rtol=1e-7, atol=1e-12 * scale,
                                       err_msg=f"{grp}.{par}")


def test_netlist_level_model_hook_solves_to_the_hand_root():
    """tests/test_va.py:449-467 of the reference: the level=99 card through
    the netlist hook, V(out) = the triode root of (1 - v)/10k = 1m (0.7 - v/2) v."""
    import voltax_amd as vx

    m = vx.va.load(vm.SQUARE_LAW)
    net = """
Vdd vdd 0 1.0
Vin in 0 1.0
Mn out in 0 0 nch w=1u l=100n
Rl vdd out 10k
.model nch nmos level=99 vth=0.3 kp=100u
"""
    c = vx.parse_netlist(net, models=vx.va.level_models(m, level=99))
    assert "nch_n" in c.elements
    vout = vx.dc(c).z[c.node("out")]
    k = 100e-6 * 10
    a, b_, cc = 0.5 * k, -(k * 0.7 + 1e-4), 1e-4
    v = (-b_ - np.sqrt(b_ ** 2 - 4 * a * cc)) / (2 * a)
    assert np.isclose(vout, v, rtol=1e-9)


def test_netlist_readme_chain_delay():
    """README.md:60-79 of the reference, parsed from text: 43.5 ps."""
    import voltax_amd as vx
    from voltax_amd import measure

    from test_host import README_CHAIN

    c = vx.parse_netlist(README_CHAIN)
    t = np.linspace(0, 1e-9, 1001)
    sol = vx.transient(c, t)
    d = measure.delay(t, sol.z[:, c.node("in")], sol.z[:, c.node("out")], 0.6,
                      targ_direction="fall")
    assert round(d * 1e12, 1) == 43.5


def test_adjoint_with_differentiable_model_parameters():
    """Model parameters made differentiable (``differentiable=``) are device
    rows: the GPU's parameter-seeded interpreter passes give the oracle's
    adjoint gradients for them (per device; the model gradient is the sum)."""
    import voltax_amd as vx
    from voltax_amd import adjoint

    def build(b):
        b.vsource("in", "0", 3.0, name="V1")
        b.resistor("in", "a", 1e3, name="R1")
        vm.place(b, vm.Pair(vm.DIODE, {"rs": 10.0}, differentiable=["is", "cj"]), ("a", "0"),
                 area=2.0)
        vm.place(b, vm.Pair(vm.SQNMOS, differentiable="all"), ("a", "a", "0"), w=0.5)
        vm.place(b, vm.Pair(vm.LANG, differentiable=["g0", "c0"]), ("a", "0"), scale=0.3)
        return b.build()

    c, o = _both(build)
    sol = vx.dc(c, options=vx.Options(**TIGHT))
    ref = oa.dc(o, options=oa.Options(**TIGHT))
    assert sol.converged and ref.converged
    np.testing.assert_allclose(sol.z, ref.z, rtol=1e-10, atol=1e-14)
    g = np.zeros(c.size)
    g[c.node("a")] = 1.0
    gr = adjoint.dc_vjp(c, sol.z, 0.0, g)
    go = oa.dc_vjp(o, ref.z, 0.0, g)
    assert set(gr) == set(go)
    assert set(go["diode"]) == {"area", "is", "cj", "m"}
    assert set(go["sqnmos"]) == {"w", "kp", "vt0", "lambda", "m"}
    for grp, d in go.items():
        for par, v in d.items():
            scale = max(np.abs(v).max(), 1e-30)
            np.testing.assert_allclose(gr[grp][par], v, rtol=1e-7, atol=1e-12 * scale,
                                       err_msg=f"{grp}.{par}")


def test_smooth_mode_stamps_and_solves_like_the_oracle():
    """Smooth min / max / abs / clamp surrogates (``smooth=``) through the GPU
    interpreter: residuals, Jacobians and the DC point against the oracle."""
    import voltax_amd as vx
    from voltax_amd import sparse

    from test_va_language import SMOOTHED

    def build(b):
        b.vsource("in", "0", 0.7, name="V1")
        b.resistor("in", "a", 300.0, name="R1")
        vm.place(b, vm.Pair(SMOOTHED, smooth=0.05), ("a", "0"), k=0.25)
        vm.place(b, vm.Pair(SMOOTHED, smooth={9: 0.02}), ("a", "0"), name="S2", k=0.1)
        return b.build()

    c, o = _both(build)
    st = osp.structure(o)
    rng = np.random.default_rng(9)
    for _ in range(3):
        z = rng.uniform(-0.5, 0.9, c.size)
        got = sparse.jacobian(c, z, 0.0, 1.0, 0.0)
        got = got.cpu().numpy() if hasattr(got, "cpu") else np.asarray(got)
        ref = osp.jacobian(o, z, 0.0, 1.0, 0.0, st)
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-15 * np.abs(ref).max())
    sol = vx.dc(c, options=vx.Options(**TIGHT))
    ref = oa.dc(o, options=oa.Options(**TIGHT))
    assert sol.converged and ref.converged
    np.testing.assert_allclose(sol.z, ref.z, rtol=1e-10, atol=1e-14)


def test_equality_limit_mode_on_the_gpu():
    """``equality="limit"``: at the special point the GPU interpreter keeps the
    special value and gives the general arm's derivative, like the oracle."""
    from voltax_amd import sparse

    from test_va_language import EQ2, EQ_CONTRIB, EQUALITY

    def build(b):
        b.resistor("a", "0", 1e3, name="R1")
        vm.place(b, vm.Pair(EQUALITY, equality="limit"), ("a", "0"))
        vm.place(b, vm.Pair(EQ2, equality="limit"), ("a", "0"))
        vm.place(b, vm.Pair(EQ2), ("a", "0"), name="E3")
        return b.build()

    def build_contrib(b):  # contributions inside the arms (currents, charges, branch terms)
        b.resistor("a", "0", 1e3, name="R1")
        vm.place(b, vm.Pair(EQ_CONTRIB, equality="limit"), ("a", "0"), name="X1")
        vm.place(b, vm.Pair(EQ_CONTRIB), ("a", "0"), name="X2")
        return b.build()

    cc, oc_ = _both(build_contrib)
    stc = osp.structure(oc_)
    for va_ in (0.0, 0.3):
        z = np.random.default_rng(1).uniform(-0.5, 0.5, cc.size)
        z[cc.node("a")] = va_
        for cf, cq in ((1.0, 0.0), (0.0, 1.0)):
            got = sparse.jacobian(cc, z, 0.0, cf, cq)
            got = got.cpu().numpy() if hasattr(got, "cpu") else np.asarray(got)
            ref = osp.jacobian(oc_, z, 0.0, cf, cq, stc)
            np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-18)
    c, o = _both(build)
    st = osp.structure(o)
    for va_ in (0.0, 0.25, 0.7):
        z = np.zeros(c.size)
        z[c.node("a")] = va_
        got = sparse.jacobian(c, z, 0.0, 1.0, 0.0)
        got = got.cpu().numpy() if hasattr(got, "cpu") else np.asarray(got)
        ref = osp.jacobian(o, z, 0.0, 1.0, 0.0, st)
        np.testing.assert_allclose(got, ref, rtol=1e-13, atol=1e-18)
        r = sparse.residual(c, z, 0.0, 1.0, 0.0)
        r = r.cpu().numpy() if hasattr(r, "cpu") else np.asarray(r)
        np.testing.assert_allclose(r, o.f(z), rtol=1e-13, atol=1e-18)
    z = np.zeros(c.size)
    got = sparse.jacobian(c, z, 0.0, 1.0, 0.0)
    got = got.cpu().numpy() if hasattr(got, "cpu") else np.asarray(got)
    # 1 mS (R1) + 1 mS (limit) + 7 mS (EQ2 limit) + 0 (EQ2 value) at the node
    assert got.reshape(-1)[0] == pytest.approx(1e-3 + 1e-3 + 7e-3, rel=1e-14)
