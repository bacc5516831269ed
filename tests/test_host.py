"""CPU tests of the product's host logic and of the C-ABI library (no GPU
compute): symbol exports, circuit layout / group naming vs the oracle,
and the C++ host pattern builder vs the oracle's CSR structure."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import circuits as tc
from oracle import circuit as oc
from oracle import sparse as osp

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(vxm_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from voltax_amd import _lib

    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    names = declared_symbols()
    assert "vxm_stamp" in names and "vxm_batch_run" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # every exported entry point has a Python signature
    assert set(names) <= set(_lib.SIGNATURES)


def test_status_strings():
    from voltax_amd._lib import lib

    assert lib.vxm_status_string(0) == b"ok"
    assert b"invalid" in lib.vxm_status_string(-1)


def test_invalid_arguments_are_rejected_without_gpu():
    from voltax_amd._lib import lib

    h = ctypes.c_void_p()
    kind = np.array([9], np.int32)  # unknown device kind
    z = np.zeros(4, np.int32)
    assert lib.vxm_topology_create(2, 2, 1, 1, kind.ctypes.data, z.ctypes.data,
                                   z.ctypes.data, z.ctypes.data, ctypes.byref(h)) == -1
    bad = np.array([0, 5, -1, -1], np.int32)  # unknown index out of range
    k0 = np.array([0], np.int32)
    assert lib.vxm_topology_create(2, 2, 1, 1, k0.ctypes.data, z.ctypes.data,
                                   bad.ctypes.data, z.ctypes.data, ctypes.byref(h)) == -1
    assert lib.vxm_stamp(None, 1, None, None, 0, 0, 0, 0, None, None, None, None, None) == -1


def _pair(fn, *a):
    import voltax_amd as vx

    return fn(vx.CircuitBuilder(), *a), fn(oc.CircuitBuilder(), *a)


@pytest.mark.parametrize("which", ["rc", "chain", "ring", "grid"])
def test_layout_and_pattern_match_oracle(which):
    import voltax_amd as vx
    from voltax_amd import sparse

    if which == "rc":
        c, o = _pair(tc.rc_ladder)
    elif which == "chain":
        c = tc.inverter_chain(vx.CircuitBuilder(), vx.Pulse(0, 1.2, 1e-10, 2e-11, 2e-11, 2e-9, 4e-9))
        o = tc.inverter_chain(oc.CircuitBuilder(), oc.Pulse(0, 1.2, 1e-10, 2e-11, 2e-11, 2e-9, 4e-9))
    elif which == "ring":
        c, o = tc.ring(vx.CircuitBuilder(), 7)[0], tc.ring(oc.CircuitBuilder(), 7)[0]
    else:
        c, o = _pair(tc.grid, 8)
    assert c.layout.node_names == o.layout.node_names
    assert c.layout.groups == o.layout.groups
    assert c.size == o.size
    st, ost = sparse.structure(c), osp.structure(o)
    assert np.array_equal(st.indptr, ost.indptr)
    assert np.array_equal(st.indices, ost.indices)
    assert np.array_equal(st.diagonal, ost.diagonal)


def test_device_table_parameter_rows():
    import voltax_amd as vx

    c = tc.inverter_chain(vx.CircuitBuilder(), vx.Pulse(0, 1.2, 1e-10, 2e-11, 2e-11, 2e-9, 4e-9))
    kind, wave, unk, poff, row = c.device_table()
    # groups sorted: Capacitor, EKVMOSFET_n, EKVMOSFET_p, VoltageSource, VoltageSource_Pulse
    assert list(kind) == [1] * 3 + [4] * 6 + [2, 2]
    assert list(wave[-2:]) == [1, 3]
    d = 3  # first NMOS
    p = row[poff[d]: poff[d] + 11]
    assert np.isclose(np.exp(p[0]), 1e-6) and p[3] == 1.0 and np.isclose(p[4], 400e-6)
    pm = row[poff[6]: poff[6] + 11]
    assert pm[3] == -1.0 and np.isclose(pm[4], 200e-6) and np.isclose(np.exp(pm[0]), 2e-6)
    pulse = row[poff[-1]: poff[-1] + 7]
    np.testing.assert_allclose(pulse, [0, 1.2, 1e-10, 2e-11, 2e-11, 2e-9, 4e-9])


def _emulate_multifrontal(plan, S, data, b, transpose=False):
    """numpy replay of the HIP multifrontal factor / solve kernels, on the
    plan's exported layout (checks the symbolic layer without a GPU)."""
    L = plan.export()
    sn, sn2 = L["sn"], L["sn2"]
    fronts, pivs = {}, {}
    order = L["level_sn"]
    for s in order:  # level order = kernel launch order
        k, r, par, lev, foff, poff, ioff, coff, choff, nch = sn[s]
        nf = k + r
        F = np.zeros(nf * nf)
        ao, na = sn2[s, 0], sn2[s, 1]
        np.add.at(F, L["apos"][ao:ao + na], data[L["aslot"][ao:ao + na]])
        F = F.reshape(nf, nf)
        for c in L["children"][choff:choff + nch]:
            kc, rc = sn[c, 0], sn[c, 1]
            cm = L["cmap"][sn[c, 7]:sn[c, 7] + rc]
            F[np.ix_(cm, cm)] += fronts[c][kc:, kc:]
        piv = np.zeros(k, int)
        for j in range(k):
            p = j + int(np.argmax(np.abs(F[j:k, j])))
            piv[j] = p
            F[[j, p]] = F[[p, j]]
            F[j + 1:, j] /= F[j, j]
            F[j + 1:, j + 1:] -= np.outer(F[j + 1:, j], F[j, j + 1:])
        fronts[s], pivs[s] = F, piv
    x = b.astype(float).copy()
    uvec = {}
    for s in order:
        k, r = sn[s, 0], sn[s, 1]
        idx = L["idx"][sn[s, 6]:sn[s, 6] + k + r]
        F = fronts[s]
        v = np.concatenate([x[idx[:k]], np.zeros(r)])
        for c in L["children"][sn[s, 8]:sn[s, 8] + sn[s, 9]]:
            cm = L["cmap"][sn[c, 7]:sn[c, 7] + sn[c, 1]]
            v[cm] += uvec[c]
        if not transpose:
            for j in range(k):
                v[[j, pivs[s][j]]] = v[[pivs[s][j], j]]
            for j in range(k):
                v[j + 1:] -= F[j + 1:, j] * v[j]
        else:
            for j in range(k):
                v[j] /= F[j, j]
                v[j + 1:] -= F[j, j + 1:] * v[j]
        x[idx[:k]] = v[:k]
        uvec[s] = v[k:]
    for s in order[::-1]:
        k, r = sn[s, 0], sn[s, 1]
        idx = L["idx"][sn[s, 6]:sn[s, 6] + k + r]
        F = fronts[s]
        if not transpose:
            v = x[idx].copy()
            v[:k] -= F[:k, k:] @ v[k:]
            for j in range(k - 1, -1, -1):
                v[j] /= F[j, j]
                v[:j] -= F[:j, j] * v[j]
            x[idx[:k]] = v[:k]
        else:
            v = x[idx].copy()
            v[:k] -= F[k:, :k].T @ v[k:]
            for j in range(k - 1, -1, -1):
                v[:j] -= F[j, :j] * v[j]
            for j in range(k - 1, -1, -1):
                v[[j, pivs[s][j]]] = v[[pivs[s][j], j]]
            x[idx[:k]] = v[:k]
    return x


@pytest.mark.parametrize("which,leaf", [("grid", 1), ("grid", 16), ("chain", 1),
                                        ("ring", 3), ("grid", 4)])
def test_multifrontal_plan_replay_matches_lapack(which, leaf):
    import voltax_amd as vx
    from voltax_amd import engine, splu

    if which == "grid":
        c, o = _pair(tc.grid, 9)
    elif which == "chain":
        c = tc.inverter_chain(vx.CircuitBuilder(), vx.Pulse(0, 1.2, 1e-10, 2e-11, 2e-11, 2e-9, 4e-9))
        o = tc.inverter_chain(oc.CircuitBuilder(), oc.Pulse(0, 1.2, 1e-10, 2e-11, 2e-11, 2e-9, 4e-9))
    else:
        c, o = tc.ring(vx.CircuitBuilder(), 9)[0], tc.ring(oc.CircuitBuilder(), 9)[0]
    topo, _ = engine.topology_of(c)
    plan = splu.Plan(topo, leaf)
    ost = osp.structure(o)
    z = np.random.default_rng(4).uniform(0, 1.2, c.size)
    data = osp.jacobian(o, z, 0.0, 1.0, 1e12, ost)
    data[ost.diagonal[: o.n_nodes]] += 1e-12
    A = ost.to_dense(data)
    b = np.random.default_rng(0).standard_normal(c.size)
    for tr in (False, True):
        x = _emulate_multifrontal(plan, c.size, data, b, tr)
        ref = np.linalg.solve(A.T if tr else A, b)
        np.testing.assert_allclose(x, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())
    if which == "grid" and leaf == 16:
        assert plan.stats["supernodes"] > 1 and plan.stats["levels"] > 2


def test_grid_inverter_variants_write_the_right_parameter_slots():
    """A variant's parameter row (written slot by slot into the one pattern)
    equals the row of the same circuit built device by device with the
    variant's values, and the topology is the same."""
    import voltax_amd as vx
    from voltax_amd.workloads import GridInverters

    fam = GridInverters(vx.CircuitBuilder, nx=6, ny=5, n_inv=7, layout_seed=3)
    P, F = fam.variants(3, seed=11)
    nr, ni = len(fam.r_slots), len(fam.i_slots)
    assert F.shape == (3, nr + ni + 7)
    for v in range(3):
        log_r, sink, bits = F[v, :nr], F[v, nr:nr + ni] * 1e-3, F[v, nr + ni:]
        c = tc.grid_inverters(vx.CircuitBuilder(), log_r, sink, bits, 6, 5, fam.sites)
        kind, wave, unk, poff, row = c.device_table()
        k0, w0, u0, p0, _ = fam.circuit.device_table()
        np.testing.assert_array_equal(kind, k0)
        np.testing.assert_array_equal(unk, u0)
        np.testing.assert_allclose(P[v], row, rtol=1e-15, atol=0)


def test_ac_stimulus_matches_oracle():
    """Phasors of AC sources land on the V-source branch rows (negated) and
    through the I-source terminals, as voltax/elements/sources.py:87-116."""
    import voltax_amd as vx

    def build(b):
        b.vsource("in", "0", 0.5, ac=1.0, ac_phase=30.0, name="Vin")
        b.resistor("in", "a", 1e3)
        b.capacitor("a", "0", 1e-12)
        b.isource("a", "0", 1e-3, ac=0.2, ac_phase=-45.0, name="Ia")
        b.isource("0", "in", 0.0, ac=0.3, name="Ib")
        b.resistor("a", "0", 2e3)
        return b.build()

    c, o = build(vx.CircuitBuilder()), build(oc.CircuitBuilder())
    np.testing.assert_allclose(c.ac_stimulus(), o.ac_stimulus(), rtol=0, atol=1e-15)
    assert abs(c.ac_stimulus()[c.size - 1] + np.exp(1j * np.pi / 6)) < 1e-15


README_CHAIN = """
.subckt inv in out vdd
M1 out in vdd vdd pmos w=2u l=0.13u
M2 out in 0   0   nmos w=1u l=0.13u
Cl out 0 5f
.ends
Vdd vdd 0 1.2
Vin in  0 PULSE(0 1.2 0.2n 20p 20p 2n 4n)
X1 in a vdd inv
X2 a  b vdd inv
X3 b out vdd inv
"""


def test_netlist_of_the_readme_chain_equals_the_builder_circuit():
    """README.md:60-71 of the reference as text: the same nodes, device names
    (X1.M1 ...), groups, parameter rows and topology as the builder calls."""
    import voltax_amd as vx

    c = vx.parse_netlist(README_CHAIN)
    ref = tc.inverter_chain(vx.CircuitBuilder(), vx.Pulse(0, 1.2, 0.2e-9, 20e-12, 20e-12, 2e-9, 4e-9))
    assert c.layout.node_names == ref.layout.node_names
    assert [d for d, _, _ in c.layout.devices] == [d for d, _, _ in ref.layout.devices]
    assert set(c.elements) == set(ref.elements)
    assert c.topology_key() == ref.topology_key()
    np.testing.assert_array_equal(c.device_table()[4], ref.device_table()[4])


@pytest.mark.parametrize("text,value", [
    ("10f", 10e-15), ("1.2u", 1.2e-6), ("4.7k", 4.7e3), ("2meg", 2e6),
    ("10pF", 10e-12), ("1e-3", 1e-3), ("3", 3.0), ("5mil", 5 * 25.4e-6),
])
def test_parse_value(text, value):
    """tests/test_builder_netlist.py:55-60 of the reference."""
    import voltax_amd as vx

    assert vx.parse_value(text) == pytest.approx(value)


def test_netlist_features_and_errors():
    import voltax_amd as vx

    c = vx.parse_netlist("""* title-less deck
.param rl=2k half={rl/2}
.model nch nmos (level=1 kp=300u vto=0.45 lambda=0.02)
.subckt div top out r1=1k
Ra top out {r1}
Rb out 0
+ {half}
.ends div
V1 s 0 DC 1.5 AC 1 90
I1 0 s SIN(0 1m 1k)
X1 s o div r1=3k
M1 d g 0 0 nch w=2u
Vg g 0 0.9
Vd d 0 PWL(0 0 1n 1.2)
.tran 1n 10n
.control
run
.endc
.end
R99 never parsed 1
""")
    assert c.get("X1.Ra", "r") == pytest.approx(3e3)
    assert c.get("X1.Rb", "r") == pytest.approx(1e3)
    assert "X1.out" not in c.layout.node_names and "o" in c.layout.node_names
    el, i = c.device("V1")
    assert el.ac[i] == 1.0 and el.ac_phase[i] == 90.0
    m, _ = c.device("M1")
    assert m.process.kp == pytest.approx(300e-6) and m.process.vth == pytest.approx(0.45)
    assert type(c.device("Vd")[0].waves[0]).__name__ == "PWL"
    for text, match in [("R1 a b", "missing resistance"), ("Z1 a b 1", "unsupported element"),
                        (".foo", "unsupported directive"), ("X1 a b nosuch", "unknown subcircuit"),
                        ("M1 a b c d nope", "unknown model")]:
        with pytest.raises(ValueError, match=f"^<netlist>:2: .*{match}"):
            vx.parse_netlist("R0 x 0 1\n" + text)


def test_netlist_level_models_hook_builds_verilog_a_devices():
    """tests/test_va.py:449-467 of the reference (the model hook): a
    ``level=99`` card becomes a Verilog-A element group named after the card
    and polarity, with the card's parameters and the instance's w and l."""
    import voltax_amd as vx

    import va_models as vm

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
    el, i = c.device("Mn")
    assert el.card["vth"] == 0.3 and el.card["kp"] == 100e-6 and el.card["type"] == 1
    assert el.get("w", i) == 1e-6 and el.get("l", i) == 100e-9


def test_options_reuse_and_circuit_linearity():
    """Options.reuse needs the FP64 LU; Circuit.is_linear (the condition for
    keeping LU factors across time steps on the sparse path) holds for
    resistor / capacitor / source circuits only."""
    import voltax_amd as vx

    assert vx.Options(reuse=True).reuse
    with pytest.raises(ValueError, match="precision"):
        vx.Options(reuse=True, precision="mixed")
    b = vx.CircuitBuilder()
    b.vsource("a", "0", vx.Pulse(0, 1, 1e-9, 1e-10, 1e-10, 1e-9, 4e-9), name="V1")
    b.resistor("a", "b", 1e3, name="R1")
    b.capacitor("b", "0", 1e-12, name="C1")
    b.isource("b", "0", 1e-6, name="I1")
    assert b.build().is_linear
    b.nmos("b", "a", "0", "0", w=1e-6, l=0.13e-6, name="M1")
    assert not b.build().is_linear


def _plan_case(which):
    import voltax_amd as vx
    from voltax_amd import workloads

    if which.startswith("grid") and which[4:].isdigit():
        return _pair(tc.grid, int(which[4:]))
    if which == "chain":
        return (tc.inverter_chain(vx.CircuitBuilder(), 0.0),
                tc.inverter_chain(oc.CircuitBuilder(), 0.0))
    if which == "ring9":
        return tc.ring(vx.CircuitBuilder(), 9)[0], tc.ring(oc.CircuitBuilder(), 9)[0]
    if which.startswith("power"):
        n = int(which[5:])
        return (workloads.power_grid(vx.CircuitBuilder(), n, n),
                workloads.power_grid(oc.CircuitBuilder(), n, n))
    if which == "gridinv":
        fam = workloads.GridInverters(vx.CircuitBuilder)
        P, F = fam.variants(1, seed=5)
        nr, ni = len(fam.r_slots), len(fam.i_slots)
        args = (F[0, :nr], F[0, nr:nr + ni] * 1e-3, F[0, nr + ni:], fam.nx, fam.ny, fam.sites)
        return (tc.grid_inverters(vx.CircuitBuilder(), *args),
                tc.grid_inverters(oc.CircuitBuilder(), *args))
    if which == "va_grid":
        import va_models as vm

        diode, rl = vm.Pair(vm.DIODE, {"rs": 5.0}), vm.Pair(vm.RL)
        return (tc.va_grid(vx.CircuitBuilder(), 12, vm.place, diode, rl),
                tc.va_grid(oc.CircuitBuilder(), 12, vm.place, diode, rl))
    raise KeyError(which)


@pytest.mark.parametrize("which", ["grid12", "grid40", "chain", "ring9", "power30", "power100",
                                   "gridinv", "va_grid"])
def test_symbolic_plan_equals_the_reference_plan(which):
    """The C++ symbolic analysis (pivot groups, nested dissection, postorder,
    borders, levels) gives exactly the plan of voltax/sparse.py:257-535 as
    the oracle restates it with scipy: the same supernode sets in the same
    order, the same parents, levels, borders and LU flop count."""
    from voltax_amd import engine, splu

    c, o = _plan_case(which)
    assert c.size == o.size
    topo, _ = engine.topology_of(c)
    plan = splu.Plan(topo, 16)
    sn, par, lev = plan.structure()
    ost = osp.structure(o)
    np.testing.assert_array_equal(topo.indptr, ost.indptr)
    np.testing.assert_array_equal(topo.indices, ost.indices)
    ref = osp.make_plan(ost, 16)
    assert plan.n_supernodes == len(ref.supernodes)
    for i, s in enumerate(ref.supernodes):
        np.testing.assert_array_equal(np.flatnonzero(sn == i), s)
    np.testing.assert_array_equal(par, ref.parents)
    ref_level = np.zeros(len(ref.parents), np.int64)
    for i, p in enumerate(ref.parents):  # postorder: children come first
        if p >= 0:
            ref_level[p] = max(ref_level[p], ref_level[i] + 1)
    np.testing.assert_array_equal(lev, ref_level)
    assert plan.n_levels == ref.stats["levels"]
    ex = plan.export()
    for i, b in enumerate(ref.borders):
        k, r, off = ex["sn"][i, 0], ex["sn"][i, 1], ex["sn"][i, 6]
        np.testing.assert_array_equal(ex["idx"][off + k: off + k + r], b)
    k = np.array([len(s) for s in ref.supernodes], float)
    r = np.array([len(b) for b in ref.borders], float)
    flops = float(np.sum(2.0 / 3.0 * k ** 3 + 2.0 * k * k * r + 2.0 * k * r * r))
    assert plan.stats["lu_flops"] == pytest.approx(flops, rel=1e-12)
    if which in ("grid40", "power100", "va_grid"):
        assert plan.n_supernodes > 10 and plan.n_levels > 4


@pytest.mark.parametrize("which", ["grid12", "chain", "ring9", "power30", "gridinv", "va_grid"])
def test_compact_grouping_plan_replay_matches_lapack(which):
    """The compact pivot groups (VXM_SPLU_COMPACT) factor the circuit's own
    Jacobian at a random state as LAPACK solves it, both orientations."""
    from voltax_amd import engine, splu

    c, o = _plan_case(which)
    topo, _ = engine.topology_of(c)
    ost = osp.structure(o)
    np.testing.assert_array_equal(topo.indptr, ost.indptr)
    np.testing.assert_array_equal(topo.indices, ost.indices)
    plan = splu.Plan(topo, 16, compact=True)
    z = np.random.default_rng(6).uniform(0, 1.2, c.size)
    data = osp.jacobian(o, z, 0.0, 1.0, 1e12, ost)
    data[ost.diagonal[: o.n_nodes]] += 1e-12
    A = ost.to_dense(data)
    b = np.random.default_rng(1).standard_normal(c.size)
    for tr in (False, True):
        x = _emulate_multifrontal(plan, c.size, data, b, tr)
        ref = np.linalg.solve(A.T if tr else A, b)
        np.testing.assert_allclose(x, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())


def test_compact_grouping_splits_the_bsim_sized_ring():
    """The reference's pivot groups glue every device with internal nodes to
    its terminals (voltax/sparse.py:274-300); on a ring of the BSIM-sized
    model all PMOS share vdd, so the whole ring is one front.  The compact
    groups keep fronts small and still replay LAPACK on data with the
    supply's branch row structurally zero on its diagonal."""
    import voltax_amd as vx
    from voltax_amd import engine, splu, va, workloads as wl

    m = va.load(wl.BSIM_SIZED_VA)
    N = m.element(polarity="n", differentiable=["vth0"], name="bsim_n")
    P = m.element(polarity="p", differentiable=["vth0"], name="bsim_p")
    c, _ = wl.bsim_ring(vx.CircuitBuilder(), 51, N, P)
    topo, _ = engine.topology_of(c)
    ref = splu.Plan(topo, 16)
    cp = splu.Plan(topo, 16, compact=True)
    assert ref.stats["largest_front"] == c.size
    assert cp.stats["largest_front"] <= 16
    assert cp.stats["lu_flops"] < 1e-3 * ref.stats["lu_flops"]
    S = c.size
    rng = np.random.default_rng(3)
    A = np.zeros((S, S))
    for r in range(S):
        cols = topo.indices[topo.indptr[r]:topo.indptr[r + 1]]
        A[r, cols] = rng.uniform(-1.0, 1.0, len(cols))
    nv = c.n_nodes
    A[np.arange(nv), np.arange(nv)] += 4.0 * np.abs(A[:nv]).sum(1)
    br = np.arange(nv, S)
    A[br, br] = 0.0  # branch currents: no diagonal
    assert len(br) >= 1
    data = np.concatenate([A[r, topo.indices[topo.indptr[r]:topo.indptr[r + 1]]]
                           for r in range(S)])
    b = rng.standard_normal(S)
    for tr in (False, True):
        x = _emulate_multifrontal(cp, S, data, b, tr)
        want = np.linalg.solve(A.T if tr else A, b)
        np.testing.assert_allclose(x, want, rtol=1e-9, atol=1e-12 * np.abs(want).max())


def test_pivot_matching_equals_scipy_on_random_graphs():
    """The C++ maximum bipartite matching behind the pivot groups picks the
    same matching as scipy's maximum_bipartite_matching (which the reference
    calls, voltax/sparse.py:291-295) on random graphs with many maximum
    matchings: floating voltage sources between random node pairs."""
    import voltax_amd as vx
    from voltax_amd import engine, splu

    rng = np.random.default_rng(12)
    for trial in range(40):
        nn = int(rng.integers(8, 40))
        nsrc = int(rng.integers(2, nn))

        def build(b):
            for k in range(nn - 1):
                b.resistor(f"n{k}", f"n{k + 1}", 1.0 + k)
            pairs = np.random.default_rng(trial).integers(0, nn, (nsrc, 2))
            for a, c in pairs:
                if a != c:
                    b.vsource(f"n{a}", f"n{c}", 0.1)
            b.vsource("n0", "0", 1.0)
            return b.build()

        c, o = build(vx.CircuitBuilder()), build(oc.CircuitBuilder())
        topo, _ = engine.topology_of(c)
        plan = splu.Plan(topo, 2)
        sn, par, _ = plan.structure()
        ref = osp.make_plan(osp.structure(o), 2)
        assert plan.n_supernodes == len(ref.supernodes), trial
        for i, s in enumerate(ref.supernodes):
            np.testing.assert_array_equal(np.flatnonzero(sn == i), s)
        np.testing.assert_array_equal(par, ref.parents)


def test_noalloc_stamp_rejects_on_the_host():
    """vxm_stamp_noalloc checks its arguments before any GPU call: a batch
    beyond the fixed scratch returns VXM_ERR_TOO_LARGE, a topology whose
    tables are not yet resident returns VXM_ERR_INVALID (it never uploads,
    since that allocates)."""
    import voltax_amd as vx
    from voltax_amd import engine
    from voltax_amd._lib import lib

    c = tc.ring(vx.CircuitBuilder(), 5)[0]
    kind, wave, unk, poff, row = c.device_table()
    topo = engine.Topology(c.size, c.n_nodes, kind, wave, unk, poff, len(row))  # fresh handle
    z = np.zeros(c.size)
    res = np.zeros(c.size)
    big = 4 * 1024 * 1024 // topo.work + 1
    assert lib.vxm_stamp_noalloc(topo.handle, big, row.ctypes.data, z.ctypes.data, 0.0, 1.0,
                                 0.0, 0.0, None, None, res.ctypes.data, None) == -5
    assert lib.vxm_stamp_noalloc(topo.handle, 1, row.ctypes.data, z.ctypes.data, 0.0, 1.0,
                                 0.0, 0.0, None, None, res.ctypes.data, None) == -1
    assert lib.vxm_stamp_noalloc(topo.handle, 1, row.ctypes.data, z.ctypes.data, 0.0, 1.0,
                                 0.0, 0.0, None, None, None, None) == -1
    assert lib.vxm_status_string(-5).decode().startswith("batch too large")


def test_error_paths_return_status_codes_without_a_gpu():
    """Entry points report failures as status codes and never throw or abort
    through the C ABI: a batch cannot be created without a device (HIP error,
    no handle returned), runs and timings on a NULL batch are invalid, and
    every status code has a name."""
    import torch

    import voltax_amd as vx
    from voltax_amd import engine
    from voltax_amd._lib import NewtonOptions, lib

    if torch.cuda.is_available():
        pytest.skip("needs a host without a GPU")
    c = tc.ring(vx.CircuitBuilder(), 5)[0]
    kind, wave, unk, poff, row = c.device_table()
    topo = engine.Topology(c.size, c.n_nodes, kind, wave, unk, poff, len(row))
    handles = (ctypes.c_void_p * 1)(topo.handle.value)
    inst_topo = np.zeros(1, np.int32)
    inst_off = np.zeros(1, np.int64)
    out = ctypes.c_void_p(12345)
    st = lib.vxm_batch_create(1, handles, 1, inst_topo.ctypes.data, inst_off.ctypes.data,
                              ctypes.byref(out))
    assert st == -2 and not out.value
    opts = NewtonOptions(100, 1e-6, 1e-9, 1.0, 1e-12, 10, 1e-2, 25)
    assert lib.vxm_batch_run(None, None, None, None, 1, 0, 1, 0.0, ctypes.byref(opts), 0,
                             None, None, None, None) == -1
    assert lib.vxm_batch_timings(None, None) == -1
    names = {s: lib.vxm_status_string(s).decode() for s in range(-6, 1)}
    assert names[-6] == "host memory exhausted" and names[-2] == "HIP runtime error"
    assert "unknown" not in " ".join(names.values())


def test_empty_batches_return_ok_without_launching():
    """A batch of 0 instances checks its arguments and enqueues nothing: the
    stamping, product, update and sparse LU entry points return VXM_OK on a
    host without a GPU (any launch or upload would fail there), and invalid
    arguments are still reported."""
    import torch

    import voltax_amd as vx
    from voltax_amd import engine
    from voltax_amd._lib import NewtonOptions, lib
    from voltax_amd.splu import Plan

    if torch.cuda.is_available():
        pytest.skip("needs a host without a GPU")
    c = tc.ring(vx.CircuitBuilder(), 5)[0]
    kind, wave, unk, poff, row = c.device_table()
    topo = engine.Topology(c.size, c.n_nodes, kind, wave, unk, poff, len(row))
    plan = Plan(topo)
    buf = np.zeros(16)
    p = buf.ctypes.data
    piv = np.zeros(4, np.int32).ctypes.data
    assert lib.vxm_stamp(topo.handle, 0, p, p, 0.0, 1.0, 0.0, 0.0, None, p, p, p, None) == 0
    assert lib.vxm_spmv(topo.handle, 0, p, p, p, None) == 0
    assert lib.vxm_spmv(topo.handle, -1, p, p, p, None) == -1
    assert lib.vxm_stamp_noalloc(topo.handle, 0, p, p, 0.0, 1.0, 0.0, 0.0, None, p, p,
                                 None) == 0
    assert lib.vxm_splu_factor(plan.handle, 0, p, p, piv, 0, None) == 0
    assert lib.vxm_splu_factor(plan.handle, 0, p, p, piv, 3, None) == -1
    assert lib.vxm_splu_solve(plan.handle, 0, p, piv, p, p, 0, 0, None) == 0
    assert lib.vxm_splu_factor_solve(plan.handle, 0, p, p, piv, p, p, None) == 0
    assert lib.vxm_splu_factor_solve(plan.handle, 0, None, p, piv, p, p, None) == -1
    assert lib.vxm_splu_solve_refined(plan.handle, 0, p, p, piv, p, p, p, 1, 0, 0, None) == 0
    opts = NewtonOptions(100, 1e-6, 1e-9, 1.0, 1e-12, 10, 1e-2, 25)
    done = np.zeros(1, np.int8).ctypes.data
    assert lib.vxm_newton_update(0, c.size, c.n_nodes, p, p, 1.0, ctypes.byref(opts), None,
                                 done, p, None) == 0
