"""GPU parity of the stamping kernels and the fused batched Newton kernel
against the oracle (CPU restatement of the reference).

Tolerances: Jacobian / residual values 1e-12 relative (summation order is
the only difference); node voltages and waveforms 1e-9 relative to the
waveform scale, with both sides converging Newton to rtol=1e-12.
"""

import numpy as np
import pytest

import circuits as tc
from oracle import analysis as oa
from oracle import circuit as oc
from oracle import measure as om
from oracle import sparse as osp

pytestmark = pytest.mark.gpu

TIGHT = dict(rtol=1e-12, atol=1e-15)


def _vx():
    import voltax_amd as vx

    return vx


def both(fn, *a, **k):
    vx = _vx()
    return fn(vx.CircuitBuilder(), *a, **k), fn(oc.CircuitBuilder(), *a, **k)


def pulse(mod):
    return mod.Pulse(0, 1.2, 0.2e-9, 20e-12, 20e-12, 2e-9, 4e-9)


@pytest.mark.parametrize("name", ["rc_ladder", "chain", "ring5", "grid", "star"])
def test_stamped_jacobian_and_residual_match_oracle(name):
    """Stamped Jacobian values and residuals against the oracle; "star" has a
    row, a diagonal slot and a pattern row of more than 32 terms, which the
    gather and product kernels sum by whole wavefronts (also checked for the
    batched SpMV against the oracle's matvec)."""
    vx = _vx()
    from voltax_amd import sparse

    if name == "rc_ladder":
        c, o = both(tc.rc_ladder)
    elif name == "chain":
        c = tc.inverter_chain(vx.CircuitBuilder(), pulse(vx))
        o = tc.inverter_chain(oc.CircuitBuilder(), pulse(oc))
    elif name == "ring5":
        c, o = tc.ring(vx.CircuitBuilder(), 5)[0], tc.ring(oc.CircuitBuilder(), 5)[0]
    elif name == "star":
        c, o = both(tc.star, 80)
    else:
        c, o = both(tc.grid, 6)
    z = np.random.default_rng(0).uniform(-0.5, 1.5, c.size)
    st = sparse.structure(c)
    ost = osp.structure(o)
    assert np.array_equal(st.indptr, ost.indptr) and np.array_equal(st.indices, ost.indices)
    for f, q, t in [(1.0, 0.0, 0.3e-9), (0.0, 1.0, 0.0), (0.5, 1e12, 2.1e-9)]:
        got = sparse.jacobian(c, z, t, f, q).cpu().numpy()
        want = osp.jacobian(o, z, t, f, q, ost)
        scale = np.abs(want).max()
        np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-14 * scale)
        rg = sparse.residual(c, z, t, f, q).cpu().numpy()
        rw = f * o.f(z, t) + q * o.q(z)
        np.testing.assert_allclose(rg, rw, rtol=1e-12, atol=1e-14 * np.abs(rw).max())
    if name == "star":
        hub = c.node("hub")
        assert st.indptr[hub + 1] - st.indptr[hub] > 32
        xs = np.random.default_rng(1).standard_normal((3, c.size))
        ds = np.stack([osp.jacobian(o, z + 0.1 * k, 0.0, 1.0, 1e12, ost) for k in range(3)])
        y = st.matvec(ds, xs).cpu().numpy()
        for k in range(3):
            want = ost.matvec(ds[k], xs[k])
            terms = ost.matvec(np.abs(ds[k]), np.abs(xs[k]))
            assert np.all(np.abs(y[k] - want) <= 8 * np.finfo(float).eps * terms)


def test_rc_ladder_golden_and_gradient_free_transient():
    """Golden state of the reference (tests/test_analysis.py:33-39)."""
    vx = _vx()
    c = tc.rc_ladder(vx.CircuitBuilder())
    ts = np.linspace(0, 1e-5, 101)
    sol = vx.transient(c, ts, ic="zero")
    assert sol.converged.all()
    np.testing.assert_allclose(sol.z[-1], [1.0, 0.921635241, 0.766714055, -7.83647588e-05],
                               atol=1e-8)
    o = tc.rc_ladder(oc.CircuitBuilder())
    ref = oa.transient(o, ts, ic="zero", options=oa.Options(**TIGHT))
    got = vx.transient(c, ts, ic="zero", options=vx.Options(**TIGHT))
    np.testing.assert_allclose(got.z, ref.z, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("method", ["be", "trap"])
def test_inverter_chain_delay_matches_oracle(method):
    """Config 1: 3-stage EKV chain, DC initial point then 1000 steps."""
    vx = _vx()
    ts = np.linspace(0, 1e-9, 1001)
    c = tc.inverter_chain(vx.CircuitBuilder(), pulse(vx))
    o = tc.inverter_chain(oc.CircuitBuilder(), pulse(oc))
    got = vx.transient(c, ts, method=method, options=vx.Options(**TIGHT))
    ref = oa.transient(o, ts, method=method, options=oa.Options(**TIGHT))
    assert got.converged.all() and ref.converged.all()
    np.testing.assert_allclose(got.z, ref.z, rtol=0, atol=1e-9 * 1.2)
    d_got = vx.measure.delay(ts, got.v("in"), got.v("out"), 0.6, targ_direction="fall")
    d_ref = om.delay(ts, ref.v("in"), ref.v("out"), 0.6, targ_direction="fall")
    assert abs(d_got - d_ref) <= 1e-9 * abs(d_ref)
    if method == "be":
        assert round(d_got * 1e12, 1) == 43.5  # README.md:78


def test_dc_operating_point_matches_oracle():
    vx = _vx()
    c = tc.inverter_chain(vx.CircuitBuilder(), pulse(vx))
    o = tc.inverter_chain(oc.CircuitBuilder(), pulse(oc))
    got = vx.dc(c, 0.0, options=vx.Options(**TIGHT))
    ref = oa.dc(o, 0.0, options=oa.Options(**TIGHT))
    assert got.converged and ref.converged
    np.testing.assert_allclose(got.z, ref.z, rtol=1e-9, atol=1e-12)


def _rings(mod, ns, seed):
    rng = np.random.default_rng(seed)
    out = []
    for n in ns:
        dv = rng.normal(0, 0.03, 2 * n)
        ws = np.exp(rng.normal(0, 0.05, 2 * n))
        c, nodes = tc.ring(mod.CircuitBuilder(), n, dvth=dv, wscale=ws)
        out.append((c, nodes))
    return out


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_ring_batch_uneven_sizes_match_oracle(precision):
    """Uneven ring oscillators (3..33 unknowns) in one launch vs the oracle."""
    vx = _vx()
    ns = [3, 5, 7, 13, 15, 21, 31, 9, 3]
    ts = np.linspace(0, 100e-12, 51)
    mine = _rings(vx, ns, 1)
    theirs = _rings(oc, ns, 1)
    opts = vx.Options(precision=precision, **TIGHT)
    sols = vx.transient_batch([c for c, _ in mine], ts,
                              ic=[tc.ring_ic(c, nd) for c, nd in mine], options=opts)
    for (o, nd), sol in zip(theirs, sols):
        ref = oa.transient(o, ts, ic=tc.ring_ic(o, nd), options=oa.Options(**TIGHT))
        assert sol.converged.all()
        np.testing.assert_allclose(sol.z[:, : o.n_nodes], ref.z[:, : o.n_nodes], rtol=0,
                                   atol=1e-9 * 1.2)


def test_concurrent_tile_classes_repeat_bit_for_bit():
    """768 rings of every size (3..31 stages: all eight tile classes, launched
    concurrently on their own streams) run twice on the plain FP64 path and
    twice with the extrapolated start and reused factors: each pair agrees
    bit for bit (devices add their Jacobian and residual entries with LDS
    atomics, whose order is fixed per wave), and every 8th ring matches the
    compiled CPU restatement of the reference dense path (1e-12 V for FP64,
    1e-9 of VDD for the reused-factor path)."""
    import torch

    vx = _vx()
    from oracle import cpu_dense
    from voltax_amd import workloads as wl
    from voltax_amd.engine import DenseBatch

    specs = wl.ring_batch_specs(768, seed=3)
    ts = np.arange(101) * 2e-12
    circuits, ics, _ = wl.build_rings(vx.CircuitBuilder, specs)
    batch = DenseBatch(circuits)
    z0 = torch.as_tensor(np.concatenate(ics)).cuda()
    tsd = torch.as_tensor(ts).cuda()
    for pr in (False, True):
        outs = []
        for _ in range(2):
            z, conv, _ = batch.run(z0, tsd, False, 0.0, vx.Options(), "be", predict=pr,
                                   reuse=pr)
            assert bool(conv.all())
            outs.append(z.cpu().numpy())
        assert np.array_equal(outs[0], outs[1])
        zs = batch.split(z, len(ts))
        for i in range(0, len(specs), 8):
            c = circuits[i]
            zc, _, conv_c = cpu_dense.transient_be(c.size, c.n_nodes, c.device_table(),
                                                   ics[i], ts)
            assert conv_c.all()
            assert np.abs(zs[i] - zc).max() <= (1e-9 * 1.2 if pr else 1e-12)


@pytest.mark.parametrize("method", ["be", "trap"])
def test_predicted_start_keeps_the_roots_and_saves_iterations(method):
    """The extrapolated Newton start (Options.predict: the cubic through the
    last four points) changes only the initial guess: the rings and the pulsed
    inverter chain keep the oracle's waveforms and need fewer iterations than
    the last-point start."""
    vx = _vx()
    ns = [3, 7, 15, 31, 21]
    ts = np.linspace(0, 200e-12, 101)
    mine = _rings(vx, ns, 2)
    theirs = _rings(oc, ns, 2)
    pulse = lambda m: m.Pulse(0, 1.2, 20e-12, 20e-12, 20e-12, 60e-12, 200e-12)  # noqa: E731
    chain, ochain = tc.inverter_chain(vx.CircuitBuilder(), pulse(vx)), \
        tc.inverter_chain(oc.CircuitBuilder(), pulse(oc))
    circuits = [c for c, _ in mine] + [chain]
    ics = [tc.ring_ic(c, nd) for c, nd in mine] + [chain.zeros()]
    plain = vx.transient_batch(circuits, ts, ic=ics, method=method, options=vx.Options(**TIGHT))
    pred = vx.transient_batch(circuits, ts, ic=ics, method=method,
                              options=vx.Options(predict=True, **TIGHT))
    refs = [oa.transient(o, ts, ic=tc.ring_ic(o, nd), method=method, options=oa.Options(**TIGHT))
            for o, nd in theirs]
    refs.append(oa.transient(ochain, ts, ic=ochain.zeros(), method=method,
                             options=oa.Options(**TIGHT)))
    for a, b, ref in zip(plain, pred, refs):
        assert a.converged.all() and b.converged.all()
        np.testing.assert_allclose(b.z, ref.z, rtol=1e-9, atol=1e-12)
        np.testing.assert_allclose(b.z, a.z, rtol=1e-10, atol=1e-13)
    it_plain = sum(int(s.iters[1:].sum()) for s in plain)
    it_pred = sum(int(s.iters[1:].sum()) for s in pred)
    if method == "be":
        assert it_pred < 0.8 * it_plain
    else:  # the extrapolated start applies to backward Euler only
        assert it_pred == it_plain


@pytest.mark.parametrize("method", ["be", "trap"])
@pytest.mark.parametrize("predict", [False, True])
def test_reused_factors_keep_the_roots_and_save_factorizations(method, predict):
    """Options.reuse (later Newton iterations of a time step reuse its LU
    factors while the updates contract): the rings and the pulsed inverter
    chain keep the oracle's waveforms, at the tight tolerance and at the
    default one (there against the plain FP64 path to 1e-9 of VDD), with
    fewer LU factorizations than iterations."""
    vx = _vx()
    ns = [3, 7, 15, 31, 21]
    ts = np.linspace(0, 200e-12, 101)
    mine = _rings(vx, ns, 2)
    theirs = _rings(oc, ns, 2)
    pulse = lambda m: m.Pulse(0, 1.2, 20e-12, 20e-12, 20e-12, 60e-12, 200e-12)  # noqa: E731
    chain, ochain = tc.inverter_chain(vx.CircuitBuilder(), pulse(vx)), \
        tc.inverter_chain(oc.CircuitBuilder(), pulse(oc))
    circuits = [c for c, _ in mine] + [chain]
    ics = [tc.ring_ic(c, nd) for c, nd in mine] + [chain.zeros()]
    refs = [oa.transient(o, ts, ic=tc.ring_ic(o, nd), method=method, options=oa.Options(**TIGHT))
            for o, nd in theirs]
    refs.append(oa.transient(ochain, ts, ic=ochain.zeros(), method=method,
                             options=oa.Options(**TIGHT)))
    for tol in (TIGHT, {}):
        plain = vx.transient_batch(circuits, ts, ic=ics, method=method,
                                   options=vx.Options(predict=predict, **tol))
        opts = vx.Options(predict=predict, reuse=True, **tol)
        z, conv, iters, batch = vx.transient_batch(circuits, ts, ic=ics, method=method,
                                                   options=opts, raw=True)
        assert bool(conv.all())
        zs = batch.split(z, len(ts))
        facts = batch.last_factorizations.cpu().numpy()[:, 1:]
        its = iters.cpu().numpy()[:, 1:]
        assert np.all(facts >= 1) and np.all(facts <= its)
        if not tol:  # (at rtol 1e-12 round-off stalls the contraction test)
            assert facts.sum() < 0.8 * its.sum()
        for got, a, ref in zip(zs, plain, refs):
            if tol:
                np.testing.assert_allclose(got, ref.z, rtol=1e-9, atol=1e-12)
                np.testing.assert_allclose(got, a.z, rtol=1e-10, atol=1e-13)
            else:
                assert np.abs(got - a.z).max() <= 1e-9 * 1.2
    with pytest.raises(ValueError, match="precision"):
        vx.Options(reuse=True, precision="mixed")


def test_dc_sweep_follows_the_oracle_point_by_point():
    """Transfer curve of the 3-stage chain, Vin swept 0 -> 1.2 V with
    continuation (voltax/analysis.py:532-572): every point equals the oracle's
    DC solve started from the oracle's previous point."""
    import voltax_amd as vx

    c = tc.inverter_chain(vx.CircuitBuilder(), 0.0)
    o = tc.inverter_chain(oc.CircuitBuilder(), 0.0)
    # both sides of the switching threshold; the threshold itself (gain ~1e7 for
    # three stages) is left out: there a 1e-9 comparison measures round-off
    vals = np.concatenate([np.linspace(0.0, 0.5, 6), np.linspace(0.7, 1.2, 6)])
    tight = dict(rtol=1e-12, atol=1e-15)
    sol = vx.dc_sweep(c, vals, "Vin", options=vx.Options(**tight))
    assert sol.converged.all() and sol.z.shape == (12, c.size)
    np.testing.assert_array_equal(sol.t, vals)
    z = o.zeros()
    for n, v in enumerate(vals):
        o.set_param("Vin", value=float(v))
        ref = oa.dc(o, guess=z, options=oa.Options(**tight))
        z = ref.z
        np.testing.assert_allclose(sol.z[n], ref.z, rtol=1e-9, atol=1e-12)
    # the swept copy leaves the caller's circuit untouched
    assert c.get("Vin", "value") == 0.0
    out = sol.z[:, c.node("out")]
    assert out[0] > 1.19 and out[-1] < 0.01  # three inversions: out = not in



def _ac_circuits(builder):
    b = builder()
    b.vsource("in", "0", 0.0, ac=1.0, name="Vin")
    b.resistor("in", "out", 1e3)
    b.capacitor("out", "0", 1e-12)
    rc = b.build()
    chain = tc.inverter_chain(builder(), 0.5, ac=1.0)
    return rc, chain


def test_ac_matches_oracle_and_closed_form():
    """`ac` (voltax/analysis.py:494-524): RC low-pass = 1 / (1 + j w RC); the
    3-stage EKV chain linearised at its DC point (Vin = 0.5 V, ac 1 V) equals
    the oracle's dense complex solve at every frequency."""
    import voltax_amd as vx

    rc, chain = _ac_circuits(vx.CircuitBuilder)
    orc, ochain = _ac_circuits(oc.CircuitBuilder)
    freqs = np.logspace(3, 11, 17)
    sol = vx.ac(rc, freqs, options=vx.Options(**TIGHT))
    h = 1.0 / (1.0 + 2j * np.pi * freqs * 1e3 * 1e-12)
    np.testing.assert_allclose(sol.v("out"), h, rtol=1e-12, atol=1e-15)
    np.testing.assert_allclose(sol.z, oa.ac(orc, freqs, options=oa.Options(**TIGHT)),
                               rtol=1e-10, atol=1e-15)
    sol = vx.ac(chain, freqs, options=vx.Options(**TIGHT))
    ref = oa.ac(ochain, freqs, options=oa.Options(**TIGHT))
    np.testing.assert_allclose(sol.z, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())
    assert sol.db("out").shape == freqs.shape



def _waveform_circuit(b):
    """Every waveform kind (PWL with delay and repeat, a damped and phased
    Sine, Pulse, Step) driving RC sections and an EKV inverter."""
    mod = oc if isinstance(b, oc.CircuitBuilder) else _vx()
    b.vsource("a", "0", mod.PWL([0.0, 0.3e-9, 0.5e-9, 0.9e-9], [0.0, 1.0, 0.4, 1.1],
                                delay=0.1e-9, repeat=0.2e-9), name="Vpwl")
    b.resistor("a", "x", 1e3, name="Ra")
    b.capacitor("x", "0", 50e-15, name="Cx")
    b.isource("0", "x", mod.Sine(0.0, 0.2e-3, 3e9, 0.05e-9, 1e9, 30.0), name="Isin")
    b.vsource("p", "0", mod.Pulse(0.0, 1.2, 0.1e-9, 30e-12, 50e-12, 0.3e-9, 0.7e-9), name="Vp")
    b.vsource("s", "0", mod.Step(0.2, 0.9, 0.6e-9), name="Vs")
    b.resistor("s", "y", 2e3, name="Rs")
    b.capacitor("y", "0", 20e-15, name="Cy")
    b.vsource("vdd", "0", 1.2, name="Vdd")
    b.pmos("q", "p", "vdd", "vdd", w=2e-6, l=0.13e-6, name="Mp")
    b.nmos("q", "p", "0", "0", w=1e-6, l=0.13e-6, name="Mn")
    b.capacitor("q", "0", 5e-15, name="Cq")
    return b.build()


@pytest.mark.parametrize("solver", ["dense", "sparse"])
@pytest.mark.parametrize("method", ["be", "trap"])
def test_waveforms_drive_both_paths_like_the_oracle(solver, method):
    """PWL / Sine / Pulse / Step sources evaluated on the GPU at every time
    step (fused dense kernel, sparse path) against the oracle; the driven
    nodes equal the waveforms' closed forms."""
    vx = _vx()
    c, o = both(_waveform_circuit)
    ts = np.linspace(0, 1.5e-9, 301)
    got = vx.transient(c, ts, ic="dc", method=method, options=vx.Options(solver=solver, **TIGHT))
    ref = oa.transient(o, ts, ic="dc", method=method, options=oa.Options(**TIGHT))
    assert got.converged.all() and ref.converged.all()
    np.testing.assert_allclose(got.z, ref.z, rtol=0, atol=1e-9 * 1.2)
    # PWL: delay 0.1 ns, then repeats [0.2, 0.9] ns -> at 1.25 ns: tau = 1.15,
    # folded to 0.45 ns, between (0.3, 1.0) and (0.5, 0.4): 1.0 - 0.6 * 0.75
    k = int(round(1.25e-9 / ts[1]))
    assert got.v("a")[k] == pytest.approx(0.55, abs=1e-12)
    assert got.v("s")[0] == 0.2 and got.v("s")[-1] == pytest.approx(0.9, abs=1e-15)
    assert got.v("p")[k] == pytest.approx(0.0, abs=1e-15)   # 1.25 ns: after the fall


@pytest.mark.parametrize("solver", ["dense", "sparse"])
def test_gmin_stepping_dc_of_the_full_adder_matches_oracle(solver):
    """gmin stepping on the GPU (fused kernel and sparse path) on the
    reference's fixture, a CMOS full adder whose direct Newton attempt fails
    (tests/test_analysis.py:138-149): like the oracle, the direct attempt
    alone does not converge, the default options converge through the
    stepping stages (iteration counts beyond the 25 direct steps), and the
    operating points equal the oracle's to 1e-9 V for every input pair."""
    vx = _vx()
    for a, b in [(0.0, 0.0), (1.2, 0.0), (1.2, 1.2)]:
        c = tc.full_adder_dc(vx.CircuitBuilder(), a, b)
        o = tc.full_adder_dc(oc.CircuitBuilder(), a, b)
        plain = vx.dc(c, options=vx.Options(solver=solver, gmin_steps=0))
        assert not bool(plain.converged)
        got = vx.dc(c, options=vx.Options(solver=solver, **TIGHT))
        ref = oa.dc(o, options=oa.Options(**TIGHT))
        assert bool(got.converged) and bool(ref.converged)
        assert int(got.iters) > 25 and int(ref.iters) > 25
        assert abs(int(got.iters) - int(ref.iters)) <= 2, (int(got.iters), int(ref.iters))
        np.testing.assert_allclose(got.z[: c.n_nodes], ref.z[: o.n_nodes], rtol=0, atol=1.2e-9)
