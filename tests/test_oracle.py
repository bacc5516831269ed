"""Pins the oracle (CPU restatement of the reference) before it is trusted.

Each check is one the reference's own tests hold or one the mathematics
fixes, chosen so a dropped term, wrong sign or transposed operand fails:

* golden RC-ladder state and gradient      tests/test_analysis.py:33-49
* transient / DC gradients vs central FD   tests/test_analysis.py:52-81
* EKV closed forms                         tests/test_elements.py:190-219
* stamped Jacobian = derivative of f, q    tests/test_sparse.py:174-189 (vs FD)
* supernodal factor/solve = LAPACK         tests/test_sparse.py:208-230
* plan invariants                          tests/test_sparse.py:233-242
* RC exponential, BE/trap orders           tests/test_elements.py:29-52
* config-1 inverter-chain delay 43.5 ps    README.md:78
* PWL td= / r= hand values                 voltax/signals.py:150-165 (SPICE definition)
"""

import numpy as np
import pytest

import circuits as tc
from oracle import analysis as oa
from oracle import circuit as oc
from oracle import measure as om
from oracle import sparse as osp


def fd_grad(f, x, eps=1e-6):
    x = np.asarray(x, float)
    g = np.zeros_like(x)
    for i in range(x.size):
        dx = np.zeros_like(x)
        dx[i] = eps
        g[i] = (f(x + dx) - f(x - dx)) / (2 * eps)
    return g


def rc(R1=1e3, R2=2e3, C1=1e-9, C2=0.5e-9, Rload=1e4):
    return tc.rc_ladder(oc.CircuitBuilder(), R1, R2, C1, C2, Rload)


def test_golden_rc_ladder_state_and_gradient():
    c = rc()
    ts = np.linspace(0, 1e-5, 101)
    sol = oa.transient(c, ts, ic="zero")
    np.testing.assert_allclose(sol.z[-1], [1.0, 0.921635241, 0.766714055, -7.83647588e-05],
                               atol=1e-8)
    g_z = np.zeros_like(sol.z)
    g_z[:, c.node("out")] = 1.0
    g = oa.transient_vjp(c, sol, g_z, ic="zero")["Resistor"]["log_r"]
    # group order R1, R2, Rload (builder insertion order)
    np.testing.assert_allclose(g, -np.array([13.35247162, 13.96802873, -11.62111701]),
                               rtol=1e-6)


@pytest.mark.parametrize("method", ["be", "trap"])
def test_transient_gradient_vs_fd(method):
    ts = np.linspace(0, 5e-6, 51)

    def run(lp):
        R1, R2, C1, C2 = np.exp(lp)
        c = rc(R1, R2, C1, C2)
        return c, oa.transient(c, ts, ic="zero", method=method)

    x = np.log([1e3, 2e3, 1e-9, 0.5e-9])
    c, sol = run(x)
    out = c.node("out")
    g_z = np.zeros_like(sol.z)
    g_z[:, out] = 2 * sol.z[:, out] / len(ts)
    gr = oa.transient_vjp(c, sol, g_z, ic="zero", method=method)
    g = np.concatenate([gr["Resistor"]["log_r"][:2], gr["Capacitor"]["log_c"]])
    fd = fd_grad(lambda p: float(np.mean(run(p)[1].z[:, out] ** 2)), x)
    np.testing.assert_allclose(g, fd, rtol=1e-5, atol=1e-10)


def test_dc_gradient_through_ekv_vs_fd():
    def build(w, vth):
        b = oc.CircuitBuilder(nmos=oc.MOSProcess.nmos(vth=vth))
        b.vsource("vdd", "0", 1.2)
        b.vsource("in", "0", 0.55)
        b.resistor("vdd", "out", 20e3)
        b.nmos("out", "in", "0", "0", w=w)
        return b.build()

    c = build(1.5e-6, 0.42)
    op = oa.dc(c, options=oa.Options(rtol=1e-12, atol=1e-15))
    g_z = np.zeros(c.size)
    g_z[c.node("out")] = 1.0
    g = oa.dc_vjp(c, op.z, 0.0, g_z)["EKVMOSFET_n"]["log_w"][0]
    fd = fd_grad(lambda lw: float(oa.dc(build(np.exp(lw[0]), 0.42),
                                        options=oa.Options(rtol=1e-12, atol=1e-15))
                                  .v("out")), np.log([1.5e-6]), 1e-6)[0]
    assert np.isclose(g, fd, rtol=1e-6)


def test_dc_gradient_of_grid_vs_fd_including_sources():
    """Resistor, current-source and voltage-source cotangents of a DC node
    voltage vs central differences (tests/test_sparse.py:336-353)."""
    c = tc.grid(oc.CircuitBuilder(), 5)
    node = c.node("n2_2")
    op = oa.dc(c, options=oa.Options(rtol=1e-12, atol=1e-15))
    g_z = np.zeros(c.size)
    g_z[node] = 1.0
    gr = oa.dc_vjp(c, op.z, 0.0, g_z)

    def v(group, key, i, eps):
        cc = tc.grid(oc.CircuitBuilder(), 5)
        arr = cc.groups[group].params[key] if key != "value" else cc.groups[group].wave["value"]
        arr[i] += eps
        return oa.dc(cc, options=oa.Options(rtol=1e-12, atol=1e-15)).z[node]

    for group, key, i, eps in [("Resistor", "log_r", 7, 1e-3), ("CurrentSource", "value", 2, 1e-4),
                               ("VoltageSource", "value", 1, 1e-3)]:
        fd = (v(group, key, i, eps) - v(group, key, i, -eps)) / (2 * eps)
        assert np.isclose(gr[group][key][i], fd, rtol=1e-5, atol=1e-14), (group, key)


def _ekv(w=1e-6, l=0.13e-6, polarity="n", process=None):
    kp = 400e-6 if polarity == "n" else 200e-6
    return oc.Group("EKVMOSFET", np.zeros((1, 4), np.int64),
                    {"log_w": np.log([w]), "log_l": np.log([l]), "dvth": np.zeros(1)},
                    polarity=polarity, process=process or oc.MOSProcess(kp=kp))


def _ids(g, vd, vg, vs, vb):
    I, _ = oc.currents(g, [np.array([x], float) for x in (vd, vg, vs, vb)], [], 0.0, g.params)
    return I[0][0]


def test_ekv_square_law_symmetry_and_mirror():
    g = _ekv(2e-6, 1e-6, process=oc.MOSProcess(kp=200e-6, vth=0.4, n=1.0, lam=0.0))
    sq = 0.5 * 200e-6 * 2 * (1.4 - 0.4) ** 2
    assert np.isclose(_ids(g, 1.5, 1.4, 0.0, 0.0), sq, rtol=0.05)
    n = _ekv()
    assert np.isclose(_ids(n, 0.3, 0.9, 0.1, 0.0), -_ids(n, 0.1, 0.9, 0.3, 0.0))
    p_same = _ekv(polarity="p", process=n.process)
    assert np.isclose(_ids(p_same, -0.3, -0.9, -0.1, 0.0), -_ids(n, 0.3, 0.9, 0.1, 0.0))
    assert _ids(_ekv(polarity="p"), 0.0, 0.0, 1.2, 1.2) < 0


def test_ekv_charge_conservation_and_cgg():
    g = _ekv(10e-6, 1e-6)
    v = [np.array([x]) for x in (1.2, 1.2, 0.0, 0.0)]
    Q, _ = oc.charges(g, v, [], g.params)
    assert abs(sum(q[0] for q in Q)) < 1e-25
    h = 1e-6
    qg = lambda vg: oc.charges(g, [v[0], np.array([vg]), v[2], v[3]], [], g.params)[0][1][0]
    cgg = (qg(1.2 + h) - qg(1.2 - h)) / (2 * h)
    p = g.process
    expected = (2 / 3 + (p.n - 1) / (3 * p.n)) * p.cox * 10e-12 + 2 * p.cov * 10e-6
    assert np.isclose(cgg, expected, rtol=0.01)


def _zoo():
    b = oc.CircuitBuilder
    ring, _ = tc.ring(b(), 5, dvth=np.linspace(-0.02, 0.02, 10))
    return {
        "rc_ladder": rc(),
        "chain": tc.inverter_chain(b(), oc.Pulse(0, 1.2, 0.2e-9, 20e-12, 20e-12, 2e-9, 4e-9)),
        "ring": ring,
        "grid": tc.grid(b(), 6),
    }


@pytest.mark.parametrize("name", ["rc_ladder", "chain", "ring", "grid"])
def test_stamped_jacobian_is_derivative_of_residual(name):
    c = _zoo()[name]
    st = osp.structure(c)
    z = np.random.default_rng(0).uniform(-0.5, 1.5, c.size)
    t = 0.3e-9
    for part, fn in (("f", lambda zz: c.f(zz, t)), ("q", c.q)):
        J = st.to_dense(osp.jacobian(c, z, t, 1.0 if part == "f" else 0,
                                     1.0 if part == "q" else 0, st))
        h = 1e-6
        fd = np.stack([(fn(z + h * e) - fn(z - h * e)) / (2 * h) for e in np.eye(c.size)], 1)
        scale = np.abs(fd).max()
        np.testing.assert_allclose(J, fd, rtol=1e-5, atol=1e-7 * scale)


def test_pattern_covers_device_blocks_only():
    c = rc()
    st = osp.structure(c)
    P = st.to_dense(np.ones(st.nnz)) != 0
    a, out, vin = c.node("a"), c.node("out"), c.layout.internal("Vin")
    assert P[a, out] and P[out, a] and P[vin, c.node("in")]
    assert not P[c.node("in"), out]
    assert np.all(np.diag(P))


@pytest.mark.parametrize("name,leaf", [("grid", 1), ("grid", 16), ("chain", 1), ("ring", 2)])
def test_factor_solve_matches_lapack(name, leaf):
    c = _zoo()[name]
    st = osp.structure(c)
    z = oa.dc(c).z
    plan = osp.make_plan(st, leaf)
    G = osp.jacobian(c, z, 0.0, 1.0, 0, st)
    G[st.diagonal[: c.n_nodes]] += 1e-12
    C = osp.jacobian(c, z, 0.0, 0, 1.0, st)
    b = np.random.default_rng(1).standard_normal(c.size)
    for data in (G, G + 2j * np.pi * 1e9 * C):
        A = st.to_dense(data)
        fac = osp.factor(plan, data)
        for tr in (False, True):
            x = osp.solve(plan, fac, b.astype(data.dtype), tr)
            ref = np.linalg.solve(A.T if tr else A, b)
            np.testing.assert_allclose(x, ref, rtol=1e-8, atol=1e-10 * np.abs(ref).max())


def test_plan_invariants():
    c = tc.grid(oc.CircuitBuilder(), 12)
    st = osp.structure(c)
    gid = osp.pivot_groups(st)
    for name in c.groups:
        if c.groups[name].kind == "VoltageSource":
            u = c.unknowns(name)
            for dev in u:
                assert gid[dev[2]] in (gid[dev[0]], gid[dev[1]] if dev[1] >= 0 else -1)
    plan = osp.make_plan(st)
    assert plan.stats["supernodes"] > 1 and plan.stats["levels"] > 2
    covered = np.sort(np.concatenate(plan.supernodes))
    assert np.array_equal(covered, np.arange(st.size))


def test_spsolve_vjp_matches_dense_adjoint():
    c = tc.grid(oc.CircuitBuilder(), 6)
    st = osp.structure(c)
    plan = osp.make_plan(st, 4)
    data = osp.jacobian(c, np.zeros(c.size), 0.0, 1.0, 0, st)
    data[st.diagonal[: c.n_nodes]] += 1e-3
    rng = np.random.default_rng(2)
    b, g = rng.standard_normal(c.size), rng.standard_normal(c.size)
    gdata, gb = osp.spsolve_vjp(st, plan, data, b, g)
    A = st.to_dense(data)
    x = np.linalg.solve(A, b)
    lam = np.linalg.solve(A.T, g)
    np.testing.assert_allclose(gb, lam, rtol=1e-9)
    np.testing.assert_allclose(gdata, -lam[st.rows] * x[st.indices], rtol=1e-8, atol=1e-14)


def test_rc_step_exponential_and_integration_orders():
    R, C = 1e3, 1e-9

    def circ():
        b = oc.CircuitBuilder()
        b.vsource("in", "0", 1.0)
        b.resistor("in", "out", R)
        b.capacitor("out", "0", C)
        return b.build()

    c = circ()
    ts = np.linspace(0, 5 * R * C, 2001)
    sol = oa.transient(c, ts, ic=c.state(v={"in": 1.0}), method="trap")
    assert np.max(np.abs(sol.v("out")[1:] - (1 - np.exp(-ts[1:] / (R * C))))) < 1e-5
    for method, order in (("be", 1), ("trap", 2)):
        errs = []
        for n in (50, 100):
            t2 = np.linspace(0, R * C, n + 1)
            v = oa.transient(c, t2, ic=c.state(v={"in": 1.0}), method=method).v("out")[-1]
            errs.append(abs(v - (1 - np.exp(-1))))
        assert np.log2(errs[0] / errs[1]) == pytest.approx(order, abs=0.15)


def test_resistor_divider_dc_closed_form():
    b = oc.CircuitBuilder()
    b.vsource("in", "0", 3.0, name="V")
    b.resistor("in", "mid", 1e3)
    b.resistor("mid", "0", 2e3)
    op = oa.dc(b.build())
    assert op.converged and np.isclose(op.v("mid"), 2.0) and np.isclose(op.i("V"), -1e-3)


def test_config1_delay_is_43_5_ps():
    """README.md:60-78: the 3-stage EKV inverter chain, 43.5 ps."""
    c = tc.inverter_chain(oc.CircuitBuilder(), oc.Pulse(0, 1.2, 0.2e-9, 20e-12, 20e-12,
                                                        2e-9, 4e-9))
    ts = np.linspace(0, 1e-9, 1001)
    sol = oa.transient(c, ts)
    assert sol.converged.all()
    d = om.delay(ts, sol.v("in"), sol.v("out"), 0.6, targ_direction="fall")
    assert round(d * 1e12, 1) == 43.5


def test_measure_period_of_sampled_sine():
    t = np.linspace(0, 10e-9, 20001)
    y = np.sin(2 * np.pi * 1e9 * t + 0.3)
    assert np.isclose(om.period(t, y), 1e-9, rtol=1e-6)


# ---------------------------------------------------------------- MLP oracle
def _mlp_case(seed=0, dims=(7, 12, 10, 5), M=9):
    rng = np.random.default_rng(seed)
    Ws = [rng.normal(0, 0.6, (dims[i + 1], dims[i])) for i in range(len(dims) - 1)]
    bs = [rng.normal(0, 0.3, dims[i + 1]) for i in range(len(dims) - 1)]
    X = rng.normal(0, 1, (M, dims[0]))
    T = rng.normal(0, 1, (M, dims[-1]))
    return Ws, bs, X, T


def test_mlp_oracle_gradients_match_central_differences():
    """The oracle's backward pass (no bf16 rounding) against central finite
    differences of its own loss in FP64, for every weight and bias entry of a
    small 3-layer network and of its linear skip."""
    from oracle import mlp

    Ws, bs, X, T = _mlp_case()
    S = np.random.default_rng(9).normal(0, 0.3, (T.shape[1], X.shape[1]))
    _, dWs, dbs, dS = mlp.backward(Ws, bs, X, T, quant=False, S=S)
    h = 1e-6
    pairs = [(S, dS)] + [(Ws[l], dWs[l]) for l in range(len(Ws))] + \
        [(bs[l], dbs[l]) for l in range(len(Ws))]
    for P, dP in pairs:
        fd = np.zeros_like(P)
        for idx in np.ndindex(P.shape):
            old = P[idx]
            P[idx] = old + h
            lp = mlp.loss(Ws, bs, X, T, S=S)
            P[idx] = old - h
            lm = mlp.loss(Ws, bs, X, T, S=S)
            P[idx] = old
            fd[idx] = (lp - lm) / (2 * h)
        np.testing.assert_allclose(dP, fd, rtol=1e-5, atol=1e-8)


def test_mlp_oracle_bf16_rounding_matches_torch():
    """bf16 rounding of the oracle = torch's float32 -> bfloat16 conversion
    (round to nearest even), including ties and values near powers of two."""
    import torch

    from oracle import mlp

    rng = np.random.default_rng(1)
    x = np.concatenate([rng.normal(0, 1, 4000), rng.normal(0, 1e-20, 100), rng.normal(0, 1e20, 100),
                        np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -1.0 - 2 ** -8, 0.0])])
    x32 = x.astype(np.float32)
    want = torch.from_numpy(x32).to(torch.bfloat16).to(torch.float64).numpy()
    np.testing.assert_array_equal(mlp.bf16(x32), want)


def test_mlp_oracle_adam_first_step_is_signed_lr():
    """A bias-corrected Adam step from zero moments moves every weight by
    -lr * sign(g) (up to eps) whatever the gradient scale (Kingma & Ba 2015, sec. 3)."""
    from oracle import mlp

    g = np.array([3.0, -1e-3, 2e2])
    w, m1, m2 = mlp.adam(np.zeros(3), g, np.zeros(3), np.zeros(3), 1, lr=0.01)
    np.testing.assert_allclose(w, -0.01 * np.sign(g), rtol=1e-4)


def test_gmin_stepping_rescues_cmos_dc():
    """The reference's fixture (tests/test_analysis.py:138-149): the direct
    Newton attempt on a CMOS full adder fails, gmin stepping converges to the
    logic levels.  Every input pair also gives the full adder's truth table
    (s = a xor c, co = a and c with carry-in 0), which a wrong gmin schedule,
    a missing restart from the guess or a stage that kept the large gmin
    would not."""
    c = tc.full_adder_dc(oc.CircuitBuilder())
    plain = oa.dc(c, options=oa.Options(gmin_steps=0))
    rescued = oa.dc(c)
    assert not plain.converged
    assert rescued.converged
    assert rescued.v("s") < 0.01 and rescued.v("co") < 0.01
    assert int(rescued.iters) > oa.Options().direct_steps  # the stepping stages ran
    for a, b in [(1.2, 0.0), (0.0, 1.2), (1.2, 1.2)]:
        sol = oa.dc(tc.full_adder_dc(oc.CircuitBuilder(), a, b))
        assert sol.converged
        s, co = (a > 0) != (b > 0), (a > 0) and (b > 0)
        assert abs(sol.v("s") - 1.2 * s) < 0.01 and abs(sol.v("co") - 1.2 * co) < 0.01
    # the schedule: stage k solves at logspace(gmin_start, gmin, steps)[k-1],
    # restarting from the guess at stage 1; with one stage logspace holds
    # gmin_start alone, so the result is plain Newton at gmin_start from the
    # guess with the full budget
    one = oa.dc(c, options=oa.Options(gmin_steps=1))
    direct = oa.newton(oa.Problem(c, "dc", 0.0), oa.Linear(c, oa.Options()), c.zeros(),
                       1e-2, oa.Options(), 100)
    np.testing.assert_array_equal(one.z, direct[0])
    assert bool(one.converged) == direct[1]


@pytest.mark.parametrize("which", ["ring5", "ring13", "chain"])
def test_compiled_cpu_restatement_equals_the_numpy_oracle(which):
    """The compiled single-thread restatement of the dense path (the timed CPU
    baseline, oracle/cpu_dense.cpp) gives the numpy oracle's transient: the
    same states to 1e-10 V and the same Newton iteration counts."""
    import voltax_amd as vx
    from oracle import cpu_dense

    if which.startswith("ring"):
        n = int(which[4:])
        c, nodes = tc.ring(vx.CircuitBuilder(), n)
        o, _ = tc.ring(oc.CircuitBuilder(), n)
        z0 = tc.ring_ic(o, nodes)
        ts = np.arange(101) * 2e-12
    else:
        pulse = lambda m: m.Pulse(0, 1.2, 0.05e-9, 20e-12, 20e-12, 2e-9, 4e-9)  # noqa: E731
        c = tc.inverter_chain(vx.CircuitBuilder(), pulse(vx))
        o = tc.inverter_chain(oc.CircuitBuilder(), pulse(oc))
        z0 = oa.dc(o).z
        ts = np.linspace(0, 0.2e-9, 101)
    ref = oa.transient(o, ts, ic=z0)
    z, iters, conv = cpu_dense.transient_be(c.size, c.n_nodes, c.device_table(), z0, ts)
    assert conv.all() and ref.converged.all()
    np.testing.assert_allclose(z, ref.z, rtol=0, atol=1e-10)
    np.testing.assert_array_equal(iters[1:], ref.iters[1:])


def _pwl(times, values, delay=0.0, repeat=-1.0):
    p = {"times": np.asarray(times, float), "values": np.asarray(values, float),
         "delay": delay, "repeat": repeat}
    return lambda t: float(oc.PWL.eval(p, t)[0])


def test_pwl_repeat_follows_the_spice_definition():
    """SPICE's PWL(... td= r=) as the reference documents it
    (voltax/signals.py:150-165): the first value before the first point,
    linear between points, the last value after them without r=; with r= the
    segment [r, t_last] repeats once past t_last, a period boundary keeping
    the left limit (the value at t_last), td shifting everything.  Hand
    values of times (0, 1, 2, 3) ns, values (0, 1, 0, 2), r = 1 ns (period
    2 ns), plus periodicity and r = t_0 (the whole waveform repeats)."""
    ns = 1e-9
    times, values = np.array([0.0, 1.0, 2.0, 3.0]) * ns, [0.0, 1.0, 0.0, 2.0]
    plain = _pwl(times, values)
    rep = _pwl(times, values, repeat=1 * ns)
    for t, v in [(-1.0, 0.0), (0.5, 0.5), (2.5, 1.0), (3.0, 2.0), (7.0, 2.0)]:
        assert plain(t * ns) == pytest.approx(v, abs=1e-15)
    # before t_last the repeat changes nothing
    for t in np.linspace(-0.5, 3.0, 15) * ns:
        assert rep(t) == pytest.approx(plain(t), abs=1e-15)
    # past t_last: tau folded back by whole periods into (1, 3] ns
    for t, v in [(3.5, 0.5), (4.0, 0.0), (4.5, 1.0), (5.0, 2.0), (5.5, 0.5), (6.0, 0.0),
                 (7.0, 2.0), (9.0, 2.0), (8.25, 0.5)]:
        assert rep(t * ns) == pytest.approx(v, abs=1e-12), t
    # periodic with period t_last - r past t_last
    for t in np.linspace(3.05, 4.95, 20) * ns:
        assert rep(t + 2 * ns) == pytest.approx(rep(t), abs=1e-12)
        assert rep(t + 6 * ns) == pytest.approx(rep(t), abs=1e-12)
    # td shifts the whole repeating waveform
    shifted = _pwl(times, values, delay=0.7 * ns, repeat=1 * ns)
    for t in np.linspace(-1.0, 9.0, 41) * ns:
        assert shifted(t + 0.7 * ns) == pytest.approx(rep(t), abs=1e-12)
    # r = t_0: the whole waveform repeats (period t_last - t_0 = 3 ns)
    whole = _pwl(times, values, repeat=0.0)
    for t, v in [(3.5, 0.5), (4.0, 1.0), (5.5, 1.0), (6.0, 2.0), (6.5, 0.5)]:
        assert whole(t * ns) == pytest.approx(v, abs=1e-12), t
