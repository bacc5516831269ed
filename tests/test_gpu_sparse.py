"""GPU parity of the supernodal LU (factor, solve, transposed solve, mixed
precision) and of the large-system Newton driver, against LAPACK and the
oracle (mirrors tests/test_sparse.py:208-303 of the reference)."""

import numpy as np
import pytest

import circuits as tc
from oracle import analysis as oa
from oracle import circuit as oc
from oracle import sparse as osp

pytestmark = pytest.mark.gpu

TIGHT = dict(rtol=1e-12, atol=1e-15)


def _sys(name, leaf):
    import voltax_amd as vx
    from voltax_amd import sparse, splu

    if name == "grid":
        c = tc.grid(vx.CircuitBuilder(), 10)
        o = tc.grid(oc.CircuitBuilder(), 10)
    elif name.startswith("grid"):
        n = int(name[4:])
        c = tc.grid(vx.CircuitBuilder(), n)
        o = tc.grid(oc.CircuitBuilder(), n)
    elif name.startswith("ladder"):
        n = int(name[6:])
        c = tc.ladder(vx.CircuitBuilder(), n)
        o = tc.ladder(oc.CircuitBuilder(), n)
    elif name == "chain":
        c = tc.inverter_chain(vx.CircuitBuilder(), vx.Pulse(0, 1.2, 1e-10, 2e-11, 2e-11, 2e-9, 4e-9))
        o = tc.inverter_chain(oc.CircuitBuilder(), oc.Pulse(0, 1.2, 1e-10, 2e-11, 2e-11, 2e-9, 4e-9))
    else:
        c, o = tc.ring(vx.CircuitBuilder(), 9)[0], tc.ring(oc.CircuitBuilder(), 9)[0]
    z = np.random.default_rng(3).uniform(0, 1.2, c.size)
    ost = osp.structure(o)
    data = osp.jacobian(o, z, 0.0, 1.0, 1e12, ost)  # BE-like matrix, oracle values
    data[ost.diagonal[: o.n_nodes]] += 1e-12
    return c, o, ost, data, splu.plan_of(c, leaf)


@pytest.mark.parametrize("name,leaf", [("grid", 1), ("grid", 16), ("chain", 1),
                                       ("ring", 2), ("ring", 64), ("grid40", 16),
                                       ("ladder100", 256), ("ladder160", 256)])
def test_factor_solve_matches_lapack(name, leaf):
    """Every front size class: register fronts of 16/32/64 unknowns (one
    wavefront, several fronts each; grid40 has all three), a 101-unknown front
    in LDS (ladder100) and a 161-unknown front eliminated in HBM (ladder160)."""
    from voltax_amd import splu

    c, o, ost, data, plan = _sys(name, leaf)
    A = ost.to_dense(data)
    b = np.random.default_rng(1).standard_normal(c.size)
    f = splu.factor(plan, data)
    for tr in (False, True):
        x = splu.solve(plan, f, b, transpose=tr).cpu().numpy()
        ref = np.linalg.solve(A.T if tr else A, b)
        np.testing.assert_allclose(x, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())
    f32 = splu.factor(plan, data, fp32=True)
    x = splu.solve(plan, f32, b, refine=3).cpu().numpy()
    ref = np.linalg.solve(A, b)
    np.testing.assert_allclose(x, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())


def test_batched_factor_solve():
    from voltax_amd import splu

    c, o, ost, data, plan = _sys("grid", 16)
    rng = np.random.default_rng(5)
    B = 7
    datas = np.stack([data * (1 + 0.1 * rng.random(len(data))) for _ in range(B)])
    # keep the matrices structurally the same but numerically distinct and nonsingular
    bs = rng.standard_normal((B, c.size))
    f = splu.factor(plan, datas)
    x = splu.solve(plan, f, bs).cpu().numpy()
    for i in range(B):
        ref = np.linalg.solve(ost.to_dense(datas[i]), bs[i])
        np.testing.assert_allclose(x[i], ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())


def test_spsolve_vjp_matches_oracle():
    from voltax_amd import splu

    c, o, ost, data, plan = _sys("grid", 4)
    rng = np.random.default_rng(2)
    b, g = rng.standard_normal(c.size), rng.standard_normal(c.size)
    gd, gb = splu.spsolve_vjp(plan, data, b, g)
    oplan = osp.make_plan(ost, 4)
    rd, rb = osp.spsolve_vjp(ost, oplan, data, b, g)
    np.testing.assert_allclose(gb.cpu().numpy(), rb, rtol=1e-9, atol=1e-12 * np.abs(rb).max())
    np.testing.assert_allclose(gd.cpu().numpy(), rd, rtol=1e-8, atol=1e-12 * np.abs(rd).max())


def test_plan_invariants():
    import voltax_amd as vx
    from voltax_amd import splu

    c = tc.grid(vx.CircuitBuilder(), 12)
    plan = splu.plan_of(c)
    assert plan.stats["supernodes"] > 1 and plan.stats["levels"] > 2
    sn, par, lev = plan.structure()
    # every voltage-source current is pivoted in the supernode of one of its nodes
    for name, el in c.elements.items():
        if type(el).__name__ == "VoltageSource":
            for dev in c.unknowns(name):
                terms = [t for t in dev[:2] if t >= 0]
                assert sn[dev[2]] in [sn[t] for t in terms]
    # postorder: parents come after children
    assert all(p == -1 or p > i for i, p in enumerate(par))


def test_sparse_dc_and_transient_match_oracle():
    import voltax_amd as vx

    c = tc.grid(vx.CircuitBuilder(), 10)
    o = tc.grid(oc.CircuitBuilder(), 10)
    got = vx.dc(c, options=vx.Options(solver="sparse", **TIGHT))
    ref = oa.dc(o, options=oa.Options(solver="sparse", **TIGHT))
    assert got.converged and ref.converged
    np.testing.assert_allclose(got.z, ref.z, rtol=1e-9, atol=1e-13)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_sparse_transient_of_nonlinear_chain_matches_dense(precision):
    import voltax_amd as vx

    pulse = vx.Pulse(0, 1.2, 0.1e-9, 20e-12, 20e-12, 2e-9, 4e-9)
    c = tc.inverter_chain(vx.CircuitBuilder(), pulse, stages=5)
    ts = np.linspace(0, 0.4e-9, 81)
    sp = vx.transient(c, ts, options=vx.Options(solver="sparse", precision=precision, **TIGHT))
    de = vx.transient(c, ts, options=vx.Options(**TIGHT))
    assert sp.converged.all() and de.converged.all()
    np.testing.assert_allclose(sp.z[:, : c.n_nodes], de.z[:, : c.n_nodes], rtol=0, atol=1.2e-9)


def test_mid_front_kernel_choice_and_autotuner():
    """Both kernels for fronts of 33..64 unknowns (register wavefront, blocked
    LDS) solve the grid40 system to LAPACK accuracy, and the autotuner picks a
    measured candidate and leaves it configured on the plan."""
    from voltax_amd import autotune, splu

    c, o, ost, data, plan = _sys("grid40", 16)
    A = ost.to_dense(data)
    b = np.random.default_rng(4).standard_normal(c.size)
    ref = np.linalg.solve(A, b)
    try:
        for mid in (True, False):
            plan.configure(mid)
            x = splu.solve(plan, splu.factor(plan, data), b).cpu().numpy()
            np.testing.assert_allclose(x, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())
        best, times = autotune.tune_leaf(c, 64, candidates=((8, False), (16, False), (16, True)))
        assert best in times and all(t > 0 for t in times.values())
        assert splu.plan_of(c, best[0]).mid_blocked == best[1]
    finally:
        plan.configure(False)


@pytest.mark.parametrize("name,leaf", [("grid", 16), ("grid40", 16), ("chain", 1), ("ring", 2),
                                       ("ladder100", 256), ("ladder160", 256), ("grid20", 8),
                                       ("grid20", 16), ("ladder80", 4)])
def test_fused_factor_solve_gives_the_same_bits(name, leaf):
    """vxm_splu_factor_solve (the forward sweep of register fronts fused into
    their factorization) equals vxm_splu_factor + vxm_splu_solve bit for bit,
    factors and solution, batched, on every front size class (grid40: all
    register classes; ladders: fronts in LDS and in HBM), and LAPACK.  The
    top levels of small fronts go through the merged top-level kernels on
    grid (10x10, 16-lane fronts from level 1), ring, grid20 (32-lane fronts
    from level 3 with leaf 8, level 2 with leaf 16) and ladder80 (16-lane
    fronts, six levels)."""
    from voltax_amd import splu

    c, o, ost, data, plan = _sys(name, leaf)
    rng = np.random.default_rng(11)
    B = 3
    batch = np.stack([data * (1 + 0.01 * rng.standard_normal(len(data))) for _ in range(B)])
    b = rng.standard_normal((B, c.size))
    f = splu.factor(plan, batch)
    x = splu.solve(plan, f, b).cpu().numpy()
    f2, x2 = splu.factor_solve(plan, batch, b)
    assert np.array_equal(f.front.cpu().numpy(), f2.front.cpu().numpy())
    assert np.array_equal(f.piv.cpu().numpy(), f2.piv.cpu().numpy())
    assert np.array_equal(x, x2.cpu().numpy())
    for i in range(B):
        ref = np.linalg.solve(ost.to_dense(batch[i]), b[i])
        np.testing.assert_allclose(x2.cpu().numpy()[i], ref, rtol=1e-9,
                                   atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("name,leaf", [("grid", 16), ("grid40", 16), ("ladder100", 256)])
def test_complex_factor_solve_matches_numpy(name, leaf):
    """Complex128 fronts (the AC path): G + j w C on the circuit's pattern,
    plain and transposed solves against numpy's complex LAPACK solve."""
    import torch

    from voltax_amd import splu

    c, o, ost, data, plan = _sys(name, leaf)
    rng = np.random.default_rng(8)
    cdata = data * (1 + 0.3j * rng.standard_normal(len(data)))
    A = ost.to_dense(cdata.real) + 1j * ost.to_dense(cdata.imag)
    b = rng.standard_normal(c.size) + 1j * rng.standard_normal(c.size)
    f = splu.factor(plan, torch.from_numpy(cdata))
    for tr in (False, True):
        x = splu.solve(plan, f, b, transpose=tr).cpu().numpy()
        ref = np.linalg.solve(A.T if tr else A, b)
        np.testing.assert_allclose(x, ref, rtol=1e-9, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("method", ["be", "trap"])
def test_sparse_transient_reused_factors_keep_the_roots(method):
    """Options.reuse on the sparse path (later Newton iterations of a step
    reuse its LU factors while the updates contract; a linear circuit keeps
    them across steps of one size): the waveform of the plain path and the
    oracle, at the tight and the default tolerance, with one factorization
    for the whole transient of this linear grid, and for a nonlinear chain
    fewer factorizations than iterations."""
    import voltax_amd as vx
    from voltax_amd.sparse_newton import transient_sparse

    def build(b, mod):
        b.isource("0", "n3_4", mod.Sine(0.0, 2e-3, 5e9), name="Ip")
        b.capacitor("n3_4", "0", 1e-13, name="Cp")
        return tc.grid(b, 8)

    c, o = build(vx.CircuitBuilder(), vx), build(oc.CircuitBuilder(), oc)
    ts = np.linspace(0, 4e-10, 41)
    ref = oa.transient(o, ts, ic="dc", method=method, options=oa.Options(solver="sparse", **TIGHT))
    for tol in (TIGHT, {}):
        plain = transient_sparse(c, ts, "dc", method, vx.Options(**tol))
        mine = transient_sparse(c, ts, "dc", method, vx.Options(reuse=True, **tol))
        assert plain.converged.all() and mine.converged.all()
        assert plain.factorizations is None
        assert mine.factorizations <= mine.iters[1:].sum()
        assert c.is_linear and mine.factorizations == 1
        np.testing.assert_allclose(mine.z, plain.z, rtol=1e-12, atol=1e-15)
        if tol:
            np.testing.assert_allclose(mine.z, ref.z, rtol=1e-9, atol=1e-12)
    # a nonlinear circuit: factors kept within a step only
    pulse = lambda m: m.Pulse(0, 1.2, 20e-12, 20e-12, 20e-12, 60e-12, 200e-12)  # noqa: E731
    c = tc.inverter_chain(vx.CircuitBuilder(), pulse(vx))
    o = tc.inverter_chain(oc.CircuitBuilder(), pulse(oc))
    ts = np.linspace(0, 200e-12, 101)
    ref = oa.transient(o, ts, ic="dc", method=method, options=oa.Options(**TIGHT))
    for tol in (TIGHT, {}):
        plain = transient_sparse(c, ts, "dc", method, vx.Options(**tol))
        mine = transient_sparse(c, ts, "dc", method, vx.Options(reuse=True, **tol))
        assert not c.is_linear and plain.converged.all() and mine.converged.all()
        assert mine.factorizations <= mine.iters[1:].sum()
        if tol:
            np.testing.assert_allclose(mine.z, plain.z, rtol=1e-10, atol=1e-13)
            np.testing.assert_allclose(mine.z, ref.z, rtol=1e-9, atol=1e-12)
        else:
            assert np.abs(mine.z - plain.z).max() <= 1e-9 * 1.2
            assert mine.factorizations < 0.8 * mine.iters[1:].sum()


def test_sparse_transient_extrapolated_start_keeps_the_roots():
    """Options.predict on the sparse path: the same backward-Euler waveform as
    the last-point start (and the oracle); on a smooth drive (a sine current)
    in fewer Newton iterations."""
    import voltax_amd as vx
    from voltax_amd.sparse_newton import transient_sparse

    def build(b, mod):
        b.isource("0", "n3_4", mod.Sine(0.0, 2e-3, 5e9), name="Ip")
        b.capacitor("n3_4", "0", 1e-13, name="Cp")
        return tc.grid(b, 8)

    c, o = build(vx.CircuitBuilder(), vx), build(oc.CircuitBuilder(), oc)
    ts = np.linspace(0, 4e-10, 41)
    plain = transient_sparse(c, ts, "dc", "be", vx.Options(**TIGHT))
    pred = transient_sparse(c, ts, "dc", "be", vx.Options(predict=True, **TIGHT))
    ref = oa.transient(o, ts, ic="dc", options=oa.Options(solver="sparse", **TIGHT))
    assert plain.converged.all() and pred.converged.all()
    np.testing.assert_allclose(pred.z, plain.z, rtol=1e-10, atol=1e-13)
    np.testing.assert_allclose(pred.z, ref.z, rtol=1e-9, atol=1e-12)
    # at the default tolerance a close enough start converges in one iteration
    plain_d = transient_sparse(c, ts, "dc", "be", vx.Options())
    pred_d = transient_sparse(c, ts, "dc", "be", vx.Options(predict=True))
    np.testing.assert_allclose(pred_d.z, plain_d.z, rtol=1e-6, atol=1e-9)
    assert pred_d.iters[1:].sum() < plain_d.iters[1:].sum()
