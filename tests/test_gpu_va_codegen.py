"""Verilog-A programs as generated code on the GPU (run-time HIP C++ from the
compiled program, csrc/va_codegen.cpp): the small test models through the
generated path against the interpreter and the oracle, and a BSIM4-sized
synthetic model (10 unknowns per device, about 600 instructions) in a
circuit large enough for the sparse plan to split, against the oracle in
stamping, DC, transient and adjoint."""

import os

import numpy as np
import pytest

import va_models as vm
from oracle import analysis as oa
from oracle import circuit as oc
from oracle import sparse as osp

pytestmark = pytest.mark.gpu

TIGHT = dict(rtol=1e-12, atol=1e-15)


def _both(build):
    import voltax_amd as vx

    return build(vx.CircuitBuilder()), build(oc.CircuitBuilder())


def test_small_models_through_generated_code_match_the_interpreter():
    """With VXM_VA_CODEGEN=all every program, however small, runs as
    generated code: the stamped residuals and Jacobians of the mixed test
    circuit equal the interpreter's and the oracle's."""
    from test_gpu_va import _mixed
    from voltax_amd import sparse

    c, o = _both(_mixed)
    st = osp.structure(o)
    rng = np.random.default_rng(5)
    zs = [rng.uniform(-0.2, 0.9, c.size) for _ in range(3)]
    coefs = ((1.0, 0.0), (0.0, 1.0), (1.0, 2e9))
    interp = [(sparse.jacobian(c, z, 0.0, cf, cq).cpu().numpy(),
               sparse.residual(c, z, 0.0, cf, cq).cpu().numpy()) for z, (cf, cq) in zip(zs, coefs)]
    os.environ["VXM_VA_CODEGEN"] = "all"
    try:
        print("codegen: mixed circuit, every program generated", flush=True)
        for (z, (cf, cq)), (ji, ri) in zip(zip(zs, coefs), interp):
            jg = sparse.jacobian(c, z, 0.0, cf, cq).cpu().numpy()
            rg = sparse.residual(c, z, 0.0, cf, cq).cpu().numpy()
            ref = osp.jacobian(o, z, 0.0, cf, cq, st)
            np.testing.assert_allclose(jg, ji, rtol=1e-13, atol=1e-16 * np.abs(ji).max())
            np.testing.assert_allclose(rg, ri, rtol=1e-13, atol=1e-16 * np.abs(ri).max())
            np.testing.assert_allclose(jg, ref, rtol=1e-12, atol=1e-15 * np.abs(ref).max())
        import voltax_amd as vx
        from voltax_amd import adjoint

        sol = vx.dc(c, options=vx.Options(**TIGHT))
        ref = oa.dc(o, options=oa.Options(**TIGHT))
        assert sol.converged and ref.converged
        np.testing.assert_allclose(sol.z, ref.z, rtol=1e-9, atol=1e-13)
        g = np.zeros(c.size)
        g[c.node("out")] = 1.0
        gr = adjoint.dc_vjp(c, sol.z, 0.0, g)
        go = oa.dc_vjp(o, ref.z, 0.0, g)
        for grp in ("diode", "sqnmos"):
            for par, v in go[grp].items():
                np.testing.assert_allclose(gr[grp][par], v, rtol=1e-7, atol=1e-12)
    finally:
        del os.environ["VXM_VA_CODEGEN"]


def _bsim_grid(b, n=6, pair=None, pulse=None):
    """An n x n resistor grid fed at two corners, three BSIM-sized devices
    sinking current from grid nodes (gates on their own sources, one of them
    a pulse), and diodes with series resistance (matched pivot pairs)."""
    nm = lambda i, j: f"n{i}_{j}"  # noqa: E731
    rng = np.random.default_rng(2)
    for i in range(n):
        for j in range(n):
            if j + 1 < n:
                b.resistor(nm(i, j), nm(i, j + 1), 0.5 * (1 + rng.random()))
            if i + 1 < n:
                b.resistor(nm(i, j), nm(i + 1, j), 0.5 * (1 + rng.random()))
    b.vsource(nm(0, 0), "0", 1.0, name="Vp0")
    b.vsource(nm(n - 1, n - 1), "0", 1.0, name="Vp1")
    for k, (i, j) in enumerate([(1, 4), (3, 2), (4, 5)]):
        gate = pulse if (k == 0 and pulse is not None) else 0.6 + 0.2 * k
        b.vsource(f"g{k}", "0", gate, name=f"Vg{k}")
        vm.place(b, pair, (nm(i, j), f"g{k}", "0", "0"), name=f"M{k}", w=(1 + k) * 1e-6)
    diode = vm.Pair(vm.DIODE, {"rs": 5.0})
    for i, j in [(2, 2), (5, 0), (0, 5)]:
        vm.place(b, diode, (nm(i, j), "0"))
    return b.build()


def test_bsim_sized_model_on_the_sparse_path_matches_the_oracle():
    import voltax_amd as vx
    from voltax_amd import adjoint, engine, sparse, splu

    pair = vm.Pair(vm.BSIMISH)
    c, o = _both(lambda b: _bsim_grid(b, pair=pair))
    assert c.size == o.size and c.size > 32
    kind, _, unk, _, _ = c.device_table()
    assert unk.shape[1] == 10 and (kind == 3 + 10).sum() == 3
    print(f"codegen: BSIM-sized grid, S={c.size}", flush=True)
    # plan: the nested dissection splits the grid; each device's unknowns are
    # glued to its terminals (voltax/sparse.py:274-300)
    topo, _ = engine.topology_of(c)
    plan = splu.Plan(topo, 16)
    sn, _, _ = plan.structure()
    ref_plan = osp.make_plan(osp.structure(o), 16)
    assert plan.n_supernodes == len(ref_plan.supernodes) > 1
    for i, s in enumerate(ref_plan.supernodes):
        np.testing.assert_array_equal(np.flatnonzero(sn == i), s)
    for k in range(3):
        u = [x for x in unk[np.flatnonzero(kind == 13)[k]] if x >= 0]
        assert len(set(sn[u])) == 1
    # stamping
    st = osp.structure(o)
    rng = np.random.default_rng(7)
    for cf, cq in ((1.0, 0.0), (0.0, 1.0), (1.0, 1e11)):
        z = rng.uniform(0.0, 1.0, c.size)
        got = sparse.jacobian(c, z, 0.0, cf, cq).cpu().numpy()
        ref = osp.jacobian(o, z, 0.0, cf, cq, st)
        np.testing.assert_allclose(got, ref, rtol=1e-12, atol=1e-14 * np.abs(ref).max())
        r = sparse.residual(c, z, 0.0, cf, cq).cpu().numpy()
        rr = cf * o.f(z) + cq * o.q(z)
        np.testing.assert_allclose(r, rr, rtol=1e-12, atol=1e-14 * np.abs(rr).max())
    # DC and its adjoint
    opts = vx.Options(solver="sparse", **TIGHT)
    sol = vx.dc(c, options=opts)
    ref = oa.dc(o, options=oa.Options(**TIGHT))
    assert sol.converged and ref.converged
    np.testing.assert_allclose(sol.z[: c.n_nodes], ref.z[: o.n_nodes], rtol=0, atol=1e-9)
    i_dev = [c.node("n1_4"), c.node("n3_2"), c.node("n4_5")]
    assert min(1.0 - sol.z[i_dev]) > 1e-4  # the devices conduct
    g = np.zeros(c.size)
    g[i_dev] = 1.0
    gr = adjoint.dc_vjp(c, sol.z, 0.0, g)
    go = oa.dc_vjp(o, ref.z, 0.0, g)
    for grp, par in (("bsimish", "log_w"), ("bsimish", "log_l"), ("bsimish", "m"),
                     ("Resistor", "log_r"), ("diode", "area")):
        np.testing.assert_allclose(gr[grp][par], go[grp][par], rtol=1e-7,
                                   atol=1e-9 * np.abs(go[grp][par]).max())
    assert np.abs(go["bsimish"]["log_w"]).min() > 0


def test_bsim_sized_transient_and_gradient_match_the_oracle():
    """Ten BE steps with the first gate pulsing: waveforms against the
    oracle, and the adjoint of the summed grid-node voltages through the
    generated code's J^T and parameter passes."""
    import voltax_amd as vx
    from voltax_amd import adjoint

    pair = vm.Pair(vm.BSIMISH)
    c, o = _both(lambda b: _bsim_grid(b, pair=pair, pulse=(vx if isinstance(b, vx.CircuitBuilder)
                                                        else oc).Pulse(
        0.2, 1.0, 2e-12, 5e-12, 5e-12, 20e-12, 60e-12)))
    ts = np.linspace(0.0, 30e-12, 11)
    opts = vx.Options(solver="sparse", **TIGHT)
    sol = vx.transient(c, ts, options=opts)
    ref = oa.transient(o, ts, options=oa.Options(**TIGHT))
    assert sol.converged.all() and ref.converged.all()
    nv = c.n_nodes
    np.testing.assert_allclose(sol.z[:, :nv], ref.z[:, :nv], rtol=0, atol=1e-9)
    g = np.zeros_like(sol.z)
    g[:, :nv] = 1.0 / nv
    gr = adjoint.transient_vjp(c, sol, g, ic="dc", options=opts)
    go = oa.transient_vjp(o, ref, g, ic="dc", options=oa.Options(**TIGHT))
    for grp, par in (("bsimish", "log_w"), ("bsimish", "log_l"), ("Resistor", "log_r")):
        np.testing.assert_allclose(gr[grp][par], go[grp][par], rtol=1e-7,
                                   atol=1e-9 * np.abs(go[grp][par]).max())


def _bsim_ring_pair(n, dvs, wn=0.2e-6, wp=0.4e-6, l=45e-9, cl=1e-15):
    """The product's ring of the BSIM-sized model (one pattern, per-instance
    parameter rows with vth0 shifted by dv: NMOS +dv, PMOS magnitude -dv) and
    the oracle's ring for every dv."""
    import voltax_amd as vx
    from oracle import va as ova
    from voltax_amd import va, workloads as wl

    m = va.load(wl.BSIM_SIZED_VA)
    N = m.element(polarity="n", differentiable=["vth0"], name="bsim_n")
    P = m.element(polarity="p", differentiable=["vth0"], name="bsim_p")
    c, nodes = wl.bsim_ring(vx.CircuitBuilder(), n, N, P, wn=wn, wp=wp, l=l, cl=cl)
    row = c.device_table()[4]
    cols = wl.vth0_columns(c)
    params = np.tile(row, (len(dvs), 1))
    for b, dv in enumerate(dvs):
        params[b, cols["n"]] += dv
        params[b, cols["p"]] -= dv
    mn = ova.Model(wl.BSIM_SIZED_VA, {"type": 1}, differentiable=["vth0"])
    mp = ova.Model(wl.BSIM_SIZED_VA, {"type": -1}, differentiable=["vth0"])
    oracles = []
    for dv in dvs:
        ob = oc.CircuitBuilder()
        ob.vsource("vdd", "0", 1.0, name="Vdd")
        for i in range(n):
            a, y = f"n{i}", f"n{(i + 1) % n}"
            ob.va(mp, (y, a, "vdd", "vdd"), w=wp, l=l, vth0=0.4 - dv)
            ob.va(mn, (y, a, "0", "0"), w=wn, l=l, vth0=0.4 + dv)
            ob.capacitor(y, "0", cl)
        oracles.append(ob.build())
    return c, nodes, params, oracles


@pytest.mark.parametrize("grouping,extra", [("reference", {}), ("compact", {}),
                                            ("compact", {"predict": True, "reuse": True})])
def test_bsim_sized_ring_batch_and_vth0_gradient_match_the_oracle(grouping, extra):
    """Rings of the BSIM-sized model (the stand-in for the BSIM4 rings of the
    device-evaluation-bound configuration), a batch of per-instance VTH0
    offsets on the sparse path with generated device code: node waveforms
    against the oracle per instance, and the adjoint of a waveform loss with
    respect to every device's vth0 (instance 1) against the oracle, with the
    reference's pivot groups (one front) and the compact ones (small fronts),
    and with predicted starts and reused factors."""
    import voltax_amd as vx
    from voltax_amd import adjoint, workloads as wl
    from voltax_amd.analysis import Solution
    from voltax_amd.sparse_newton import transient_sparse_batch

    dvs = [0.0, 0.02]
    c, nodes, params, oracles = _bsim_ring_pair(3, dvs)
    z0 = wl.bsim_ring_ic(c, nodes)
    ts = np.arange(41) * 1e-12
    opts = vx.Options(solver="sparse", grouping=grouping, **extra, **TIGHT)
    out, conv, _ = transient_sparse_batch(c, params, np.tile(z0, (len(dvs), 1)), ts, options=opts)
    assert bool(conv.all())
    zs = out.cpu().numpy()
    nv = c.n_nodes
    refs = []
    for b, o in enumerate(oracles):
        ic = np.zeros(o.size)
        ic[:nv] = z0[:nv]
        refs.append(oa.transient(o, ts, ic=ic, options=oa.Options(**TIGHT)))
        np.testing.assert_allclose(zs[:, b, :nv], refs[b].z[:, :nv], rtol=1e-9, atol=1e-12)
    assert np.abs(zs[-1, 1, :nv] - zs[-1, 0, :nv]).max() > 1e-4  # the offsets matter
    # adjoint of sum_n w_n v_n0(t_n) with respect to vth0, instance 1
    b = 1
    for name, el in c.elements.items():
        if getattr(el, "va", None) is not None:
            pol = "p" if name.endswith("_p") else "n"
            el.set(slice(None), vth0=0.4 - dvs[b] if pol == "p" else 0.4 + dvs[b])
    assert np.allclose(c.device_table()[4], params[b])
    sol = Solution(ts, zs[:, b, :], np.ones(len(ts), bool), c.layout)
    o, ref = oracles[b], refs[b]
    g = np.zeros((len(ts), c.size))
    g[:, c.node("n0")] = np.linspace(0.0, 1.0, len(ts))
    go_ = np.zeros((len(ts), o.size))
    go_[:, o.node("n0")] = g[:, c.node("n0")]
    gr = adjoint.transient_vjp(c, sol, g, ic="zero", options=opts)
    go = oa.transient_vjp(o, ref, go_, ic="zero", options=oa.Options(**TIGHT))
    for pol, ogrp in (("n", "bsimish_2"), ("p", "bsimish")):
        got = gr[f"bsim_{pol}_{pol}"]["vth0"]
        want = go[ogrp]["vth0"]
        np.testing.assert_allclose(got, want, rtol=1e-7, atol=1e-9 * np.abs(want).max())
