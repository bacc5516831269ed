"""GPU parity of the static-pivot band LU of the fused dense kernel
(Options(lu="band")) against partial pivoting and the oracle.

The band LU pivots on the diagonal of the band numbering where partial
pivoting would pick the same rows (transient Jacobians of the rings are
diagonally dominant), so the two give the same waveforms to round-off; the
test allows 1e-12 V.  Where a static pivot fails its multiplier test the
kernel refactors with partial pivoting: forced here from the host with a
multiplier bound no factorization can meet.
"""

import numpy as np
import pytest

import circuits as tc
from oracle import analysis as oa
from oracle import circuit as oc

pytestmark = pytest.mark.gpu

TIGHT = dict(rtol=1e-12, atol=1e-15)


def _vx():
    import voltax_amd as vx

    return vx


def _rings(mod, ns, seed):
    rng = np.random.default_rng(seed)
    out = []
    for n in ns:
        dv, ws = rng.normal(0, 0.03, 2 * n), np.exp(rng.normal(0, 0.05, 2 * n))
        out.append(tc.ring(mod.CircuitBuilder(), n, dvth=dv, wscale=ws))
    return out


NS = [3, 5, 9, 13, 17, 21, 25, 27, 29, 31]


@pytest.mark.parametrize("method", ["be", "trap"])
def test_band_lu_equals_partial_pivoting_on_the_rings(method):
    """Every ring size of the benchmark (all tile classes): plain Newton,
    the extrapolated start and reused factors, band against partial."""
    vx = _vx()
    from voltax_amd.engine import DenseBatch

    mine = _rings(vx, NS, 1)
    circuits = [c for c, _ in mine]
    ics = [tc.ring_ic(c, nd) for c, nd in mine]
    batch = DenseBatch(circuits)
    assert batch.band_patterns().all()
    ts = np.linspace(0, 300e-12, 151)
    for kw in (dict(), dict(predict=True), dict(predict=True, reuse=True)):
        ref = vx.transient_batch(circuits, ts, ic=ics, method=method,
                                 options=vx.Options(**kw, **TIGHT))
        got = vx.transient_batch(circuits, ts, ic=ics, method=method,
                                 options=vx.Options(lu="band", **kw, **TIGHT))
        for a, b in zip(ref, got):
            assert a.converged.all() and b.converged.all()
            assert np.abs(a.z - b.z).max() <= 1e-12, kw


def test_band_lu_matches_the_oracle():
    vx = _vx()
    ns = [5, 27, 31]
    mine, theirs = _rings(vx, ns, 2), _rings(oc, ns, 2)
    ts = np.linspace(0, 200e-12, 101)
    circuits = [c for c, _ in mine]
    ics = [tc.ring_ic(c, nd) for c, nd in mine]
    got = vx.transient_batch(circuits, ts, ic=ics, options=vx.Options(lu="band", **TIGHT))
    for g, (o, nd) in zip(got, theirs):
        ref = oa.transient(o, ts, ic=tc.ring_ic(o, nd), options=oa.Options(**TIGHT))
        np.testing.assert_allclose(g.z, ref.z, rtol=1e-9, atol=1e-12)


@pytest.mark.parametrize("reuse", [False, True])
def test_failed_static_pivots_fall_back_to_partial_pivoting(reuse):
    """A multiplier bound no factorization meets sends every band
    factorization to the partial-pivoting fallback (in LDS): the waveforms
    still equal partial pivoting.  Reuse runs use the band LU in the 64-lane
    class (the 31-stage ring)."""
    vx = _vx()
    mine = _rings(vx, [7, 31], 3)
    circuits = [c for c, _ in mine]
    ics = [tc.ring_ic(c, nd) for c, nd in mine]
    ts = np.linspace(0, 100e-12, 51)
    kw = dict(predict=True, reuse=True) if reuse else {}
    ref = vx.transient_batch(circuits, ts, ic=ics, options=vx.Options(**kw, **TIGHT))
    got = vx.transient_batch(circuits, ts, ic=ics,
                             options=vx.Options(lu="band", band_max_mult=1e-300, **kw, **TIGHT))
    for a, b in zip(ref, got):
        assert b.converged.all()
        np.testing.assert_allclose(b.z, a.z, rtol=1e-12, atol=1e-13)


def test_mixed_batch_and_dc_with_the_band_lu():
    """A batch mixing band patterns (rings) with one without (the two-source
    inverter chain) in one tile class, and DC operating points (where static
    pivots may fail their test), equal partial pivoting and the oracle."""
    vx = _vx()
    from voltax_amd.engine import DenseBatch

    pulse = lambda m: m.Pulse(0, 1.2, 20e-12, 20e-12, 20e-12, 60e-12, 200e-12)  # noqa: E731
    chain = tc.inverter_chain(vx.CircuitBuilder(), pulse(vx))
    ochain = tc.inverter_chain(oc.CircuitBuilder(), pulse(oc))
    mine = _rings(vx, [3, 5], 4)
    circuits = [chain] + [c for c, _ in mine]
    assert DenseBatch(circuits).band_patterns().tolist() == [0, 1, 1]
    ts = np.linspace(0, 200e-12, 101)
    ics = [chain.zeros()] + [tc.ring_ic(c, nd) for c, nd in mine]
    ref = vx.transient_batch(circuits, ts, ic=ics, options=vx.Options(**TIGHT))
    got = vx.transient_batch(circuits, ts, ic=ics, options=vx.Options(lu="band", **TIGHT))
    for a, b in zip(ref, got):
        assert np.abs(a.z - b.z).max() <= 1e-12
    oref = oa.transient(ochain, ts, ic=ochain.zeros(), options=oa.Options(**TIGHT))
    np.testing.assert_allclose(got[0].z, oref.z, rtol=1e-9, atol=1e-12)
    # DC of a ring and of the full adder: band against partial and the oracle
    for make in (lambda m: tc.ring(m.CircuitBuilder(), 9)[0], lambda m: tc.full_adder_dc(
            m.CircuitBuilder())):
        c, o = make(vx), make(oc)
        a = vx.dc(c, options=vx.Options(**TIGHT))
        b = vx.dc(c, options=vx.Options(lu="band", **TIGHT))
        r = oa.dc(o, options=oa.Options(**TIGHT))
        assert bool(a.converged) and bool(b.converged)
        np.testing.assert_allclose(b.z, a.z, rtol=1e-10, atol=1e-13)
        np.testing.assert_allclose(b.z, r.z, rtol=1e-9, atol=1e-12)
