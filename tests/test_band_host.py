"""Host checks of the static-pivot band LU of the fused dense kernel
(Options(lu="band"), VXM_BAND_LU): the band numbering of the ring and chain
patterns, and a numpy replay of the kernel's elimination order on the
oracle's stamped Jacobians.

The replay follows the kernel step for step (lu_factor_band in
voltax_amd/csrc/dense_newton.hpp): static pivots on the diagonal for
positions below smax-2, updating only the next two columns and the border,
then partial pivoting inside the 2x2 border block.  It pins the claim the
kernel rests on -- in the band numbering every column it skips is exactly
zero in the pivot row -- and that the result is the solution LAPACK gives.
"""

import ctypes

import numpy as np
import pytest

import circuits as tc
from oracle import circuit as oc
from oracle import sparse as osp

WIDTH, BORDER = 2, 2  # kBandWidth, kBandBorder (voltax_amd/csrc/kernels.hpp)


def _band_pos(c, smax):
    from voltax_amd import engine
    from voltax_amd._lib import lib

    topo, _ = engine.topology_of(c)
    pos = np.full(c.size, -1, np.int32)
    st = lib.vxm_topology_band(topo.handle, smax, pos.ctypes.data_as(ctypes.c_void_p))
    return st, pos


def _smax(S):
    from voltax_amd.engine import dense_classes

    smax, _ = dense_classes()
    return int(smax[np.argmax(smax >= S)])


def _ring_pair(n, seed=0):
    import voltax_amd as vx

    rng = np.random.default_rng(seed)
    dv, ws = rng.normal(0, 0.03, 2 * n), np.exp(rng.normal(0, 0.05, 2 * n))
    c, nd = tc.ring(vx.CircuitBuilder(), n, dvth=dv, wscale=ws)
    o, ond = tc.ring(oc.CircuitBuilder(), n, dvth=dv, wscale=ws)
    return c, o, nd


@pytest.mark.parametrize("n", [3, 5, 11, 21, 27, 29, 31])
def test_ring_patterns_have_a_band_numbering(n):
    c, _, _ = _ring_pair(n)
    smax = _smax(c.size)
    st, pos = _band_pos(c, smax)
    assert st == 0
    assert len(set(pos.tolist())) == c.size and pos.min() >= 0 and pos.max() < smax
    # the supply node and its branch current form the border, node first
    vdd, ivdd = c.node("vdd"), c.size - 1
    assert pos[vdd] == smax - 2 and pos[ivdd] == smax - 1
    # every other coupling lies within two positions
    kind, _, unk, _, _ = c.device_table()
    for u in unk:
        u = [x for x in u if x >= 0]
        for i in u:
            for j in u:
                if pos[i] < smax - BORDER and pos[j] < smax - BORDER:
                    assert abs(int(pos[i]) - int(pos[j])) <= WIDTH


def test_chain_with_two_sources_has_no_band_numbering():
    import voltax_amd as vx

    c = tc.inverter_chain(vx.CircuitBuilder(), vx.Pulse(0, 1.2, 0.2e-9, 20e-12, 20e-12, 2e-9,
                                                         4e-9))
    st, _ = _band_pos(c, _smax(c.size))
    assert st != 0  # four source unknowns do not fit a border of two


def band_lu_replay(A, b):
    """The kernel's band LU and back substitution on the permuted system A
    (smax x smax, padding rows identity).  Returns (x, largest multiplier,
    largest magnitude met in a skipped column of a pivot row)."""
    n = A.shape[0]
    KB = n - BORDER
    a, y = A.astype(float).copy(), b.astype(float).copy()
    assigned = np.zeros(n, bool)
    myk = np.full(n, n)
    lmax = skipped = 0.0
    for k in range(KB):
        assigned[k], myk[k] = True, k
        rows = ~assigned
        l = a[rows, k] / a[k, k]
        lmax = max(lmax, float(np.abs(l).max(initial=0.0)))
        a[rows, k] = l
        cols = list(range(k + 1, min(k + 1 + WIDTH, KB))) + list(range(KB, n))
        skip = [j for j in range(k + 1, KB) if j not in cols]
        skipped = max(skipped, float(np.abs(a[k, skip]).max(initial=0.0)))
        a[np.ix_(rows, cols)] -= np.outer(l, a[k, cols])
        y[rows] -= l * y[k]
    for k in range(KB, n):  # border block: partial pivoting among unused rows
        cand = np.flatnonzero(~assigned)
        p = cand[np.argmax(np.abs(a[cand, k]))]
        assigned[p], myk[p] = True, k
        rows = ~assigned
        l = a[rows, k] / a[p, k]
        a[rows, k] = l
        a[np.ix_(rows, range(k + 1, n))] -= np.outer(l, a[p, k + 1:])
        y[rows] -= l * y[p]
    x = np.zeros(n)
    for k in range(n - 1, -1, -1):  # pivot row of step k, then U above it
        p = int(np.flatnonzero(myk == k)[0])
        x[k] = y[p] / a[p, k]
        upper = myk < k
        y[upper] -= a[upper, k] * x[k]
    return x, lmax, skipped


@pytest.mark.parametrize("n", [3, 11, 27, 31])
@pytest.mark.parametrize("cq", [0.0, 1.0 / 2e-12])
def test_band_lu_replay_on_stamped_jacobians_equals_lapack(n, cq):
    c, o, _ = _ring_pair(n, seed=n)
    smax = _smax(c.size)
    st, pos = _band_pos(c, smax)
    assert st == 0
    rng = np.random.default_rng(1)
    z = rng.uniform(0.0, 1.2, o.size)
    z[o.size - 1] = rng.normal(0, 1e-4)
    ost = osp.structure(o)
    data = osp.jacobian(o, z, 1e-10, 1.0, cq, ost)
    J = np.zeros((o.size, o.size))
    for r in range(o.size):
        J[r, ost.indices[ost.indptr[r]:ost.indptr[r + 1]]] = data[ost.indptr[r]:ost.indptr[r + 1]]
    J[np.arange(o.n_nodes), np.arange(o.n_nodes)] += 1e-12  # gmin on node diagonals
    A = np.eye(smax)
    A[np.ix_(pos, pos)] = J
    rhs = rng.normal(size=o.size)
    bp = np.zeros(smax)
    bp[pos] = rhs
    with np.errstate(all="ignore"):
        x, lmax, skipped = band_lu_replay(A, bp)
    want = np.linalg.solve(J, rhs)
    if cq > 0:
        # transient Jacobians (C/h on the diagonals): the static pivots pass
        # the kernel's multiplier test and solve to round-off
        assert skipped == 0.0  # the numbering's promise: nothing is lost by skipping
        assert lmax <= 100.0
        np.testing.assert_allclose(x[pos], want, rtol=1e-12, atol=1e-13 * np.abs(want).max())
    elif lmax <= 100.0:
        np.testing.assert_allclose(x[pos], want, rtol=1e-9, atol=1e-12 * np.abs(want).max())
    else:
        # a DC Jacobian at a random state: static pivots blow up (the kernel's
        # test sees the multiplier and refactors with partial pivoting)
        assert n >= 11
