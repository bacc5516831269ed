"""The rest of the reference's analysis API on the GPU: `linearize`
(voltax/analysis.py:483-491), `solve_root` with its implicit-function
gradient (:269-306), and the batched dense solve behind them."""

import numpy as np
import pytest

import circuits as tc
from oracle import analysis as oa
from oracle import circuit as oc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [1, 3, 8, 13, 16, 31, 32, 33, 57, 64])
def test_dense_solve_matches_lapack(n):
    """Batched LU with partial pivoting against numpy's LAPACK solve, plain and
    transposed, on matrices that need pivoting (random, zero diagonal)."""
    import voltax_amd as vx

    rng = np.random.default_rng(n)
    B = 37
    A = rng.standard_normal((B, n, n))
    A[:, np.arange(n), np.arange(n)] = 0.0 if n > 1 else 1.0
    b = rng.standard_normal((B, n))
    for tr in (False, True):
        x = vx.dense_solve(A, b, transpose=tr)
        ref = np.linalg.solve(np.swapaxes(A, 1, 2) if tr else A, b[..., None])[..., 0]
        np.testing.assert_allclose(x, ref, rtol=1e-9, atol=1e-11 * np.abs(ref).max())
    x1 = vx.dense_solve(A[0], b[0])
    np.testing.assert_allclose(x1, np.linalg.solve(A[0], b[0]), rtol=1e-9,
                               atol=1e-11 * np.abs(x1).max())


def test_solve_root_generic():
    """The reference's fixture (tests/test_analysis.py:178-186): z^2 = p gives
    sqrt(p) and d sqrt(p)/dp = 0.25 at p = 4."""
    import voltax_amd as vx

    def residual(z, p, args):
        return z ** 2 - p

    z, ok = vx.solve_root(residual, np.array([4.0]), np.array([1.0]))
    assert ok and z[0] == pytest.approx(2.0, rel=1e-12)
    g = vx.solve_root_vjp(residual, np.array([4.0]), z, np.array([1.0]))
    assert g[0] == pytest.approx(0.25, rel=1e-10)


def test_solve_root_system_and_gradient_vs_closed_form_and_fd():
    """A diode with a series resistor, written as two unknowns (node voltage
    and current): the root equals the scalar root found by bisection, the
    damping caps the first steps at max_dv, and the gradient with respect to
    (log R, log Is) equals central differences of the solve."""
    from scipy.optimize import brentq

    import voltax_amd as vx

    VT, V = 0.025852, 1.0

    def residual(z, p, args):
        v, i = z[0], z[1]
        r, is_ = np.exp(p[0]), np.exp(p[1])
        return np.array([i - is_ * (np.exp(v / VT) - 1.0), (V - v) / r - i])

    p = np.log([1e3, 1e-14])
    z, ok = vx.solve_root(residual, p, np.array([0.0, 0.0]), options=vx.Options(rtol=1e-12,
                                                                             atol=1e-15))
    assert ok
    vd = brentq(lambda v: (V - v) / 1e3 - 1e-14 * (np.exp(v / VT) - 1.0), 0.0, 1.0, xtol=1e-15)
    assert z[0] == pytest.approx(vd, rel=1e-10)
    g_z = np.array([1.0, 0.0])  # L = v
    g = vx.solve_root_vjp(residual, p, z, g_z)
    eps = 1e-6
    for k in range(2):
        dp = np.zeros(2)
        dp[k] = eps
        up = vx.solve_root(residual, p + dp, z, options=vx.Options(rtol=1e-13, atol=1e-16))[0]
        dn = vx.solve_root(residual, p - dp, z, options=vx.Options(rtol=1e-13, atol=1e-16))[0]
        assert g[k] == pytest.approx((up[0] - dn[0]) / (2 * eps), rel=1e-6)


def test_linearize_matches_the_reference_fixture_and_the_oracle():
    """tests/test_analysis.py:169-175 on the RC ladder, and G, C of the EKV
    chain at its operating point equal the oracle's."""
    import voltax_amd as vx

    c = tc.rc_ladder(vx.CircuitBuilder())
    G, C = vx.linearize(c, vx.dc(c).z)
    a, out = c.node("a"), c.node("out")
    assert G[a, a] == pytest.approx(1 / 1e3 + 1 / 2e3, rel=1e-14)
    assert G[a, out] == pytest.approx(-1 / 2e3, rel=1e-14)
    assert C[out, out] == pytest.approx(0.5e-9, rel=1e-14)
    pulse = lambda m: m.Pulse(0, 1.2, 0.2e-9, 20e-12, 20e-12, 2e-9, 4e-9)  # noqa: E731
    c = tc.inverter_chain(vx.CircuitBuilder(), pulse(vx))
    o = tc.inverter_chain(oc.CircuitBuilder(), pulse(oc))
    z = oa.dc(o).z
    G, C = vx.linearize(c, z)
    Go, Co = oa.linearize(o, z)
    np.testing.assert_allclose(G, Go, rtol=1e-12, atol=1e-12 * np.abs(Go).max())
    np.testing.assert_allclose(C, Co, rtol=1e-12, atol=1e-12 * np.abs(Co).max())
