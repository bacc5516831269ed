"""Forward-mode dual numbers over numpy arrays (oracle / test infrastructure).

The reference builds per-device Jacobian blocks with ``jax.jvp`` seeded on
each local unknown (voltax/sparse.py:240-244) and parameter cotangents with
``jax.vjp`` (voltax/analysis.py:286-288).  JAX is not installed here, so the
oracle carries exact first derivatives with this small dual type instead:
``Dual(v, d)`` holds values ``v`` of shape ``(N,)`` and tangents ``d`` of
shape ``(N, P)`` (P seed directions).  Every rule below is the textbook
derivative of the elementary function; nothing is approximated.
"""

from __future__ import annotations

import numpy as np


class Dual:
    __slots__ = ("v", "d")
    __array_priority__ = 100  # make numpy defer to our reflected operators

    def __init__(self, v, d):
        self.v = np.asarray(v, dtype=float)
        self.d = np.asarray(d, dtype=float)

    # -- construction helpers
    @staticmethod
    def seed(values, k, p):
        """Variable number ``k`` of ``p`` seed directions."""
        v = np.asarray(values, dtype=float)
        d = np.zeros(v.shape + (p,))
        d[..., k] = 1.0
        return Dual(v, d)

    def _lift(self, o):
        if isinstance(o, Dual):
            return o
        o = np.asarray(o, dtype=float)
        return Dual(o, np.zeros(np.broadcast(o, self.v).shape + self.d.shape[-1:]))

    # -- arithmetic
    def __add__(self, o):
        o = self._lift(o)
        return Dual(self.v + o.v, self.d + o.d)

    __radd__ = __add__

    def __sub__(self, o):
        o = self._lift(o)
        return Dual(self.v - o.v, self.d - o.d)

    def __rsub__(self, o):
        return self._lift(o) - self

    def __neg__(self):
        return Dual(-self.v, -self.d)

    def __mul__(self, o):
        o = self._lift(o)
        return Dual(self.v * o.v, self.d * o.v[..., None] + o.d * self.v[..., None])

    __rmul__ = __mul__

    def __truediv__(self, o):
        o = self._lift(o)
        q = self.v / o.v
        return Dual(q, (self.d - o.d * q[..., None]) / o.v[..., None])

    def __rtruediv__(self, o):
        return self._lift(o) / self

    def __pow__(self, n):
        # integer or real constant exponent only
        return Dual(self.v ** n, (n * self.v ** (n - 1))[..., None] * self.d)


def val(x):
    return x.v if isinstance(x, Dual) else np.asarray(x, dtype=float)


def tan(x, shape):
    """Tangent of ``x`` broadcast to ``shape`` (zeros for constants)."""
    if isinstance(x, Dual):
        return np.broadcast_to(x.d, shape)
    return np.zeros(shape)


def _unary(x, f, df):
    if isinstance(x, Dual):
        return Dual(f(x.v), df(x.v)[..., None] * x.d)
    return f(np.asarray(x, dtype=float))


def exp(x):
    return _unary(x, np.exp, np.exp)


def log(x):
    return _unary(x, np.log, lambda v: 1.0 / v)


def sqrt(x):
    return _unary(x, np.sqrt, lambda v: 0.5 / np.sqrt(v))


def sigmoid_np(v):
    return 0.5 * (1.0 + np.tanh(0.5 * v))


def softplus(x):
    """``log(1 + exp(x))`` evaluated as ``max(x, 0) + log1p(exp(-|x|))``
    (the overflow-safe form of jax.nn.softplus); derivative ``sigmoid(x)``."""
    f = lambda v: np.maximum(v, 0.0) + np.log1p(np.exp(-np.abs(v)))  # noqa: E731
    return _unary(x, f, sigmoid_np)


def sin(x):
    return _unary(x, np.sin, np.cos)


def select(cond, a, b):
    """``where(cond, a, b)`` on values; the tangent of the branch taken (the
    derivative JAX's jvp gives away from the switch point)."""
    cond = np.asarray(cond)
    if not isinstance(a, Dual) and not isinstance(b, Dual):
        return np.where(cond, a, b)
    ref = a if isinstance(a, Dual) else b
    a, b = ref._lift(a), ref._lift(b)
    shape = np.broadcast(a.v, b.v, cond).shape
    return Dual(np.where(cond, a.v, b.v),
                np.where(cond[..., None], np.broadcast_to(a.d, shape + a.d.shape[-1:]),
                         np.broadcast_to(b.d, shape + b.d.shape[-1:])))


def clip01(x):
    v = val(x)
    return select(v < 0.0, 0.0, select(v > 1.0, 1.0, x))


def minimum(a, b):
    return select(val(b) < val(a), b, a)


def floor_mod(x, y):
    """``x - y * floor(x / y)`` (numpy's remainder, jnp's ``%``); derivatives
    1 in x and ``-floor(x / y)`` in y."""
    if not isinstance(x, Dual) and not isinstance(y, Dual):
        return np.remainder(x, y)
    ref = x if isinstance(x, Dual) else y
    x, y = ref._lift(x), ref._lift(y)
    k = np.floor(x.v / y.v)
    return Dual(np.remainder(x.v, y.v), x.d - k[..., None] * y.d)


def item(x, idx):
    """Entry ``idx`` of a Dual (or array) along its value axes."""
    if isinstance(x, Dual):
        return Dual(x.v[idx], x.d[idx])
    return np.asarray(x, dtype=float)[idx]
