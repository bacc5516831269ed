"""Analyses with the reference's API (voltax/analysis.py:50-572), run on the GPU.

`dc` and `transient` keep the reference's signatures, options and results.
Below 65 unknowns they run the fused batched Newton kernel (one wavefront
per instance group, the whole time loop inside one launch); larger systems
run stamping + supernodal sparse LU kernels driven step by step.
`dc_batch` / `transient_batch` are the vmap-style entry points: thousands of
instances of uneven size, grouped by pattern, in one launch.
"""

from __future__ import annotations

import os
from dataclasses import dataclass

import numpy as np

from .circuit import Circuit, Layout

SPARSE_THRESHOLD = 200
DEFAULT_SOLVER = os.environ.get("VOLTAX_SOLVER", "auto")
DENSE_MAX = 64  # largest system of the fused dense kernel


@dataclass(frozen=True)
class Options:
    """Newton settings (voltax/analysis.py:50-89)."""

    max_steps: int = 100
    rtol: float = 1e-6
    atol: float = 1e-9
    max_dv: float = 1.0
    gmin: float = 1e-12
    gmin_steps: int = 10
    gmin_start: float = 1e-2
    direct_steps: int = 25
    solver: str = DEFAULT_SOLVER
    sparse_threshold: int = SPARSE_THRESHOLD
    precision: str = "fp64"  # "fp64" LU, or "mixed": FP32 LU + FP64 refinement
    # backward-Euler steps on the fused dense path start Newton from the cubic
    # through the last four points, extrapolated (same roots; fewer iterations)
    predict: bool = False
    # time steps on the fused dense path reuse the step's FP64 LU factors in
    # later Newton iterations while the updates contract (simplified Newton;
    # same equations and convergence test; precision "fp64" only)
    reuse: bool = False
    # LU of the fused dense path: "partial" pivoting (the reference's
    # jnp.linalg.solve), or "band": static pivots in a band numbering of the
    # pattern where one exists, checked per factorization (multipliers within
    # 100, else partial pivoting); same roots up to round-off
    lu: str = "partial"
    # "band": the largest |multiplier| a static pivot may produce before the
    # factorization falls back to partial pivoting
    band_max_mult: float = 100.0
    # pivot groups of the sparse LU's plan: "reference" (each device with
    # internal unknowns glued to its terminals, voltax/sparse.py:257-305) or
    # "compact" (internal unknowns together, unknowns without a diagonal
    # glued to their neighbours): smaller fronts where devices with internal
    # nodes share a supply; same roots up to round-off
    grouping: str = "reference"

    def __post_init__(self):
        if self.solver not in ("auto", "dense", "sparse"):
            raise ValueError("solver must be 'auto', 'dense' or 'sparse', "
                             f"got {self.solver!r}")
        if self.precision not in ("fp64", "mixed"):
            raise ValueError(f"precision must be 'fp64' or 'mixed', got {self.precision!r}")
        if self.reuse and self.precision != "fp64":
            raise ValueError("reuse needs precision 'fp64'")
        if self.lu not in ("partial", "band"):
            raise ValueError(f"lu must be 'partial' or 'band', got {self.lu!r}")
        if self.grouping not in ("reference", "compact"):
            raise ValueError(f"grouping must be 'reference' or 'compact', got {self.grouping!r}")


@dataclass
class Solution:
    """Result of `dc` or `transient` (voltax/analysis.py:92-139)."""

    t: np.ndarray
    z: np.ndarray
    converged: np.ndarray
    layout: Layout
    iters: np.ndarray | None = None

    def v(self, node=None):
        if node is None:
            return self.z[..., : self.layout.n_nodes]
        idx = self.layout.node(node)
        return np.zeros(self.z.shape[:-1]) if idx < 0 else self.z[..., idx]

    def i(self, device, which=None):
        return self.z[..., self.layout.internal(device, which)]

    def check(self):
        bad = ~(np.asarray(self.converged) & np.all(np.isfinite(self.z), axis=-1))
        if np.any(bad):
            where = "" if np.ndim(bad) == 0 else f" at t={self.t[np.argmax(bad)]}"
            raise RuntimeError(f"Newton did not converge{where}.")
        return self

    def __getitem__(self, name):
        if "." in name:
            device, which = name.rsplit(".", 1)
            return self.i(device, which)
        return self.v(name)


def _precision(options):
    return 1 if options.precision == "mixed" else 0


def _initial(circuit, ic, guess=None):
    if isinstance(ic, Solution):
        return np.asarray(ic.z, float), False
    if isinstance(ic, str):
        if ic == "dc":
            return (circuit.zeros() if guess is None else np.asarray(guess, float)), True
        if ic == "zero":
            return circuit.zeros(), False
        raise ValueError(f"ic must be 'dc', 'zero' or an array, got {ic!r}")
    return np.asarray(ic, float), False


def _dense_ok(circuit, options):
    # Verilog-A devices are evaluated by the stamping kernels (sparse path)
    return circuit.size <= DENSE_MAX and options.solver != "sparse" and not circuit.has_va


def dc_batch(circuits, t=0.0, guesses=None, options=Options()):
    """DC operating points of many circuits in one launch (list of Solutions)."""
    from .engine import DenseBatch

    circuits = list(circuits)
    guesses = guesses or [c.zeros() for c in circuits]
    batch = DenseBatch(circuits)
    z0 = np.concatenate([np.asarray(g, float) for g in guesses])
    z, conv, iters = batch.run(z0, None, True, t, options, precision=_precision(options))
    zs = batch.split(z, 1)
    conv, iters = conv.cpu().numpy(), iters.cpu().numpy()
    return [Solution(np.asarray(t, float), zs[i][0], bool(conv[i, 0]), c.layout, iters[i, 0])
            for i, c in enumerate(circuits)]


def transient_batch(circuits, ts, ic="dc", method="be", options=Options(), batch=None,
                    raw=False):
    """Transient of many circuits on one time grid, one fused launch.

    ``ic`` is "dc", "zero", or a list of initial states.  With ``raw=True``
    returns the flat device outputs (z, conv, iters) and the batch instead of
    Solutions (what the benchmark times).
    """
    from .engine import DenseBatch

    circuits = list(circuits)
    ts = np.asarray(ts, float)
    batch = batch or DenseBatch(circuits)
    if isinstance(ic, (list, tuple)):
        z0 = np.concatenate([np.asarray(x, float) for x in ic])
        do_dc = False
    else:
        starts = [_initial(c, ic) for c in circuits]
        z0 = np.concatenate([s[0] for s in starts])
        do_dc = starts[0][1]
    z, conv, iters = batch.run(z0, ts, do_dc, ts[0], options, method,
                               precision=_precision(options))
    if raw:
        return z, conv, iters, batch
    zs = batch.split(z, len(ts))
    conv, iters = conv.cpu().numpy().astype(bool), iters.cpu().numpy()
    return [Solution(ts, zs[i], conv[i], c.layout, iters[i]) for i, c in enumerate(circuits)]


def dc(circuit: Circuit, t=0.0, guess=None, options=Options()) -> Solution:
    """DC operating point with the gmin-stepping fallback (voltax/analysis.py:410-431)."""
    if _dense_ok(circuit, options):
        return dc_batch([circuit], t, None if guess is None else [guess], options)[0]
    from .sparse_newton import dc_sparse

    return dc_sparse(circuit, t, guess, options)


def transient(circuit: Circuit, ts, ic="dc", method="be", options=Options()) -> Solution:
    """Transient on a fixed grid (voltax/analysis.py:434-480)."""
    if method not in ("be", "trap"):
        raise ValueError(f"method must be 'be' or 'trap', got {method!r}")
    if _dense_ok(circuit, options):
        z0, do_dc = _initial(circuit, ic)
        return transient_batch([circuit], ts, [z0] if not do_dc else "dc", method,
                               options)[0]
    from .sparse_newton import transient_sparse

    return transient_sparse(circuit, ts, ic, method, options)


def dc_sweep(circuit: Circuit, values, apply, t=0.0, guess=None,
             options=Options()) -> Solution:
    """DC sweep with continuation (voltax/analysis.py:532-572): point n starts
    Newton (with the gmin-stepping fallback) from the root of point n-1, so the
    sweep follows one solution branch.

    `apply` is a device name (sets that source's DC value) or
    ``fn(circuit, value) -> circuit``; it acts on a private copy of `circuit`.
    The returned Solution has ``t`` = the sweep values and ``z`` (N, S)."""
    import copy

    values = np.asarray(values, float)
    work = copy.deepcopy(circuit)
    if isinstance(apply, str):
        name = apply
        apply = lambda c, v: c.set(name, value=v)  # noqa: E731
    z = circuit.zeros() if guess is None else np.asarray(guess, float)
    zs, oks, its = [], [], []
    for v in values:
        sol = dc(apply(work, float(v)), t, guess=z, options=options)
        z = np.asarray(sol.z, float)
        zs.append(z)
        oks.append(bool(np.all(sol.converged)))
        its.append(int(np.sum(sol.iters)))
    return Solution(values, np.stack(zs), np.array(oks), circuit.layout, np.array(its))


@dataclass
class ACSolution:
    """Result of `ac` (voltax/analysis.py:142-175): complex phasors per frequency."""

    freqs: np.ndarray
    z: np.ndarray  # (F, S) complex small-signal state
    op: Solution
    layout: Layout

    def v(self, node=None):
        if node is None:
            return self.z[..., : self.layout.n_nodes]
        idx = self.layout.node(node)
        return np.zeros(self.z.shape[:-1], complex) if idx < 0 else self.z[..., idx]

    def i(self, device, which=None):
        return self.z[..., self.layout.internal(device, which)]

    def db(self, node):
        return 20 * np.log10(np.abs(self.v(node)))

    def phase(self, node, unwrap=True):
        ang = np.angle(self.v(node))
        return np.rad2deg(np.unwrap(ang) if unwrap else ang)


def ac(circuit: Circuit, freqs, op=None, options=Options()) -> ACSolution:
    """Small-signal AC analysis around the DC operating point
    (voltax/analysis.py:494-524): ``(G + j 2 pi f C) z = -b`` for every
    frequency, G = df/dz and C = dq/dz stamped on the GPU at the operating
    point, all frequencies factored and solved as one complex batch of the
    supernodal LU (one plan per topology, complex128 fronts)."""
    import torch

    from . import engine, splu
    from ._lib import check, lib, ptr, stream_ptr

    op = dc(circuit, options=options) if op is None else op
    freqs = np.atleast_1d(np.asarray(freqs, float))
    topo, row = engine.topology_of(circuit)
    dev = dict(dtype=torch.float64, device="cuda")
    params = torch.as_tensor(row if len(row) else np.zeros(1), **dev)
    z = torch.as_tensor(np.asarray(op.z, float).reshape(-1)[: circuit.size], **dev)
    g = torch.empty(topo.nnz, **dev)
    c = torch.empty(topo.nnz, **dev)
    res = torch.empty(topo.S, **dev)
    work = torch.empty(max(topo.work, 1), **dev)
    t = float(np.asarray(op.t).reshape(-1)[0]) if np.size(op.t) else 0.0
    s = stream_ptr()
    for cf, cq, out in ((1.0, 0.0, g), (0.0, 1.0, c)):
        check(lib.vxm_stamp(topo.handle, 1, ptr(params), ptr(z), t, cf, cq, 0.0, None, ptr(out),
                            ptr(res), ptr(work), s), "vxm_stamp")
    b = torch.from_numpy(circuit.ac_stimulus().astype(np.complex128)).cuda()
    F = len(freqs)
    fr = torch.as_tensor(freqs, **dev)
    data = torch.empty((F, topo.nnz), dtype=torch.complex128, device="cuda")
    rhs = torch.empty((F, topo.S), dtype=torch.complex128, device="cuda")
    check(lib.vxm_ac_system(topo.nnz, topo.S, F, ptr(g), ptr(c), ptr(fr), ptr(b), ptr(data),
                            ptr(rhs), s), "vxm_ac_system")
    plan = splu.plan_of(circuit, compact=options.grouping == "compact")
    x = splu.solve(plan, splu.factor(plan, data), rhs)
    return ACSolution(freqs, x.cpu().numpy(), op, circuit.layout)


def linearize(circuit: Circuit, z, t=0.0):
    """Small-signal matrices ``G = df/dz`` and ``C = dq/dz`` at state `z` as dense
    (S, S) numpy arrays (voltax/analysis.py:483-491), stamped on the GPU
    (`sparse.jacobian` gives the values on the pattern)."""
    from . import sparse

    st = sparse.structure(circuit)
    z = np.asarray(z, float)
    G = sparse.jacobian(circuit, z, t, f=1.0, q=0.0, structure=st)
    C = sparse.jacobian(circuit, z, t, f=0.0, q=1.0, structure=st)
    return st.to_dense(G), st.to_dense(C)


def dense_solve(A, b, transpose=False):
    """x = A^-1 b (A^-T b) for one (n, n) or a batch (B, n, n) of matrices, n <= 64,
    by the GPU's batched LU with partial pivoting (vxm_dense_solve); numpy in and out."""
    import torch

    from . import engine
    from ._lib import check, lib, ptr, stream_ptr

    engine.require_gpu()
    A = np.asarray(A, float)
    b = np.asarray(b, float)
    one = A.ndim == 2
    A3 = A.reshape((-1,) + A.shape[-2:])
    n = A3.shape[-1]
    At = torch.as_tensor(np.ascontiguousarray(A3)).cuda()
    bt = torch.as_tensor(np.ascontiguousarray(b.reshape(A3.shape[0], n))).cuda()
    x = torch.empty_like(bt)
    check(lib.vxm_dense_solve(A3.shape[0], n, ptr(At), ptr(bt), ptr(x), int(bool(transpose)),
                              stream_ptr()), "vxm_dense_solve")
    out = x.cpu().numpy()
    return out[0] if one else out.reshape(b.shape)


_CSTEP = 1e-30  # complex-step size: derivative error h^2 |f'''| / 6, far below rounding


def _cstep_jacobian(fn, x):
    """d fn / d x by complex steps (Squire and Trapp 1998): exact to rounding for
    residuals written with operations that extend analytically to complex
    arguments (+ - * /, powers, exp, log, ...)."""
    x = np.asarray(x, float)
    cols = []
    for j in range(x.size):
        xc = x.astype(complex)
        xc.flat[j] += 1j * _CSTEP
        cols.append(np.imag(np.asarray(fn(xc), complex)).ravel() / _CSTEP)
    return np.stack(cols, axis=-1) if cols else np.zeros((0, 0))


def solve_root(residual, params, guess, args=None, n_nodes=None, options=Options(),
               jacobian=None):
    """Solve ``residual(z, params, args) = 0`` for z (voltax/analysis.py:291-306).

    The reference's damped Newton (:208-233): ``dz = -J^-1 r``, the update scaled
    by ``min(1, max_dv / max|dz[:n_nodes]|)``, converged when every
    ``|dz| <= atol + rtol |z_new|`` with no damping; at most `options.max_steps`
    iterations.  The linear solves run on the GPU (`dense_solve`; n <= 64).
    `residual` takes and returns numpy arrays; the Jacobian ``d residual / dz``
    comes from `jacobian(z, params, args)` when given, else from complex steps
    of `residual`.  Returns ``(z, converged)``; gradients with respect to
    `params` come from `solve_root_vjp` (the implicit-function VJP of
    :282-288)."""
    z = np.array(guess, float).reshape(-1)
    n = z.size if n_nodes is None else int(n_nodes)
    jac = jacobian or (lambda zz, p, a: _cstep_jacobian(lambda u: residual(u, p, a), zz))
    done = False
    for _ in range(options.max_steps):
        r = np.asarray(residual(z, params, args), float).reshape(-1)
        J = np.asarray(jac(z, params, args), float).reshape(z.size, z.size)
        dz = -dense_solve(J, r)
        dv = np.max(np.abs(dz[:n]), initial=0.0)
        scale = min(1.0, options.max_dv / (dv + 1e-300))
        zn = z + scale * dz
        small = np.all(np.abs(dz) <= options.atol + options.rtol * np.abs(zn))
        done = bool(small and scale == 1.0 and np.all(np.isfinite(zn)))
        z = zn
        if done:
            break
    return z, done


def solve_root_vjp(residual, params, z, g_z, args=None, jacobian=None, param_jacobian=None):
    """Gradient of a loss with dL/dz = g_z at a root z of `solve_root`
    (voltax/analysis.py:282-288): ``lam = J^-T g_z`` (a transposed GPU solve at the
    root), ``dL/dparams = -lam^T d residual / d params``.  `params` is an array;
    the parameter Jacobian comes from `param_jacobian(z, params, args)` or from
    complex steps of `residual` in `params`."""
    z = np.asarray(z, float).reshape(-1)
    p = np.asarray(params, float)
    jac = jacobian or (lambda zz, pp, a: _cstep_jacobian(lambda u: residual(u, pp, a), zz))
    pjac = param_jacobian or (
        lambda zz, pp, a: _cstep_jacobian(lambda q: residual(zz, q.reshape(pp.shape), a), pp))
    J = np.asarray(jac(z, p, args), float).reshape(z.size, z.size)
    lam = dense_solve(J, np.asarray(g_z, float).reshape(-1), transpose=True)
    Jp = np.asarray(pjac(z, p, args), float).reshape(z.size, p.size)
    return (-(lam @ Jp)).reshape(p.shape)
