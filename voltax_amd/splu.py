"""Supernodal sparse LU on the GPU, with the reference's factor/solve
interface (voltax/sparse.py:543-621).

`Plan` is the symbolic analysis of one topology (host C++, once);
`factor(plan, data)` returns device factors, `solve(plan, factors, b,
transpose)` solves ``A x = b`` or ``A^T x = b`` (the adjoint), and
`spsolve` wraps both with the reference's VJP formula.
"""

from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass

import numpy as np
import torch

from . import engine
from ._lib import check, lib, ptr, stream_ptr

LEAF_SIZE = int(os.environ.get("VOLTAX_SPARSE_LEAF", 16))


class Plan:
    """Owns a ``vxm_splu`` handle (symbolic analysis of one topology)."""

    def __init__(self, topo, leaf: int = LEAF_SIZE, compact: bool = False):
        self.topo = topo
        self.leaf = leaf
        self.compact = bool(compact)
        self.handle = C.c_void_p()
        check(lib.vxm_splu_create_ex(topo.handle, int(leaf), 1 if compact else 0,
                                     C.byref(self.handle)), "vxm_splu_create_ex")
        sz = np.zeros(8, np.int64)
        check(lib.vxm_splu_sizes(self.handle, sz.ctypes.data), "vxm_splu_sizes")
        (self.front_total, self.piv_total, self.r_total, self.n_supernodes, self.n_levels,
         self.max_front, self.max_k, self.work_per) = (int(x) for x in sz)
        st = np.zeros(2)
        check(lib.vxm_splu_stats(self.handle, st.ctypes.data), "vxm_splu_stats")
        self.size = topo.S
        self.stats = dict(supernodes=self.n_supernodes, levels=self.n_levels,
                          lu_flops=float(st[0]), factor_entries=float(st[1]),
                          largest_front=self.max_front, largest_supernode=self.max_k)

    def structure(self):
        """(supernode of each unknown, parent of each supernode, level)."""
        sn = np.zeros(self.size, np.int32)
        par = np.zeros(self.n_supernodes, np.int32)
        lev = np.zeros(self.n_supernodes, np.int32)
        check(lib.vxm_splu_structure(self.handle, sn.ctypes.data, par.ctypes.data,
                                     lev.ctypes.data), "vxm_splu_structure")
        return sn, par, lev

    def configure(self, mid_blocked: bool):
        """Kernel for fronts of 33..64 unknowns: register wavefront (False) or
        blocked LDS workgroup (True); a schedule choice of the autotuner."""
        check(lib.vxm_splu_configure(self.handle, int(bool(mid_blocked))), "vxm_splu_configure")
        self.mid_blocked = bool(mid_blocked)

    def export(self):
        """Host copy of the symbolic layout (dict of numpy arrays)."""
        cnt = np.zeros(6, np.int64)
        z = None
        check(lib.vxm_splu_export(self.handle, cnt.ctypes.data, z, z, z, z, z, z, z, z, z),
              "vxm_splu_export")
        ns, ni, nc, nch, na, nl = (int(x) for x in cnt)
        out = dict(sn=np.zeros((ns, 10), np.int64), sn2=np.zeros((ns, 3), np.int32),
                   idx=np.zeros(ni, np.int32), cmap=np.zeros(nc, np.int32),
                   children=np.zeros(nch, np.int32), aslot=np.zeros(na, np.int32),
                   apos=np.zeros(na, np.int32), level_ptr=np.zeros(self.n_levels + 1, np.int32),
                   level_sn=np.zeros(nl, np.int32))
        check(lib.vxm_splu_export(self.handle, cnt.ctypes.data,
                                  *[out[k].ctypes.data for k in ("sn", "sn2", "idx", "cmap",
                                                                 "children", "aslot", "apos",
                                                                 "level_ptr", "level_sn")]),
              "vxm_splu_export")
        return out

    def __del__(self):
        h = getattr(self, "handle", None)
        try:
            if h is not None and h.value:
                lib.vxm_splu_destroy(h)
        except Exception:  # interpreter shutdown
            pass


@dataclass
class Factors:
    front: torch.Tensor
    piv: torch.Tensor
    batch: int
    fp32: bool
    data: torch.Tensor  # FP64 values, kept for refinement
    dtype: int = 0      # element type of the factors (F64, F32 or C128)


_plans: dict = {}


def plan_of(circuit, leaf: int = LEAF_SIZE, compact: bool = False, topo=None) -> Plan:
    """The cached plan of a circuit's topology (`topo`, if the caller holds
    it: saves building the device table again)."""
    if topo is None:
        topo, _ = engine.topology_of(circuit)
    key = (id(topo), leaf, bool(compact))
    p = _plans.get(key)
    if p is None or p.topo is not topo:
        p = Plan(topo, leaf, compact)
        _plans[key] = p
    return p


F64, F32, C128 = 0, 1, 2  # element types of the C ABI (VXM_SPLU_*)


def _is_complex(a):
    return torch.is_complex(a) if isinstance(a, torch.Tensor) else np.iscomplexobj(a)


def factor(plan: Plan, data, fp32: bool = False) -> Factors:
    """Numeric factorization of values `data` ([nnz] or [B, nnz]): FP64, FP32
    factors of FP64 values (fp32=True), or complex128 (complex `data`, as for
    the small-signal systems of `ac`)."""
    engine.require_gpu()
    cplx = _is_complex(data)
    dt = torch.complex128 if cplx else torch.float64
    data = torch.as_tensor(data, dtype=dt).cuda().contiguous()
    B = max(1, data.numel() // plan.topo.nnz)
    dtype = C128 if cplx else (F32 if fp32 else F64)
    ftype = {F64: torch.float64, F32: torch.float32, C128: torch.complex128}[dtype]
    front = torch.empty(B * plan.front_total, dtype=ftype, device="cuda")
    piv = torch.empty(max(B * plan.piv_total, 1), dtype=torch.int32, device="cuda")
    check(lib.vxm_splu_factor(plan.handle, B, ptr(data), ptr(front), ptr(piv), dtype,
                              stream_ptr()), "vxm_splu_factor")
    return Factors(front, piv, B, fp32 and not cplx, data, dtype)


def factor_solve(plan: Plan, data, b):
    """FP64 `factor` and `solve` of b ([S] or [B, S]) in one pass of the
    levels (the forward sweep of the register fronts fused into their
    factorization); returns (Factors, x), the same as factor then solve."""
    engine.require_gpu()
    data = torch.as_tensor(data, dtype=torch.float64).cuda().contiguous()
    B = max(1, data.numel() // plan.topo.nnz)
    x = torch.as_tensor(b, dtype=torch.float64).cuda().clone().contiguous()
    front = torch.empty(B * plan.front_total, dtype=torch.float64, device="cuda")
    piv = torch.empty(max(B * plan.piv_total, 1), dtype=torch.int32, device="cuda")
    work = torch.empty(max(B * plan.work_per, 1), dtype=torch.float64, device="cuda")
    check(lib.vxm_splu_factor_solve(plan.handle, B, ptr(data), ptr(front), ptr(piv), ptr(x),
                                    ptr(work), stream_ptr()), "vxm_splu_factor_solve")
    return Factors(front, piv, B, False, data, F64), x


def solve(plan: Plan, f: Factors, b, transpose: bool = False, refine: int = 0):
    """``A^-1 b`` (or ``A^-T b``); b is [S] or [B, S] (FP64).  With FP32
    factors, `refine` steps of FP64 iterative refinement (non-transposed)."""
    dtype = f.dtype
    cplx = dtype == C128
    b = torch.as_tensor(b, dtype=torch.complex128 if cplx else torch.float64).cuda().contiguous()
    shape = b.shape
    B = f.batch
    if refine and not transpose and not cplx:
        x = torch.empty_like(b)
        work = torch.empty(B * (plan.work_per + 2 * plan.size), dtype=torch.float64,
                           device="cuda")
        check(lib.vxm_splu_solve_refined(plan.handle, B, ptr(f.data), ptr(f.front), ptr(f.piv),
                                         ptr(b), ptr(x), ptr(work), int(refine), 0, int(f.fp32),
                                         stream_ptr()), "vxm_splu_solve_refined")
        return x.reshape(shape)
    x = b.clone()
    work = torch.empty(max(B * plan.work_per, 1) * (2 if cplx else 1), dtype=torch.float64,
                       device="cuda")
    check(lib.vxm_splu_solve(plan.handle, B, ptr(f.front), ptr(f.piv), ptr(x), ptr(work),
                             int(transpose), dtype, stream_ptr()), "vxm_splu_solve")
    return x.reshape(shape)


def spsolve(plan: Plan, data, b):
    """``x = A^-1 b`` for values `data` on the plan's pattern."""
    return solve(plan, factor(plan, data), b)


def spsolve_vjp(plan: Plan, data, b, g):
    """Cotangents ``(g_data, g_b)`` of ``x = A^-1 b`` (voltax/sparse.py:614-618):
    ``g_b = A^-T g``, ``g_data = -g_b[rows] * x[cols]``."""
    f = factor(plan, data)
    x = solve(plan, f, b)
    gb = solve(plan, f, g, transpose=True)
    topo = plan.topo
    rows = torch.from_numpy(np.repeat(np.arange(topo.S), np.diff(topo.indptr))).cuda()
    cols = torch.from_numpy(topo.indices.astype(np.int64)).cuda()
    return -gb[rows] * x[cols], gb
