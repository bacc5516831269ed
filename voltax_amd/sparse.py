"""Sparse Jacobian interface with the reference's names (voltax/sparse.py).

* `structure(circuit)` -> `Structure` (CSR pattern, diagonal, rows;
  voltax/sparse.py:73-193), cached per topology.
* `jacobian(circuit, z, t, f, q)` -> stamped values of ``f*df/dz + q*dq/dz``
  on the pattern (voltax/sparse.py:220-249), computed by the HIP stamping
  kernels (one thread per device, deterministic gather per slot).
* `residual(circuit, z, t, f, q)` -> ``f*f(z,t) + q*q(z)`` (Circuit.f/q).
* `matvec` -> FP64 SpMV on the GPU.

Numeric factorization / solves live in `voltax_amd.splu`.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import torch

from . import engine
from ._lib import check, lib, ptr, stream_ptr


@dataclass(eq=False)
class Structure:
    size: int
    n_nodes: int
    indptr: np.ndarray
    indices: np.ndarray
    topo: object

    @property
    def nnz(self):
        return len(self.indices)

    @property
    def rows(self):
        return np.repeat(np.arange(self.size), np.diff(self.indptr))

    @property
    def diagonal(self):
        out = np.empty(self.size, np.int64)
        for s in range(self.size):
            lo, hi = self.indptr[s], self.indptr[s + 1]
            out[s] = lo + np.searchsorted(self.indices[lo:hi], s)
        return out

    def to_dense(self, data):
        data = data.cpu().numpy() if hasattr(data, "cpu") else np.asarray(data)
        A = np.zeros((self.size, self.size), data.dtype)
        A[self.rows, self.indices] = data
        return A

    def to_scipy(self, data):
        import scipy.sparse as sp

        data = data.cpu().numpy() if hasattr(data, "cpu") else np.asarray(data)
        return sp.csr_matrix((data, self.indices, self.indptr), shape=(self.size, self.size))

    def matvec(self, data, x):
        """``A @ x`` (FP64, on the GPU) for device tensors data [.., nnz], x [.., S]."""
        data = torch.as_tensor(data, dtype=torch.float64).cuda().contiguous()
        x = torch.as_tensor(x, dtype=torch.float64).cuda().contiguous()
        B = x.numel() // self.size
        y = torch.empty_like(x)
        check(lib.vxm_spmv(self.topo.handle, B, ptr(data), ptr(x), ptr(y), stream_ptr()),
              "vxm_spmv")
        return y


_cache: dict = {}


def structure(circuit) -> Structure:
    topo, _ = engine.topology_of(circuit)
    st = _cache.get(id(topo))
    if st is None or st.topo is not topo:
        st = Structure(topo.S, topo.nv, topo.indptr.astype(np.int64),
                       topo.indices.astype(np.int64), topo)
        _cache[id(topo)] = st
    return st


def _stamp(circuit, z, t, f, q, gmin, want_data, want_res, params=None, hist=None):
    engine.require_gpu()
    topo, row = engine.topology_of(circuit)
    z = torch.as_tensor(z, dtype=torch.float64).cuda().contiguous()
    B = z.numel() // topo.S
    if params is None:
        params = torch.from_numpy(np.tile(row, B) if len(row) else np.zeros(1)).cuda()
    work = torch.empty(max(B * topo.work, 1), dtype=torch.float64, device="cuda")
    data = torch.empty((B, topo.nnz), dtype=torch.float64, device="cuda") if want_data else None
    res = torch.empty((B, topo.S), dtype=torch.float64, device="cuda") if want_res else None
    check(lib.vxm_stamp(topo.handle, B, ptr(params), ptr(z), float(t), float(f), float(q),
                        float(gmin), ptr(hist), ptr(data), ptr(res), ptr(work), stream_ptr()),
          "vxm_stamp")
    return data, res


def jacobian(circuit, z, t=0.0, f=1.0, q=0.0, structure=None, gmin=0.0):
    """Stamped values of ``f*df/dz + q*dq/dz`` (+ gmin on node diagonals);
    shape [nnz] for one state, [B, nnz] for a batch of states."""
    data, _ = _stamp(circuit, z, t, f, q, gmin, True, False)
    return data[0] if torch.as_tensor(z).dim() == 1 else data


def residual(circuit, z, t=0.0, f=1.0, q=0.0, gmin=0.0):
    """``f*f(z,t) + q*q(z) + gmin*v`` on the GPU (Circuit.f / Circuit.q)."""
    _, res = _stamp(circuit, z, t, f, q, gmin, False, True)
    return res[0] if torch.as_tensor(z).dim() == 1 else res
