"""GPU resources behind the analyses: topology handles (cached per circuit
pattern) and fused dense batches.  Device memory and streams come from
torch; all arithmetic happens in the HIP kernels of libvoltax_mi355x.so.
"""

from __future__ import annotations

import ctypes as C
import os
from functools import lru_cache

import numpy as np
import torch

from . import _lib
from ._lib import check, lib, ptr, stream_ptr


def require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("voltax_amd needs a HIP GPU (MI355X); no CPU fallback exists")


class Topology:
    """Owns a ``vxm_topology``: pattern + stamping maps of one circuit shape."""

    def __init__(self, size, n_nodes, kind, wave, unknowns, poff, n_params):
        self.handle = C.c_void_p()
        D = len(kind)
        unknowns = np.asarray(unknowns, np.int32).reshape(D, -1) if D else np.zeros((0, 4),
                                                                                   np.int32)
        self.width = max(4, unknowns.shape[1])
        if unknowns.shape[1] < self.width:
            unknowns = np.concatenate(
                [unknowns, np.full((D, self.width - unknowns.shape[1]), -1, np.int32)], 1)
        self._arrays = [np.ascontiguousarray(a, np.int32) for a in (kind, wave, unknowns, poff)]
        check(lib.vxm_topology_create_wide(size, n_nodes, D, n_params, self.width,
                                           *[a.ctypes.data for a in self._arrays],
                                           C.byref(self.handle)), "vxm_topology_create_wide")
        sizes = np.zeros(6, np.int64)
        check(lib.vxm_topology_sizes(self.handle, sizes.ctypes.data), "vxm_topology_sizes")
        self.S, self.nv, self.D, self.P, self.nnz, self.work = (int(x) for x in sizes)
        self.indptr = np.zeros(self.S + 1, np.int32)
        self.indices = np.zeros(self.nnz, np.int32)
        check(lib.vxm_topology_pattern(self.handle, self.indptr.ctypes.data,
                                       self.indices.ctypes.data), "vxm_topology_pattern")
        self._dev = None

    @property
    def pattern_dev(self):
        if self._dev is None:
            self._dev = (torch.from_numpy(self.indptr).cuda(), torch.from_numpy(self.indices).cuda())
        return self._dev

    def __del__(self):
        h = getattr(self, "handle", None)
        try:
            if h is not None and h.value:
                lib.vxm_topology_destroy(h)
        except Exception:  # interpreter shutdown
            pass


@lru_cache(maxsize=256)
def _topology_cached(key, size, n_nodes, kind_b, wave_b, unk_b, poff_b, n_params, code_b=b""):
    arr = lambda b, shape=None: np.frombuffer(b, np.int32)  # noqa: E731
    kind = arr(kind_b)
    topo = Topology(size, n_nodes, kind, arr(wave_b), arr(unk_b).reshape(len(kind), -1),
                    arr(poff_b), n_params)
    if code_b:
        code = np.frombuffer(code_b, np.int32).copy()
        topo._code = code
        check(lib.vxm_topology_set_code(topo.handle, len(code), code.ctypes.data),
              "vxm_topology_set_code")
    return topo


def topology_of(circuit):
    """(Topology, parameter row) of a circuit; topologies are cached per pattern."""
    kind, wave, unk, poff, row = circuit.device_table()
    code = circuit.va_code()[0] if circuit.has_va else np.zeros(0, np.int32)
    # the Verilog-A evaluation path is chosen when a topology takes its programs
    # (VXM_VA_CODEGEN=all: generated code for every program), so it keys the cache
    path = os.environ.get("VXM_VA_CODEGEN", "") if circuit.has_va else ""
    topo = _topology_cached((circuit.topology_key(), path), circuit.size, circuit.n_nodes,
                            kind.tobytes(), wave.tobytes(), unk.tobytes(), poff.tobytes(),
                            len(row), code.tobytes())
    return topo, row


@lru_cache(maxsize=1)
def dense_classes():
    """(smax, lanes) arrays of the fused dense kernel's tile classes."""
    n = C.c_int32()
    check(lib.vxm_dense_classes(C.byref(n), None, None), "vxm_dense_classes")
    smax = np.zeros(n.value, np.int32)
    lanes = np.zeros(n.value, np.int32)
    check(lib.vxm_dense_classes(C.byref(n), smax.ctypes.data, lanes.ctypes.data),
          "vxm_dense_classes")
    return smax, lanes


def options_struct(o) -> _lib.NewtonOptions:
    return _lib.NewtonOptions(o.max_steps, o.rtol, o.atol, o.max_dv, o.gmin, o.gmin_steps,
                              o.gmin_start, o.direct_steps, getattr(o, "band_max_mult", 100.0))


class DenseBatch:
    """Many small circuits (S <= 64), possibly of different patterns, run by
    the fused Newton kernel in one launch per tile class."""

    def __init__(self, circuits):
        require_gpu()
        self.circuits = list(circuits)
        topos, index, inst_topo, rows = [], {}, [], []
        for c in self.circuits:
            t, row = topology_of(c)
            if id(t) not in index:
                index[id(t)] = len(topos)
                topos.append(t)
            inst_topo.append(index[id(t)])
            rows.append(row)
        self.topos = topos
        self.inst_topo = np.asarray(inst_topo, np.int32)
        lens = np.array([len(r) for r in rows], np.int64)
        self.param_off = np.concatenate([[0], np.cumsum(lens)[:-1]]).astype(np.int64)
        self.params = torch.from_numpy(np.concatenate(rows) if len(rows) else np.zeros(1)).cuda()
        self.sizes = np.array([topos[i].S for i in inst_topo], np.int64)
        self.state_off = np.concatenate([[0], np.cumsum(self.sizes)[:-1]]).astype(np.int64)
        handles = (C.c_void_p * len(topos))(*[t.handle.value for t in topos])
        self.handle = C.c_void_p()
        check(lib.vxm_batch_create(len(topos), handles, len(self.circuits),
                                   self.inst_topo.ctypes.data, self.param_off.ctypes.data,
                                   C.byref(self.handle)), "vxm_batch_create")

    def set_params(self, rows):
        """Replace every instance's parameter row (same layout)."""
        self.params.copy_(torch.from_numpy(np.concatenate(rows)))

    def run(self, z0, ts, do_dc, t_dc, options, method="be", precision=0, stream=None,
            predict=None, reuse=None, lu=None):
        """Returns (z [sum T*S_i] flat, conv [n, T], iters [n, T]) device tensors.
        ``predict`` (default: ``options.predict``) starts each time step's Newton
        iteration from the cubic through the last four points, extrapolated;
        ``reuse`` (default: ``options.reuse``) lets later iterations of a time
        step reuse its LU factors while the updates contract; ``lu`` (default:
        ``options.lu``) "band" factors the instances whose pattern has a band
        numbering (`band_patterns`) with checked static pivots."""
        if isinstance(ts, torch.Tensor):
            ts_t = ts
        else:
            ts_t = torch.as_tensor(np.asarray(ts, float)).cuda()
        T = int(ts_t.numel()) if ts is not None else 1
        z0_t = z0 if isinstance(z0, torch.Tensor) else torch.as_tensor(
            z0, dtype=torch.float64).cuda()
        n = len(self.circuits)
        out_z = torch.empty(int(self.sizes.sum()) * T, dtype=torch.float64, device="cuda")
        conv = torch.empty((n, T), dtype=torch.int8, device="cuda")
        iters = torch.empty((n, T), dtype=torch.int32, device="cuda")
        opts = options_struct(options)
        if predict is None:
            predict = getattr(options, "predict", False)
        if reuse is None:
            reuse = getattr(options, "reuse", False)
        if lu is None:
            lu = getattr(options, "lu", "partial")
        check(lib.vxm_batch_run(self.handle, ptr(self.params), ptr(z0_t), ptr(ts_t), T,
                                (0 if method == "be" else 1) | (2 if predict else 0)
                                | (4 if reuse else 0) | (8 if lu == "band" else 0),
                                int(bool(do_dc)), float(t_dc),
                                C.byref(opts), int(precision), ptr(out_z), ptr(conv),
                                ptr(iters), stream_ptr(stream)), "vxm_batch_run")
        if reuse:  # bits 16..30: LU factorizations
            self.last_factorizations = iters >> 16
            iters = iters & 0xFFFF
        else:
            self.last_factorizations = iters
        return out_z, conv, iters

    def band_patterns(self):
        """Per topology of the batch (`topos`): 1 if it has a band numbering."""
        out = np.zeros(len(self.topos), np.int32)
        check(lib.vxm_batch_band(self.handle, out.ctypes.data), "vxm_batch_band")
        return out

    def timings(self):
        """Per tile class (`dense_classes`) launch time of the last run, ms."""
        ms = np.zeros(len(dense_classes()[0]), np.float32)
        check(lib.vxm_batch_timings(self.handle, ms.ctypes.data), "vxm_batch_timings")
        return ms

    def split(self, out_z, T):
        """Per-instance (T, S_i) numpy views of a flat output."""
        z = out_z.cpu().numpy()
        return [z[o * T:(o + s) * T].reshape(T, s) for o, s in zip(self.state_off, self.sizes)]

    def __del__(self):
        h = getattr(self, "handle", None)
        try:
            if h is not None and h.value:
                lib.vxm_batch_destroy(h)
        except Exception:  # interpreter shutdown
            pass
