"""ctypes binding of the compiled CPU restatement of the dense path
(`oracle/cpu_dense.cpp`, built by `oracle/Makefile` into libvxm_cpu.so).

Test infrastructure: the timed CPU baseline of the benchmark and its parity
tests against the numpy oracle.  Inputs are device tables in the layout
`voltax_amd.Circuit.device_table()` returns."""

from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB_PATH = HERE / "libvxm_cpu.so"
_lib = None


def build():
    """Compile libvxm_cpu.so (g++, seconds)."""
    subprocess.run(["make", "-s", "-B", "-C", str(HERE)], check=True)


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        _lib = C.CDLL(str(LIB_PATH))
        _lib.vxm_cpu_transient_be.restype = C.c_int
        _lib.vxm_cpu_transient_be.argtypes = (
            [C.c_int] * 3 + [C.c_void_p] * 7 + [C.c_int] * 2 + [C.c_double] * 4
            + [C.c_void_p] * 3)
    return _lib


def transient_be(size, n_nodes, table, z0, ts, max_steps=100, rtol=1e-6, atol=1e-9,
                 max_dv=1.0, gmin=1e-12):
    """Backward-Euler transient from z0 over ts: (z [T, S], iters [T], conv [T])."""
    kind, wave, unk, poff, row = (np.ascontiguousarray(a) for a in table)
    kind, wave, poff = (np.ascontiguousarray(a, np.int32) for a in (kind, wave, poff))
    unk = np.ascontiguousarray(unk, np.int32).reshape(-1, 4)
    row = np.ascontiguousarray(row, np.float64)
    ts = np.ascontiguousarray(ts, np.float64)
    z0 = np.ascontiguousarray(z0, np.float64)
    T = len(ts)
    z = np.zeros((T, size))
    iters = np.zeros(T, np.int32)
    conv = np.zeros(T, np.int8)
    st = lib().vxm_cpu_transient_be(size, n_nodes, len(kind), kind.ctypes.data,
                                    wave.ctypes.data, unk.ctypes.data, poff.ctypes.data,
                                    row.ctypes.data, z0.ctypes.data, ts.ctypes.data, T,
                                    max_steps, rtol, atol, max_dv, gmin, z.ctypes.data,
                                    iters.ctypes.data, conv.ctypes.data)
    if st != 0:
        raise ValueError(f"vxm_cpu_transient_be: status {st} (unsupported device or waveform)")
    return z, iters, conv.astype(bool)
