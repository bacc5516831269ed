"""ctypes binding of ``libvoltax_mi355x.so`` (the C ABI of
include/voltax_mi355x.h).

The product has no CPU fallback: if the library is missing this module
raises at import, and every GPU entry point requires a HIP device.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libvoltax_mi355x.so"

VXM_OK = 0


class NewtonOptions(C.Structure):
    _fields_ = [
        ("max_steps", C.c_int32),
        ("rtol", C.c_double),
        ("atol", C.c_double),
        ("max_dv", C.c_double),
        ("gmin", C.c_double),
        ("gmin_steps", C.c_int32),
        ("gmin_start", C.c_double),
        ("direct_steps", C.c_int32),
        ("band_max_mult", C.c_double),
    ]


_P = C.c_void_p
_I = C.c_int32
_D = C.c_double
_F = C.c_float

# name -> argtypes (restype is always int32 unless listed in _RESTYPE)
SIGNATURES = {
    "vxm_status_string": [C.c_int],
    "vxm_topology_create": [_I, _I, _I, _I, _P, _P, _P, _P, C.POINTER(_P)],
    "vxm_topology_create_wide": [_I, _I, _I, _I, _I, _P, _P, _P, _P, C.POINTER(_P)],
    "vxm_topology_set_code": [_P, _I, _P],
    "vxm_topology_destroy": [_P],
    "vxm_topology_sizes": [_P, _P],
    "vxm_topology_pattern": [_P, _P, _P],
    "vxm_stamp": [_P, _I, _P, _P, _D, _D, _D, _D, _P, _P, _P, _P, _P],
    "vxm_stamp_noalloc": [_P, _I, _P, _P, _D, _D, _D, _D, _P, _P, _P, _P],
    "vxm_spmv": [_P, _I, _P, _P, _P, _P],
    "vxm_dense_solve": [_I, _I, _P, _P, _P, _I, _P],
    "vxm_stamp_dev": [_P, _I, _P, _P, _P, _P, _P, _P, _P, _P],
    "vxm_param_vjp": [_P, _I, _P, _P, _D, _D, _D, _P, _D, _P, _P],
    "vxm_jt_vec": [_P, _I, _P, _P, _D, _D, _D, _D, _P, _P, _P, _P],
    "vxm_batch_create": [_I, _P, _I, _P, _P, C.POINTER(_P)],
    "vxm_batch_destroy": [_P],
    "vxm_batch_timings": [_P, _P],
    "vxm_batch_band": [_P, _P],
    "vxm_time_vjp": [_P, _I, _P, C.c_double, C.c_double, _P, C.c_double, _P, _P],
    "vxm_topology_band": [_P, _I, _P],
    "vxm_batch_run": [_P, _P, _P, _P, _I, _I, _I, _D, C.POINTER(NewtonOptions), _I, _P, _P,
                      _P, _P],
    "vxm_splu_create": [_P, _I, C.POINTER(_P)],
    "vxm_splu_create_ex": [_P, _I, _I, C.POINTER(_P)],
    "vxm_splu_factor_solve": [_P, _I, _P, _P, _P, _P, _P, _P],
    "vxm_splu_destroy": [_P],
    "vxm_splu_sizes": [_P, _P],
    "vxm_splu_stats": [_P, _P],
    "vxm_splu_structure": [_P, _P, _P, _P],
    "vxm_splu_export": [_P, _P, _P, _P, _P, _P, _P, _P, _P, _P, _P],
    "vxm_splu_configure": [_P, _I],
    "vxm_dense_classes": [_P, _P, _P],
    "vxm_splu_factor": [_P, _I, _P, _P, _P, _I, _P],
    "vxm_splu_solve": [_P, _I, _P, _P, _P, _P, _I, _I, _P],
    "vxm_splu_solve_refined": [_P, _I, _P, _P, _P, _P, _P, _P, _I, _I, _I, _P],
    "vxm_newton_update": [_I, _I, _I, _P, _P, _D, C.POINTER(NewtonOptions), _P, _P, _P, _P],
    "vxm_newton_update_reuse": [_I, _I, _I, _P, _P, _D, C.POINTER(NewtonOptions), _P, _P, _P,
                                _P, _I, _P],
    "vxm_ac_system": [_I, _I, _I, _P, _P, _P, _P, _P, _P, _P],
    "vxm_gemm_bf16": [_I, _I, _I, _P, _I, _P, _I, _F, _F, _P, _P, _I, _I, _P, _I, _P, _I, _P],
    "vxm_transpose_bf16": [_I, _I, _P, _I, _P, _I, _P],
    "vxm_rowsum_bf16": [_I, _I, _P, _I, _P, _P],
    "vxm_mse_grad": [_I, _I, _I, _I, _P, _I, _P, _I, _F, _P, _I, _P, _P],
    "vxm_adam": [C.c_int64, _P, _P, _P, _P, _F, _F, _F, _F, _F, _I, _P, _P],
    "vxm_normalize": [_I, _I, _I, _P, _I, _P, _P, _P, _P, _I, _P],
    "vxm_denormalize": [_I, _I, _P, _I, _P, _P, _P, _I, _P],
}
_RESTYPE = {"vxm_status_string": C.c_char_p}


class VXMError(RuntimeError):
    pass


def _load():
    if not LIB_PATH.exists():
        raise ImportError(
            f"{LIB_PATH} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` "
            "or `python voltax_amd/_build.py` (no CPU fallback exists)")
    lib = C.CDLL(str(LIB_PATH))
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPE.get(name, C.c_int32)
    return lib


lib = _load()


def check(status: int, what: str = "") -> None:
    if status != VXM_OK:
        msg = lib.vxm_status_string(status).decode()
        raise VXMError(f"{what or 'voltax_mi355x'} failed: {msg} ({status})")


def ptr(t) -> int:
    """Device (or host numpy) pointer of a torch tensor / numpy array."""
    if t is None:
        return None
    if hasattr(t, "data_ptr"):
        return t.data_ptr()
    return t.ctypes.data


def stream_ptr(stream=None) -> int:
    import torch

    s = stream if stream is not None else torch.cuda.current_stream()
    return s.cuda_stream
