"""CPU tests of the product's host logic and of the C-ABI library (no GPU
compute): symbol exports, circuit layout / group naming vs the oracle,
and the C++ host pattern builder vs the oracle's CSR structure."""

import ctypes
import re
from pathlib import Path

import numpy as np
import pytest

import circuits as tc
from oracle import circuit as oc
from oracle import sparse as osp

ROOT = Path(__file__).resolve().parents[1]


def declared_symbols():
    names = set()
    for h in (ROOT / "include").glob("*.h"):
        text = h.read_text()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(vxm_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from voltax_amd import _lib

    lib = ctypes.CDLL(str(_lib.LIB_PATH))
    names = declared_symbols()
    assert "vxm_stamp" in names and "vxm_batch_run" in names
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # every exported entry point has a Python signature
    assert set(names) <= set(_lib.SIGNATURES)


def test_status_strings():
    from voltax_amd._lib import lib

    assert lib.vxm_status_string(0) == b"ok"
    assert b"invalid" in lib.vxm_status_string(-1)


def test_invalid_arguments_are_rejected_without_gpu():
    from voltax_amd._lib import lib

    h = ctypes.c_void_p()
    kind = np.array([9], np.int32)  # unknown device kind
    z = np.zeros(4, np.int32)
    assert lib.vxm_topology_create(2, 2, 1, 1, kind.ctypes.data, z.ctypes.data,
                                   z.ctypes.data, z.ctypes.data, ctypes.byref(h)) == -1
    bad = np.array([0, 5, -1, -1], np.int32)  # unknown index out of range
    k0 = np.array([0], np.int32)
    assert lib.vxm_topology_create(2, 2, 1, 1, k0.ctypes.data, z.ctypes.data,
                                   bad.ctypes.data, z.ctypes.data, ctypes.byref(h)) == -1
    assert lib.vxm_stamp(None, 1, None, None, 0, 0, 0, 0, None, None, None, None, None) == -1


def _pair(fn, *a):
    import voltax_amd as vx

    return fn(vx.CircuitBuilder(), *a), fn(oc.CircuitBuilder(), *a)


@pytest.mark.parametrize("which", ["rc", "chain", "ring", "grid"])
def test_layout_and_pattern_match_oracle(which):
    import voltax_amd as vx
    from voltax_amd import sparse

    if which == "rc":
        c, o = _pair(tc.rc_ladder)
    elif which == "chain":
        c = tc.inverter_chain(vx.CircuitBuilder(), vx.Pulse(0, 1.2, 1e-10, 2e-11, 2e-11, 2e-9, 4e-9))
        o = tc.inverter_chain(oc.CircuitBuilder(), oc.Pulse(0, 1.2, 1e-10, 2e-11, 2e-11, 2e-9, 4e-9))
    elif which == "ring":
        c, o = tc.ring(vx.CircuitBuilder(), 7)[0], tc.ring(oc.CircuitBuilder(), 7)[0]
    else:
        c, o = _pair(tc.grid, 8)
    assert c.layout.node_names == o.layout.node_names
    assert c.layout.groups == o.layout.groups
    assert c.size == o.size
    st, ost = sparse.structure(c), osp.structure(o)
    assert np.array_equal(st.indptr, ost.indptr)
    assert np.array_equal(st.indices, ost.indices)
    assert np.array_equal(st.diagonal, ost.diagonal)


def test_device_table_parameter_rows():
    import voltax_amd as vx

    c = tc.inverter_chain(vx.CircuitBuilder(), vx.Pulse(0, 1.2, 1e-10, 2e-11, 2e-11, 2e-9, 4e-9))
    kind, wave, unk, poff, row = c.device_table()
    # groups sorted: Capacitor, EKVMOSFET_n, EKVMOSFET_p, VoltageSource, VoltageSource_Pulse
    assert list(kind) == [1] * 3 + [4] * 6 + [2, 2]
    assert list(wave[-2:]) == [1, 3]
    d = 3  # first NMOS
    p = row[poff[d]: poff[d] + 11]
    assert np.isclose(np.exp(p[0]), 1e-6) and p[3] == 1.0 and np.isclose(p[4], 400e-6)
    pm = row[poff[6]: poff[6] + 11]
    assert pm[3] == -1.0 and np.isclose(pm[4], 200e-6) and np.isclose(np.exp(pm[0]), 2e-6)
    pulse = row[poff[-1]: poff[-1] + 7]
    np.testing.assert_allclose(pulse, [0, 1.2, 1e-10, 2e-11, 2e-11, 2e-9, 4e-9])
