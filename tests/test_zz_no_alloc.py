"""vxm_stamp_noalloc (fixed device scratch, no allocation) against vxm_stamp
(caller's work buffer): the same kernels, so the Jacobian values and
residuals must agree bit for bit, on two calls in a row with the second on a
smaller batch, with and without the history term and the Jacobian."""

import ctypes as C

import numpy as np
import pytest

import circuits as tc

pytestmark = pytest.mark.gpu


def test_noalloc_stamp_equals_stamp_on_two_calls():
    import torch

    import voltax_amd as vx
    from voltax_amd import engine, sparse
    from voltax_amd._lib import lib, ptr, stream_ptr

    c = tc.ring(vx.CircuitBuilder(), 9)[0]
    topo, row = engine.topology_of(c)
    sparse.jacobian(c, c.zeros())  # makes the topology's tables resident
    dev = dict(dtype=torch.float64, device="cuda")
    rng = np.random.default_rng(3)
    print("noalloc: ring9, batches 96 then 5", flush=True)
    for B in (96, 5):
        params = np.tile(row, (B, 1))
        params[:, :] *= 1.0 + 0.01 * rng.standard_normal(params.shape) * (params != 0)
        pt = torch.as_tensor(params, **dev).contiguous()
        z = torch.as_tensor(rng.uniform(0.0, 1.2, (B, c.size)), **dev)
        hist = torch.as_tensor(rng.standard_normal((B, c.size)) * 1e-3, **dev)
        for h, want_data in ((None, True), (hist, True), (hist, False)):
            work = torch.empty(B * topo.work, **dev)
            d1 = torch.full((B, topo.nnz), np.nan, **dev) if want_data else None
            r1 = torch.full((B, topo.S), np.nan, **dev)
            d2 = torch.full((B, topo.nnz), np.nan, **dev) if want_data else None
            r2 = torch.full((B, topo.S), np.nan, **dev)
            args = (1e-10, 1.0, 5e11, 1e-12)
            assert lib.vxm_stamp(topo.handle, B, ptr(pt), ptr(z), *args, ptr(h), ptr(d1),
                                 ptr(r1), ptr(work), stream_ptr()) == 0
            assert lib.vxm_stamp_noalloc(topo.handle, B, ptr(pt), ptr(z), *args, ptr(h),
                                         ptr(d2), ptr(r2), stream_ptr()) == 0
            torch.cuda.synchronize()
            assert torch.equal(r1, r2) and bool(torch.isfinite(r2).all())
            if want_data:
                assert torch.equal(d1, d2) and bool(torch.isfinite(d2).all())
    # beyond the fixed scratch: a clear error, nothing launched
    big = 4 * 1024 * 1024 // topo.work + 1
    assert lib.vxm_stamp_noalloc(topo.handle, big, ptr(pt), ptr(z), 0.0, 1.0, 0.0, 0.0, None,
                                 None, ptr(r2), stream_ptr()) == -5
