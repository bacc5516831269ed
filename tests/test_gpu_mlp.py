"""GPU checks of the learned warm start's BF16-MFMA kernels: the GEMM and its
epilogues against numpy on the same bf16 operands, the small kernels, and the
network's forward pass, gradients and Adam step against the oracle
(oracle/mlp.py, itself pinned by finite differences)."""

import numpy as np
import pytest

from oracle import mlp as om

pytestmark = pytest.mark.gpu


def _bf16_tensor(a):
    import torch

    return torch.from_numpy(np.asarray(a, np.float32)).to(torch.bfloat16).cuda()


@pytest.mark.parametrize("m,n,k", [(64, 64, 32), (128, 192, 96), (192, 64, 4096),
                                   (2048, 2048, 64)])
def test_gemm_bf16_matches_numpy(m, n, k):
    import torch

    from voltax_amd import warmstart as ws

    rng = np.random.default_rng(m + n + k)
    A = om.bf16(rng.normal(0, 1, (m, k)))
    B = om.bf16(rng.normal(0, 1, (n, k)))
    bias = rng.normal(0, 1, n).astype(np.float32)
    ref = A @ B.T
    a, b = _bf16_tensor(A), _bf16_tensor(B)
    c = torch.empty((m, n), dtype=torch.float32, device="cuda")
    ws.gemm(m, n, k, a, k, b, k, c=c, ldc=n)
    tol = dict(rtol=1e-5, atol=1e-6 * k)
    np.testing.assert_allclose(c.cpu().numpy(), ref, **tol)
    # alpha, bias, ReLU, fp32 and bf16 outputs together
    cb = torch.empty((m, n), dtype=torch.bfloat16, device="cuda")
    bt = torch.from_numpy(bias).cuda()
    ws.gemm(m, n, k, a, k, b, k, alpha=0.5, bias=bt, relu=True, c=c, ldc=n, cb=cb, ldcb=n)
    want = np.maximum(0.5 * ref + bias, 0)
    got = c.cpu().numpy()
    np.testing.assert_allclose(got, want, **tol)
    np.testing.assert_array_equal(cb.float().cpu().numpy(), om.bf16(got))
    # ReLU-derivative mask from a bf16 matrix
    Mk = om.bf16(rng.normal(0, 1, (m, n)))
    ws.gemm(m, n, k, a, k, b, k, mask=_bf16_tensor(Mk), ldm=n, c=c, ldc=n)
    np.testing.assert_allclose(c.cpu().numpy(), ref * (Mk > 0), **tol)
    # accumulation into the fp32 output
    ws.gemm(m, n, k, a, k, b, k, beta=-2.0, c=c, ldc=n)
    np.testing.assert_allclose(c.cpu().numpy(), ref - 2.0 * ref * (Mk > 0), **tol)


def test_gemm_strided_operands_and_invalid_shapes():
    import torch

    from voltax_amd import warmstart as ws
    from voltax_amd._lib import VXMError

    rng = np.random.default_rng(5)
    A = om.bf16(rng.normal(0, 1, (64, 160)))
    B = om.bf16(rng.normal(0, 1, (128, 160)))
    a, b = _bf16_tensor(A), _bf16_tensor(B)
    c = torch.zeros((64, 136), dtype=torch.float32, device="cuda")
    # contract over the first 96 columns only, output pitch 136
    ws.gemm(64, 128, 96, a, 160, b, 160, c=c, ldc=136)
    np.testing.assert_allclose(c[:, :128].cpu().numpy(), A[:, :96] @ B[:, :96].T,
                               rtol=1e-5, atol=1e-4)
    for bad in [(63, 128, 96), (64, 100, 96), (64, 128, 40)]:
        with pytest.raises(VXMError):
            ws.gemm(*bad, a, 160, b, 160, c=c, ldc=136)


def test_transpose_rowsum_normalize_denormalize():
    import torch

    from voltax_amd import warmstart as ws
    from voltax_amd._lib import check, lib, ptr, stream_ptr

    rng = np.random.default_rng(2)
    X = om.bf16(rng.normal(0, 1, (100, 77)))
    src = _bf16_tensor(X)
    dst = torch.zeros((77, 128), dtype=torch.bfloat16, device="cuda")
    ws.transpose(100, 77, src, 77, dst, 128)
    np.testing.assert_array_equal(dst[:, :100].float().cpu().numpy(), X.T)
    out = torch.zeros(100, dtype=torch.float32, device="cuda")
    check(lib.vxm_rowsum_bf16(100, 77, ptr(src), 77, ptr(out), stream_ptr()))
    np.testing.assert_allclose(out.cpu().numpy(), X.sum(1), rtol=1e-5, atol=1e-5)

    x = rng.normal(3, 2, (50, 30))
    mu, sd = x.mean(0), x.std(0)
    xd = torch.from_numpy(x).cuda()
    d = dict(dtype=torch.float64, device="cuda")
    mu_t, inv_t, sd_t = (torch.as_tensor(v, **d) for v in (mu, 1 / sd, sd))
    ob = torch.zeros((50, 64), dtype=torch.bfloat16, device="cuda")
    of = torch.full((50, 64), 7.0, dtype=torch.float32, device="cuda")
    check(lib.vxm_normalize(50, 30, 64, ptr(xd), 30, ptr(mu_t), ptr(inv_t), ptr(ob), ptr(of), 64,
                            stream_ptr()))
    xn = (x - mu) / sd
    np.testing.assert_allclose(of[:, :30].cpu().numpy(), xn, rtol=1e-6, atol=1e-6)
    assert not of[:, 30:].any()
    np.testing.assert_array_equal(ob[:, :30].float().cpu().numpy(), om.bf16(of[:, :30].cpu().numpy()))
    back = torch.zeros((50, 30), **d)
    check(lib.vxm_denormalize(50, 30, ptr(of), 64, ptr(mu_t), ptr(sd_t), ptr(back), 30,
                              stream_ptr()))
    np.testing.assert_allclose(back.cpu().numpy(), x, rtol=1e-6, atol=1e-5)


def _mlp_and_batch(M=128, d_in=100, d_out=70, width=128, seed=0, skip=False):
    import torch

    from voltax_amd.warmstart import MLP

    net = MLP(d_in, d_out, width=width, layers=4, seed=seed, lr=1e-3, skip=skip)
    if skip:  # a nonzero skip so that it takes part in the forward pass
        S = np.random.default_rng(seed + 2).normal(0, 0.05, net.S.shape).astype(np.float32)
        S[d_out:] = 0
        S[:, d_in:] = 0
        net.S.copy_(torch.from_numpy(S))
        net.Sb.copy_(net.S.to(torch.bfloat16))
    rng = np.random.default_rng(seed + 1)
    X = rng.normal(0, 1, (M, d_in))
    T = rng.normal(0, 1, (M, d_out))
    return net, X, T, torch.from_numpy(X).cuda(), torch.from_numpy(T).cuda()


@pytest.mark.parametrize("M,skip", [(128, False), (100, False), (100, True)])
def test_mlp_forward_and_gradients_match_oracle(M, skip):
    """Forward output and every weight/bias gradient of a 4-layer network (and
    of its linear skip) equal the oracle evaluated with bf16 rounding at the
    same points (differences: fp32 vs FP64 accumulation and rare one-ulp bf16
    flips)."""
    net, X, T, Xd, Td = _mlp_and_batch(M, skip=skip)
    Mp = net._ensure(M)
    net.stage_inputs(Xd, Mp)
    net.stage_targets(Td, Mp)
    net.forward(Mp)
    net.backward(Mp, M)
    Ws, bs = net.weights()
    Ws64 = [w.astype(np.float64) for w in Ws]
    bs64 = [b.astype(np.float64) for b in bs]
    S = net.skip_weights().astype(np.float64) if skip else None
    _, Y = om.forward(Ws64, bs64, X, quant=True, S=S)
    Yg = net.Y[:M, : net.dims[-1]].cpu().numpy()
    assert np.linalg.norm(Yg - Y) <= 1e-3 * np.linalg.norm(Y)
    out = om.backward(Ws64, bs64, X, T, quant=True, S=S)
    loss, dWs, dbs = out[:3]
    if skip:
        dSg = net.dS[: net.dims[-1], : net.dims[0]].cpu().numpy()
        assert np.linalg.norm(dSg - out[3]) <= 1e-2 * np.linalg.norm(out[3])
    lossg = float(net.loss_rows[:M].sum()) / (M * net.dims[-1])
    assert abs(lossg - loss) <= 1e-4 * loss
    for l in range(net.L):
        dWg = net.dW[l][: net.dims[l + 1], : net.dims[l]].cpu().numpy()
        dbg = net.db[l][: net.dims[l + 1]].cpu().numpy()
        assert np.linalg.norm(dWg - dWs[l]) <= 1e-2 * np.linalg.norm(dWs[l]), l
        assert np.linalg.norm(dbg - dbs[l]) <= 1e-2 * np.linalg.norm(dbs[l]), l
        # padding stays inert
        assert not net.dW[l][net.dims[l + 1]:].any()


@pytest.mark.parametrize("wd", [0.0, 0.1])
def test_mlp_adam_step_matches_oracle(wd):
    net, X, T, Xd, Td = _mlp_and_batch()
    net.wd = wd
    W0 = [w.cpu().numpy().astype(np.float64) for w in net.W]
    net.train_batch(Xd, Td)
    for l in range(net.L):
        g = net.dW[l].cpu().numpy().astype(np.float64)
        w1, _, _ = om.adam(W0[l], g, 0 * g, 0 * g, 1, net.lr, wd=wd)
        np.testing.assert_allclose(net.W[l].cpu().numpy(), w1, rtol=1e-5, atol=1e-7)
        np.testing.assert_array_equal(net.Wb[l].float().cpu().numpy(),
                                      om.bf16(net.W[l].cpu().numpy()))
        np.testing.assert_array_equal(net.WbT[l].float().cpu().numpy(),
                                      om.bf16(net.W[l].cpu().numpy()).T)


def test_mlp_training_fits_a_smooth_map():
    """Training on the GPU drives the error of a smooth 16 -> 8 map down by
    more than 20x; prediction returns FP64 in the original units."""
    import torch

    from voltax_amd.warmstart import MLP

    rng = np.random.default_rng(7)
    A = rng.normal(0, 1, (16, 8)) / 4
    X = rng.uniform(-1, 1, (4096, 16))
    Z = np.tanh(X @ A) * 3 + 10
    Xd, Zd = torch.from_numpy(X).cuda(), torch.from_numpy(Z).cuda()
    net = MLP(16, 8, width=128, layers=4, seed=3, lr=3e-3)
    losses = net.fit(Xd, Zd, epochs=60, batch=256, seed=0, lr_decay="cosine")
    assert losses[-1] < losses[0] / 20
    pred = net.predict(Xd[:500]).cpu().numpy()
    err = np.sqrt(np.mean((pred - Z[:500]) ** 2))
    # predicting the mean would give err = std
    assert err < 0.15 * Z.std()


def test_mlp_training_run_follows_the_oracle():
    """Twelve Adam steps on the GPU (bf16 MFMA forward/backward, fp32 Adam)
    against the same run of the oracle (FP64 with bf16 rounding at the same
    points, same initial weights, batches and learning-rate schedule): the
    per-epoch losses and the final weights agree."""
    import torch

    from voltax_amd.warmstart import MLP

    rng = np.random.default_rng(11)
    n, d_in, d_out = 512, 48, 24
    X = rng.normal(0, 1, (n, d_in))
    Z = np.sin(X[:, :d_out]) + 0.1 * X[:, d_out:2 * d_out] ** 2
    net = MLP(d_in, d_out, width=64, layers=4, seed=4, lr=2e-3, weight_decay=0.01)
    Ws, bs = (list(map(lambda a: a.astype(np.float64), v)) for v in net.weights())
    losses = net.fit(torch.from_numpy(X).cuda(), torch.from_numpy(Z).cuda(), epochs=3,
                     batch=128, seed=0, lr_decay="cosine")
    mu_x, isd_x = net.mu_x.cpu().numpy(), net.inv_sd_x.cpu().numpy()
    mu_y, isd_y = net.mu_y.cpu().numpy(), net.inv_sd_y.cpu().numpy()
    Xn = (X - mu_x) * isd_x
    Yn = ((Z - mu_y) * isd_y).astype(np.float32).astype(np.float64)
    mW = [0 * w for w in Ws]
    vW = [0 * w for w in Ws]
    mb = [0 * b for b in bs]
    vb = [0 * b for b in bs]
    g = torch.Generator(device="cpu")
    g.manual_seed(0)
    total, it, want = 3 * (n // 128), 0, []
    for ep in range(3):
        perm = torch.randperm(n, generator=g).numpy()
        acc = 0.0
        for s in range(0, n, 128):
            idx = perm[s:s + 128]
            lr = 2e-3 * 0.5 * (1 + np.cos(np.pi * it / total))
            it += 1
            loss, dW, db = om.backward(Ws, bs, Xn[idx], Yn[idx], quant=True)
            acc += loss
            for l in range(4):
                Ws[l], mW[l], vW[l] = om.adam(Ws[l], dW[l], mW[l], vW[l], it, lr, wd=0.01)
                bs[l], mb[l], vb[l] = om.adam(bs[l], db[l], mb[l], vb[l], it, lr)
        want.append(acc / (n // 128))
    np.testing.assert_allclose(losses, want, rtol=2e-3)
    got_W, got_b = net.weights()
    for l in range(4):
        assert np.linalg.norm(got_W[l] - Ws[l]) <= 2e-3 * np.linalg.norm(Ws[l]), l
        assert np.linalg.norm(got_b[l] - bs[l]) <= 2e-2 * np.linalg.norm(bs[l]) + 1e-6, l
