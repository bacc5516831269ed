"""Learned Newton warm start: a multilayer perceptron (ReLU hidden layers,
linear output) that maps the parameters of a circuit variant to a predicted
operating point, trained on the GPU on operating points the plain solver
computed.

The network runs entirely in the HIP kernels of ``csrc/mlp.hip``: bf16
operands on the MFMA units with fp32 accumulation, fp32 master weights updated
by Adam.  Every product is an "NT" GEMM (both operands contiguous along the
contracted axis), so the backward pass keeps transposed bf16 copies of the
weights (refreshed after each update) and transposes activations and output
gradients with a tiled kernel:

    forward   H_{l+1} = relu(H_l W_l^T + b_l)          (Y = H_{L-1} W^T + b, fp32)
    backward  dW_l = G_{l+1}^T H_l,  db_l = sum_m G_{l+1}
              G_l  = (G_{l+1} W_l) * [H_l > 0]

Inputs and targets are standardised per column; the prediction is mapped back
to FP64 volts and amperes.  The reference has no learned part (it starts
Newton from the caller's guess, voltax/analysis.py:291-317); the Newton
solve that consumes the prediction is unchanged, so every answer is still
converged to the same tolerance.
"""

from __future__ import annotations

import numpy as np
import torch

from . import engine
from ._lib import check, lib, ptr, stream_ptr


def _pad(n, q=64):
    return -(-int(n) // q) * q


def gemm(m, n, k, a, lda, b, ldb, alpha=1.0, beta=0.0, bias=None, mask=None, ldm=0,
         relu=False, c=None, ldc=0, cb=None, ldcb=0):
    """C[m, n] = alpha A[m, k] B[n, k]^T (+ bias + beta C, mask, relu) on the MFMA units."""
    check(lib.vxm_gemm_bf16(m, n, k, ptr(a), lda, ptr(b), ldb, float(alpha), float(beta),
                            ptr(bias),
                            ptr(mask), ldm, int(relu), ptr(c), ldc, ptr(cb), ldcb,
                            stream_ptr()), "vxm_gemm_bf16")


def transpose(rows, cols, src, lds, dst, ldd):
    check(lib.vxm_transpose_bf16(rows, cols, ptr(src), lds, ptr(dst), ldd, stream_ptr()),
          "vxm_transpose_bf16")


class MLP:
    """`layers` affine maps d_in -> width -> ... -> width -> d_out, plus (with
    `skip`) a linear map d_in -> d_out added to the output:
    Y = net(X) + X S^T.  The linear part carries what is linear in the
    parameters (node voltages set by sources, the grid's response to its
    sinks); the ReLU layers learn the rest."""

    def __init__(self, d_in, d_out, width=512, layers=4, seed=0, lr=1e-3, betas=(0.9, 0.999),
                 eps=1e-8, weight_decay=0.0, skip=False):
        engine.require_gpu()
        self.dims = [int(d_in)] + [int(width)] * (layers - 1) + [int(d_out)]
        self.pd = [_pad(d) for d in self.dims]
        self.L = layers
        self.lr, self.betas, self.eps, self.wd = lr, betas, eps, weight_decay
        self.step_count = 0
        rng = np.random.default_rng(seed)
        f32 = dict(dtype=torch.float32, device="cuda")
        self.W, self.b, self.Wb, self.WbT = [], [], [], []
        self.mW, self.vW, self.mb, self.vb, self.dW, self.db = [], [], [], [], [], []
        for l in range(layers):
            fi, fo = self.dims[l], self.dims[l + 1]
            # He-uniform for ReLU inputs; the output layer starts small so the
            # first predictions sit near the standardised mean
            bound = np.sqrt(6.0 / fi) if l < layers - 1 else np.sqrt(1.0 / fi)
            w = np.zeros((self.pd[l + 1], self.pd[l]), np.float32)
            w[:fo, :fi] = rng.uniform(-bound, bound, (fo, fi))
            W = torch.from_numpy(w).cuda()
            self.W.append(W)
            self.b.append(torch.zeros(self.pd[l + 1], **f32))
            for lst in (self.mW, self.vW, self.dW):
                lst.append(torch.zeros_like(W))
            for lst in (self.mb, self.vb, self.db):
                lst.append(torch.zeros(self.pd[l + 1], **f32))
            self.Wb.append(torch.empty(W.shape, dtype=torch.bfloat16, device="cuda"))
            self.WbT.append(torch.empty((W.shape[1], W.shape[0]), dtype=torch.bfloat16,
                                        device="cuda"))
        self.skip = skip
        if skip:
            # linear skip, zero at the start; trained like the other weights
            self.S = torch.zeros((self.pd[-1], self.pd[0]), **f32)
            self.Sb = torch.zeros(self.S.shape, dtype=torch.bfloat16, device="cuda")
            self.mS, self.vS, self.dS = (torch.zeros_like(self.S) for _ in range(3))
        self._refresh_bf16()
        self._M = None
        # standardisation (identity until `set_scaling`)
        self.set_scaling(np.zeros(self.dims[0]), np.ones(self.dims[0]),
                         np.zeros(self.dims[-1]), np.ones(self.dims[-1]))

    # -- parameters
    def _refresh_bf16(self):
        for l in range(self.L):
            n, k = self.W[l].shape
            # bf16 copies of the initial weights (round to nearest even, as
            # the Adam kernel rounds its updates)
            self.Wb[l].copy_(self.W[l].to(torch.bfloat16))
            transpose(n, k, self.Wb[l], k, self.WbT[l], n)

    def set_scaling(self, mu_x, sd_x, mu_y, sd_y, floor=1e-10):
        """Per-column standardisation.  Columns whose spread is below `floor`
        (in their own units: volts, amperes) are treated as constants: their
        targets are zero and their prediction is the mean, so rounding noise of
        the solver (e.g. branch currents that are exactly zero in DC) is not
        learned."""
        d = dict(dtype=torch.float64, device="cuda")
        sd_x = np.asarray(sd_x, float)
        sd_y = np.asarray(sd_y, float)
        sd_x = np.where(sd_x > floor, sd_x, 0.0)
        sd_y = np.where(sd_y > floor, sd_y, 0.0)
        self.mu_x = torch.as_tensor(np.asarray(mu_x, float), **d)
        self.inv_sd_x = torch.as_tensor(np.where(sd_x > 0, 1.0 / np.where(sd_x > 0, sd_x, 1), 0),
                                        **d)
        self.mu_y = torch.as_tensor(np.asarray(mu_y, float), **d)
        self.sd_y = torch.as_tensor(sd_y, **d)
        self.inv_sd_y = torch.as_tensor(np.where(sd_y > 0, 1.0 / np.where(sd_y > 0, sd_y, 1), 0),
                                        **d)

    def weights(self):
        """(Ws, bs) as numpy fp32 without padding (for checks)."""
        Ws = [self.W[l][: self.dims[l + 1], : self.dims[l]].cpu().numpy() for l in range(self.L)]
        bs = [self.b[l][: self.dims[l + 1]].cpu().numpy() for l in range(self.L)]
        return Ws, bs

    def skip_weights(self):
        return self.S[: self.dims[-1], : self.dims[0]].cpu().numpy() if self.skip else None

    # -- activations of a batch of M rows
    def _ensure(self, M):
        Mp = _pad(M)
        if self._M == Mp:
            return Mp
        bf = dict(dtype=torch.bfloat16, device="cuda")
        self.H = [torch.zeros((Mp, self.pd[l]), **bf) for l in range(self.L)]
        self.Y = torch.zeros((Mp, self.pd[-1]), dtype=torch.float32, device="cuda")
        self.T = torch.zeros((Mp, self.pd[-1]), dtype=torch.float32, device="cuda")
        self.G = [torch.zeros((Mp, self.pd[l + 1]), **bf) for l in range(self.L)]
        maxd = max(self.pd)
        self.GT = torch.zeros((maxd, Mp), **bf)
        self.HT = torch.zeros((maxd, Mp), **bf)
        self.XT = torch.zeros((self.pd[0], Mp), **bf) if self.skip else None
        self.loss_rows = torch.zeros(Mp, dtype=torch.float32, device="cuda")
        self._M = Mp
        return Mp

    def forward(self, Mp):
        """Y = network(H[0]) for the Mp rows staged in H[0]."""
        if self.skip:
            gemm(Mp, self.pd[-1], self.pd[0], self.H[0], self.pd[0], self.Sb, self.pd[0],
                 c=self.Y, ldc=self.pd[-1])
        for l in range(self.L):
            k, n = self.pd[l], self.pd[l + 1]
            last = l == self.L - 1
            gemm(Mp, n, k, self.H[l], k, self.Wb[l], k, bias=self.b[l], relu=not last,
                 beta=1.0 if last and self.skip else 0.0,
                 c=self.Y if last else None, ldc=n, cb=None if last else self.H[l + 1], ldcb=n)

    def backward(self, Mp, m_valid):
        """Gradients of the mean squared error over the first m_valid rows."""
        n_out = self.dims[-1]
        N = self.pd[-1]
        check(lib.vxm_mse_grad(Mp, m_valid, N, n_out, ptr(self.Y), N, ptr(self.T), N,
                               2.0 / (m_valid * n_out), ptr(self.G[-1]), N, ptr(self.loss_rows),
                               stream_ptr()), "vxm_mse_grad")
        if self.skip:
            transpose(Mp, self.pd[0], self.H[0], self.pd[0], self.XT, Mp)
        for l in range(self.L - 1, -1, -1):
            k, n = self.pd[l], self.pd[l + 1]
            G = self.G[l]
            transpose(Mp, n, G, n, self.GT, Mp)
            if l == self.L - 1 and self.skip:
                gemm(n, self.pd[0], Mp, self.GT, Mp, self.XT, Mp, c=self.dS, ldc=self.pd[0])
            if l == 0 and self.skip:
                HT = self.XT
            else:
                HT = self.HT
                transpose(Mp, k, self.H[l], k, HT, Mp)
            gemm(n, k, Mp, self.GT, Mp, HT, Mp, c=self.dW[l], ldc=k)
            check(lib.vxm_rowsum_bf16(n, Mp, ptr(self.GT), Mp, ptr(self.db[l]), stream_ptr()),
                  "vxm_rowsum_bf16")
            if l > 0:
                gemm(Mp, k, n, G, n, self.WbT[l], n, mask=self.H[l], ldm=k, cb=self.G[l - 1],
                     ldcb=k)

    def update(self):
        self.step_count += 1
        b1, b2 = self.betas
        s = stream_ptr()
        for l in range(self.L):
            check(lib.vxm_adam(self.W[l].numel(), ptr(self.W[l]), ptr(self.dW[l]),
                               ptr(self.mW[l]), ptr(self.vW[l]), self.lr, b1, b2, self.eps,
                               self.wd, self.step_count, ptr(self.Wb[l]), s), "vxm_adam")
            check(lib.vxm_adam(self.b[l].numel(), ptr(self.b[l]), ptr(self.db[l]),
                               ptr(self.mb[l]), ptr(self.vb[l]), self.lr, b1, b2, self.eps,
                               0.0, self.step_count, None, s), "vxm_adam")
            n, k = self.W[l].shape
            transpose(n, k, self.Wb[l], k, self.WbT[l], n)
        if self.skip:
            check(lib.vxm_adam(self.S.numel(), ptr(self.S), ptr(self.dS), ptr(self.mS),
                               ptr(self.vS), self.lr, b1, b2, self.eps, self.wd,
                               self.step_count, ptr(self.Sb), s), "vxm_adam")

    # -- data staging
    def stage_inputs(self, X, Mp):
        """Standardised bf16 copy of X [m, d_in] (FP64, device) into H[0]."""
        m = X.shape[0]
        check(lib.vxm_normalize(m, self.dims[0], self.pd[0], ptr(X), X.stride(0), ptr(self.mu_x),
                                ptr(self.inv_sd_x), ptr(self.H[0]), None, self.pd[0],
                                stream_ptr()), "vxm_normalize")
        if m < Mp:
            self.H[0][m:].zero_()

    def stage_targets(self, Z, Mp):
        m = Z.shape[0]
        N = self.pd[-1]
        check(lib.vxm_normalize(m, self.dims[-1], N, ptr(Z), Z.stride(0), ptr(self.mu_y),
                                ptr(self.inv_sd_y), None, ptr(self.T), N, stream_ptr()),
              "vxm_normalize")
        if m < Mp:
            self.T[m:].zero_()

    def train_batch(self, X, Z):
        """One Adam step on the batch (X [m, d_in], Z [m, d_out], FP64 device);
        returns the batch's mean squared error in standardised units (device)."""
        m = X.shape[0]
        Mp = self._ensure(m)
        self.stage_inputs(X, Mp)
        self.stage_targets(Z, Mp)
        self.forward(Mp)
        self.backward(Mp, m)
        self.update()
        return self.loss_rows[:m].sum() / (m * self.dims[-1])

    def predict(self, X, out=None):
        """FP64 predictions [m, d_out] for FP64 inputs X [m, d_in] (device)."""
        m = X.shape[0]
        if out is None:
            out = torch.empty((m, self.dims[-1]), dtype=torch.float64, device="cuda")
        Mp = self._ensure(m)
        self.stage_inputs(X, Mp)
        self.forward(Mp)
        check(lib.vxm_denormalize(m, self.dims[-1], ptr(self.Y), self.pd[-1], ptr(self.mu_y),
                                  ptr(self.sd_y), ptr(out), out.stride(0), stream_ptr()),
              "vxm_denormalize")
        return out

    def fit(self, X, Z, epochs=10, batch=256, seed=0, lr_decay=None, log=None):
        """Trains on rows of X [n, d_in] -> Z [n, d_out] (FP64, device): standardises
        both, then `epochs` passes of shuffled minibatches.  Returns per-epoch losses."""
        mu_x, sd_x = X.mean(0), X.std(0)
        mu_y, sd_y = Z.mean(0), Z.std(0)
        self.set_scaling(mu_x.cpu().numpy(), sd_x.cpu().numpy(), mu_y.cpu().numpy(),
                         sd_y.cpu().numpy())
        n = X.shape[0]
        g = torch.Generator(device="cpu")
        g.manual_seed(seed)
        losses = []
        lr0 = self.lr
        total = epochs * max(1, n // batch)
        it = 0
        for ep in range(epochs):
            perm = torch.randperm(n, generator=g).cuda()
            acc = torch.zeros((), dtype=torch.float32, device="cuda")
            nb = 0
            for s in range(0, n - batch + 1, batch):
                idx = perm[s:s + batch]
                if lr_decay == "cosine":
                    self.lr = lr0 * 0.5 * (1 + np.cos(np.pi * it / total))
                acc += self.train_batch(X.index_select(0, idx), Z.index_select(0, idx))
                nb += 1
                it += 1
            losses.append(float(acc) / max(nb, 1))
            if log:
                log(f"epoch {ep}: loss {losses[-1]:.3e}")
        self.lr = lr0
        return losses
