"""Pins the U-Net oracle (oracle/unet_ref.py): forward and gradients against
torch's float64 conv3d/autograd (an independent implementation), and the
hand-written backward against central finite differences."""

from __future__ import annotations

import numpy as np
import pytest
import torch
import torch.nn.functional as F

from oracle import unet_ref as U


def _case(shape, seed):
    rng = np.random.default_rng(seed)
    code = rng.choice([0, 1, 2, 3, 4], size=shape, p=[0.15, 0.15, 0.4, 0.2, 0.1]).astype(np.uint8)
    r = np.where(code >= 2, rng.standard_normal(shape), 0.0).astype(np.float32)
    e = np.where(code >= 2, rng.standard_normal(shape), 0.0).astype(np.float32)
    return code, r, e


def _torch_forward(params, r, code):
    """The same network written with torch ops; returns z (float64 tensor)."""
    P = U.split_params(params)
    Wt = [torch.tensor(w, dtype=torch.float64, requires_grad=True) for w, _ in P]
    Bt = [torch.tensor(b, dtype=torch.float64, requires_grad=True) for _, b in P]

    def conv(x, li):
        w = Wt[li].permute(2, 1, 0).reshape(Wt[li].shape[2], Wt[li].shape[1], 3, 3, 3)
        return F.conv3d(x, w, Bt[li], padding=1)

    def pool(x):
        n = x.shape[2:]
        pad = [0, n[2] % 2, 0, n[1] % 2, 0, n[0] % 2]
        xs = F.pad(x, pad)
        ones = F.pad(torch.ones_like(x[:, :1]), pad)
        s = F.avg_pool3d(xs, 2) * 8
        c = F.avg_pool3d(ones, 2) * 8
        return s / c

    def up(x, shape):
        y = x.repeat_interleave(2, 2).repeat_interleave(2, 3).repeat_interleave(2, 4)
        return y[:, :, :shape[0], :shape[1], :shape[2]]

    x0, rms = U.features(r, code)
    x = torch.tensor(x0, dtype=torch.float64).permute(3, 0, 1, 2)[None]
    h1 = F.relu(conv(x, 0))
    h2 = F.relu(conv(pool(h1), 1))
    h3 = F.relu(conv(F.relu(conv(pool(h2), 2)), 3))
    s2 = F.relu(conv(up(h3, h2.shape[2:]), 4)) + h2
    s1 = F.relu(conv(up(s2, h1.shape[2:]), 5)) + h1
    y = conv(F.relu(conv(s1, 6)), 7)[0, 0]
    fl = torch.tensor(code >= 2)
    z = torch.where(fl, y * float(np.float32(rms)), torch.zeros_like(y))
    return z, Wt, Bt


@pytest.mark.parametrize("shape", [(7, 6, 5), (8, 9, 4)])
def test_forward_and_grad_match_torch(shape):
    code, r, e = _case(shape, 3)
    params = U.init_params(1).astype(np.float64)
    params[U.num_params() - 1] = 0.3  # a non-zero final bias
    z, cache = U.forward(params, r, code)
    zt, Wt, Bt = _torch_forward(params, r, code)
    np.testing.assert_allclose(z, zt.detach().numpy(), rtol=1e-10, atol=1e-12)
    # loss and its gradient
    d = zt - torch.tensor(e, dtype=torch.float64)
    A = lambda v: torch.tensor(U.apply_A(v.detach().numpy(), code))
    # (z-e)^T A (z-e) with A linear: build it through autograd-friendly ops
    fl = torch.tensor(code >= 2)
    cp = np.pad(code, 1)
    acc = torch.zeros_like(d)
    dp = F.pad(d[None, None], (1, 1, 1, 1, 1, 1))[0, 0]
    for di, dj, dk in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
        sl = (slice(1 + di, dp.shape[0] - 1 + di), slice(1 + dj, dp.shape[1] - 1 + dj),
              slice(1 + dk, dp.shape[2] - 1 + dk))
        cn = torch.tensor(cp[sl].astype(np.int64))
        acc = acc + torch.where(cn >= 2, dp[sl], torch.where(cn == 1, d, torch.zeros_like(d)))
        del cn
    Ad = torch.where(fl, 6.0 * d - acc, torch.zeros_like(d))
    L = (d * Ad).sum() / float(np.sum(e.astype(np.float64) * r))
    assert abs(L.item() - U.loss(z, r, e, code)) < 1e-10 * max(1.0, abs(L.item()))
    L.backward()
    g = U.backward(cache, r, e)
    gt = np.concatenate([np.concatenate([w.grad.numpy().ravel(), b.grad.numpy().ravel()]) for w, b in zip(Wt, Bt)])
    np.testing.assert_allclose(g, gt, rtol=1e-8, atol=1e-10 * np.abs(gt).max())
    assert A(d).shape == d.shape


def test_backward_finite_differences():
    shape = (6, 5, 4)
    code, r, e = _case(shape, 7)
    params = U.init_params(2).astype(np.float64)
    z, cache = U.forward(params, r, code)
    g = U.backward(cache, r, e)
    rng = np.random.default_rng(0)
    idx = rng.choice(U.num_params(), size=12, replace=False)
    for i in idx:
        h = 1e-6 * max(1.0, abs(params[i]))
        pp, pm = params.copy(), params.copy()
        pp[i] += h
        pm[i] -= h
        lp = U.loss(U.forward(pp, r, code)[0], r, e, code)
        lm = U.loss(U.forward(pm, r, code)[0], r, e, code)
        fd = (lp - lm) / (2 * h)
        assert abs(fd - g[i]) <= 1e-5 * max(1e-3, abs(g[i]), abs(fd)), (i, fd, g[i])


def test_bf16_rounding_helper():
    x = np.array([1.0, 1.0 + 2**-8, 1.0 + 2**-7, 1.0 + 3 * 2**-8, -3.1415926, 65504.0], np.float32)
    got = U.to_bf16(x)
    want = torch.tensor(x).to(torch.bfloat16).to(torch.float64).numpy()
    assert np.array_equal(got, want)
