"""GPU parity of the U-Net kernels (MFMA implicit-GEMM convolutions, pooling,
upsampling, loss gradient, weight gradients, Adam) against the FP64 oracle,
and of the inference forward pass (upsampling read inside the convolutions)
against the training one.

The oracle rounds activations to BF16 where the kernels store BF16, so the
forward pass is compared at a tolerance set by FP32-vs-FP64 accumulation and
the occasional flipped BF16 rounding; the weight gradients pass through BF16
activation gradients and are compared per layer at 3e-2 relative L2."""

from __future__ import annotations

import numpy as np
import pytest

from oracle import unet_ref as U

pytestmark = pytest.mark.gpu


def _case(shape, seed):
    rng = np.random.default_rng(seed)
    code = rng.choice([0, 1, 2, 3, 4], size=shape, p=[0.1, 0.1, 0.5, 0.2, 0.1]).astype(np.uint8)
    r = np.where(code >= 2, rng.standard_normal(shape), 0.0).astype(np.float32)
    e = np.where(code >= 2, rng.standard_normal(shape), 0.0).astype(np.float32)
    return code, r, e


@pytest.mark.parametrize("shape,seed", [((20, 18, 14), 1), ((11, 13, 9), 2), ((33, 30, 21), 3)])
def test_forward_backward_match_oracle(torch_gpu, shape, seed):
    torch = torch_gpu
    from boilsim.unet import UNet, init_params

    code, r, e = _case(shape, seed)
    params = init_params(seed)
    params[U.num_params() - 1] = 0.05
    net = UNet(params)
    n_fluid = int((code >= 2).sum())
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    r_d, e_d, code_d = dev(r), dev(e), dev(code)
    z_d = torch.zeros(shape, dtype=torch.float32, device="cuda")
    ws = torch.zeros(UNet.workspace_bytes(shape, training=True), dtype=torch.uint8, device="cuda")
    net.forward(r_d, code_d, n_fluid, z_d, ws, training=True)
    grads = torch.zeros(U.num_params(), dtype=torch.float32, device="cuda")
    loss2 = torch.zeros(2, dtype=torch.float64, device="cuda")
    net.backward(r_d, e_d, code_d, grads, loss2, ws)
    torch.cuda.synchronize()
    z = z_d.cpu().numpy()
    zo, cache = U.forward(params.astype(np.float64), r, code, bf16=True)
    zx, _ = U.forward(params.astype(np.float64), r, code, bf16=False)
    rel = np.linalg.norm(z - zo) / np.linalg.norm(zo)
    rel_exact = np.linalg.norm(z - zx) / np.linalg.norm(zx)
    print(f"{shape}: forward rel L2 vs bf16 oracle {rel:.2e}, vs exact FP64 network {rel_exact:.2e}")
    assert rel < 1e-2
    assert np.all(z[code < 2] == 0)
    # loss
    l_num, er = loss2.cpu().numpy()
    want_loss = U.loss(z.astype(np.float64), r, e, code)
    assert abs(l_num / er - want_loss) <= 1e-4 * abs(want_loss)
    # gradients, per layer
    g = grads.cpu().numpy().astype(np.float64)
    go = U.backward(cache, r, e)
    o = 0
    for li, (ci, co) in enumerate(U.CONVS):
        n = 27 * ci * co + co
        a, b = g[o:o + n], go[o:o + n]
        o += n
        err = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        print(f"  conv{li + 1} ({ci}->{co}) grad rel L2 {err:.2e}")
        assert err < 3e-2, (li, err)


@pytest.mark.parametrize("shape,seed", [((20, 18, 14), 4), ((11, 13, 9), 5), ((110, 110, 70), 6)])
def test_inference_forward_matches_training_forward(torch_gpu, shape, seed):
    """Inference reads the decoder's upsampled inputs on the fly inside the
    convolutions' staging, forms the features of r while staging the first
    convolution's input and pools inside the two encoder convolutions;
    training writes the feature, upsampled and pooled tensors with their own
    kernels (its backward needs them).  The same values and float operations,
    so the two forward passes agree bit for bit, including odd grid sizes
    (coarse cells with missing children); with the oracle comparison of the
    training forward above this pins the inference path too."""
    torch = torch_gpu
    from boilsim.unet import UNet, init_params

    code, r, _ = _case(shape, seed)
    params = init_params(seed)
    params[U.num_params() - 1] = 0.05
    net = UNet(params)
    n_fluid = int((code >= 2).sum())
    dev = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    r_d, code_d = dev(r), dev(code)
    out = {}
    for training in (True, False):
        z_d = torch.full(shape, np.nan, dtype=torch.float32, device="cuda")
        ws = torch.zeros(UNet.workspace_bytes(shape, training=training), dtype=torch.uint8, device="cuda")
        net.forward(r_d, code_d, n_fluid, z_d, ws, training=training)
        torch.cuda.synchronize()
        out[training] = z_d.cpu().numpy()
    assert np.isfinite(out[False]).all()
    assert np.array_equal(out[True].view(np.uint32), out[False].view(np.uint32))


def test_adam_step_matches_numpy(torch_gpu):
    torch = torch_gpu
    from sim_core import _lib

    n = U.num_params()
    rng = np.random.default_rng(0)
    w = rng.standard_normal(n).astype(np.float32)
    g = rng.standard_normal(n).astype(np.float32)
    m = np.zeros(n, np.float32)
    v = np.zeros(n, np.float32)
    wd, gd, md, vd = (torch.from_numpy(a.copy()).cuda() for a in (w, g, m, v))
    for step in (1, 2):
        _lib.call("bs_unet_adam", wd.data_ptr(), gd.data_ptr(), md.data_ptr(), vd.data_ptr(), step, 1e-3, 0.5, 0)
        gg = g * 0.5
        m = 0.9 * m + 0.1 * gg
        v = 0.999 * v + 0.001 * gg * gg
        w = w - 1e-3 * (m / (1 - 0.9 ** step)) / (np.sqrt(v / (1 - 0.999 ** step)) + 1e-8)
    torch.cuda.synchronize()
    np.testing.assert_allclose(wd.cpu().numpy(), w, rtol=1e-5, atol=1e-6)


def test_build_codes_match_oracle_diag(torch_gpu):
    torch = torch_gpu
    import oracle
    from boilsim.unet import build_codes

    rng = np.random.default_rng(4)
    mat = rng.choice([0, 1, 2, 3], size=(9, 10, 11), p=[0.6, 0.15, 0.15, 0.1]).astype(np.int32)
    code = build_codes(torch.from_numpy(mat).cuda()).cpu().numpy()
    d = oracle.diag(mat)
    want = np.where(mat == 0, 8 - np.where(d > 0, d, 0).astype(np.int64), np.where(mat == 2, 0, 1))
    assert np.array_equal(code.astype(np.int64), want)
