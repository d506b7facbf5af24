"""ORACLE — test infrastructure only.

FP64 numpy restatement of the learned preconditioner's U-Net (the network the
HIP kernels in boiling-sim_amd/csrc/unet.hip implement; it has no counterpart
in the reference) with a hand-written backward pass.  Parameter layout matches
the library: for each conv, W[27][cin][cout] then bias[cout], taps ordered
t = (di+1)*9 + (dj+1)*3 + (dk+1).

``bf16=True`` rounds every stored activation (and the weights used by the
MFMA convolutions) to bfloat16 exactly where the HIP path stores BF16, so the
GPU can be compared at a tight tolerance; ``bf16=False`` is the exact network.

Pinning: tests check forward and gradients against torch's conv3d/autograd in
float64 (an independent implementation) and the gradients against central
finite differences on tiny grids.
"""

from __future__ import annotations

import numpy as np

CONVS = [(8, 8), (8, 16), (16, 32), (32, 32), (32, 16), (16, 8), (8, 8), (8, 1)]


def num_params() -> int:
    return sum(27 * ci * co + co for ci, co in CONVS)


def split_params(params):
    out = []
    o = 0
    for ci, co in CONVS:
        w = params[o:o + 27 * ci * co].reshape(27, ci, co)
        o += 27 * ci * co
        b = params[o:o + co]
        o += co
        out.append((w, b))
    return out


def to_bf16(x):
    """Round float64/float32 values to the nearest bfloat16 (ties to even)."""
    f = np.asarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    u = (u + 0x7FFF + lsb) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def _shift(x, di, dj, dk):
    """y[i,j,k] = x[i+di, j+dj, k+dk], zero outside."""
    nx, ny, nz = x.shape[:3]
    y = np.zeros_like(x)
    si = slice(max(0, -di), min(nx, nx - di))
    sj = slice(max(0, -dj), min(ny, ny - dj))
    sk = slice(max(0, -dk), min(nz, nz - dk))
    ti = slice(max(0, di), min(nx, nx + di))
    tj = slice(max(0, dj), min(ny, ny + dj))
    tk = slice(max(0, dk), min(nz, nz + dk))
    y[si, sj, sk] = x[ti, tj, tk]
    return y


TAPS = [(t // 9 - 1, (t // 3) % 3 - 1, t % 3 - 1) for t in range(27)]


def conv(x, w, b):
    """x: (nx,ny,nz,cin) -> (nx,ny,nz,cout); zero padding."""
    y = np.zeros(x.shape[:3] + (w.shape[2],))
    for t, (di, dj, dk) in enumerate(TAPS):
        y += np.einsum("ijkc,cd->ijkd", _shift(x, di, dj, dk), w[t])
    return y + b


def conv_bwd(x, w, gy):
    """Gradients of conv w.r.t. x, w, b given gy = dL/dy."""
    gx = np.zeros_like(x)
    gw = np.zeros_like(w)
    for t, (di, dj, dk) in enumerate(TAPS):
        xs = _shift(x, di, dj, dk)
        gw[t] = np.einsum("ijkc,ijkd->cd", xs, gy)
        gx += _shift(np.einsum("ijkd,cd->ijkc", gy, w[t]), -di, -dj, -dk)
    return gx, gw, gy.sum(axis=(0, 1, 2))


def coarse_shape(s):
    return tuple((v + 1) // 2 for v in s)


def pool(x):
    nx, ny, nz, c = x.shape
    cs = coarse_shape((nx, ny, nz))
    acc = np.zeros(cs + (c,))
    cnt = np.zeros(cs + (1,))
    for a in range(2):
        for b_ in range(2):
            for d in range(2):
                sub = x[a::2, b_::2, d::2]
                acc[:sub.shape[0], :sub.shape[1], :sub.shape[2]] += sub
                cnt[:sub.shape[0], :sub.shape[1], :sub.shape[2]] += 1
    return acc / cnt, cnt


def pool_bwd(gc, cnt, fine_shape):
    g = gc / cnt
    out = np.zeros(tuple(fine_shape) + (gc.shape[3],))
    for a in range(2):
        for b_ in range(2):
            for d in range(2):
                sub = out[a::2, b_::2, d::2]
                sub[...] = g[:sub.shape[0], :sub.shape[1], :sub.shape[2]]
    return out


def up(xc, fine_shape):
    i = np.arange(fine_shape[0]) // 2
    j = np.arange(fine_shape[1]) // 2
    k = np.arange(fine_shape[2]) // 2
    return xc[i][:, j][:, :, k]


def up_bwd(gf, coarse_s):
    gc = np.zeros(tuple(coarse_s) + (gf.shape[3],))
    for a in range(2):
        for b_ in range(2):
            for d in range(2):
                sub = gf[a::2, b_::2, d::2]
                gc[:sub.shape[0], :sub.shape[1], :sub.shape[2]] += sub
    return gc


def features(r, code):
    """[r/rms, fluid, air, solid, 0, 0, 0, 0]; code: 0 air, 1 solid, >=2 fluid."""
    fl = code >= 2
    nfl = max(int(fl.sum()), 1)
    rms = float(np.sqrt(np.sum(np.asarray(r, np.float64) ** 2) / nfl))
    inv = 1.0 / rms if rms > 0 else 0.0
    x = np.zeros(code.shape + (8,))
    x[..., 0] = np.where(fl, np.asarray(r, np.float32) * np.float32(inv), 0.0)
    x[..., 1] = fl
    x[..., 2] = code == 0
    x[..., 3] = code == 1
    return x, rms


def forward(params, r, code, bf16=False):
    """Returns (z, cache).  z = rms * y on fluid cells."""
    P = split_params(np.asarray(params, np.float64))
    q = to_bf16 if bf16 else (lambda a: a)
    W = [(q(w), b) for w, b in P[:7]] + [P[7]]
    relu = lambda a: np.maximum(a, 0.0)
    x0, rms = features(r, code)
    x0 = q(x0)
    s1_shape = code.shape
    h1 = q(relu(conv(x0, *W[0])))
    p1, cnt1 = pool(h1)
    p1 = q(p1)
    h2 = q(relu(conv(p1, *W[1])))
    p2, cnt2 = pool(h2)
    p2 = q(p2)
    h3a = q(relu(conv(p2, *W[2])))
    h3 = q(relu(conv(h3a, *W[3])))
    up3 = up(h3, h2.shape[:3])
    u2 = q(relu(conv(up3, *W[4])))
    s2 = q(u2 + h2)
    up2 = up(s2, s1_shape)
    u1 = q(relu(conv(up2, *W[5])))
    s1 = q(u1 + h1)
    d1 = q(relu(conv(s1, *W[6])))
    y = conv(d1, *W[7])[..., 0]
    z = np.where(code >= 2, y * np.float32(rms), 0.0)
    cache = dict(W=W, x0=x0, h1=h1, p1=p1, cnt1=cnt1, h2=h2, p2=p2, cnt2=cnt2, h3a=h3a, h3=h3, up3=up3,
                 u2=u2, s2=s2, up2=up2, u1=u1, s1=s1, d1=d1, y=y, rms=rms, code=code)
    return z, cache


def apply_A(z, code):
    """The PCG operator A = 6I - S with class codes (FP64)."""
    fl = code >= 2
    zp = np.pad(z, 1)
    cp = np.pad(code, 1)  # 0 = air outside
    out = np.zeros_like(z, dtype=np.float64)
    for di, dj, dk in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
        sl = (slice(1 + di, zp.shape[0] - 1 + di), slice(1 + dj, zp.shape[1] - 1 + dj),
              slice(1 + dk, zp.shape[2] - 1 + dk))
        cn = cp[sl]
        out += np.where(cn >= 2, zp[sl], np.where(cn == 1, z, 0.0))
    return np.where(fl, 6.0 * z - out, 0.0)


def loss(z, r, e, code):
    """A-norm loss (z-e)^T A (z-e) / (e.r); e.r = e^T A e when A e = r."""
    d = np.asarray(z, np.float64) - np.asarray(e, np.float64)
    er = np.sum(np.asarray(e, np.float64) * np.asarray(r, np.float64))
    return float(np.sum(d * apply_A(d, code)) / er)


def backward(cache, r, e):
    """Gradients (flat, library layout) of loss(forward(params, r), r, e)."""
    W = cache["W"]
    code = cache["code"]
    rms = cache["rms"]
    fl = code >= 2
    z = np.where(fl, cache["y"] * np.float32(rms), 0.0)
    er = float(np.sum(np.asarray(e, np.float64) * np.asarray(r, np.float64)))
    gz = 2.0 * apply_A(z - np.asarray(e, np.float64), code) / er
    gy = np.where(fl, gz * np.float32(rms), 0.0)[..., None]
    grads = [None] * 8
    relu_g = lambda g, a: np.where(a > 0, g, 0.0)
    g_d1, gw, gb = conv_bwd(cache["d1"], W[7][0], gy)
    grads[7] = (gw, gb)
    g_d1 = relu_g(g_d1, cache["d1"])
    g_s1, gw, gb = conv_bwd(cache["s1"], W[6][0], g_d1)
    grads[6] = (gw, gb)
    g_u1 = relu_g(g_s1, cache["u1"])
    g_up2, gw, gb = conv_bwd(cache["up2"], W[5][0], g_u1)
    grads[5] = (gw, gb)
    g_s2 = up_bwd(g_up2, cache["s2"].shape[:3])
    g_u2 = relu_g(g_s2, cache["u2"])
    g_up3, gw, gb = conv_bwd(cache["up3"], W[4][0], g_u2)
    grads[4] = (gw, gb)
    g_h3 = relu_g(up_bwd(g_up3, cache["h3"].shape[:3]), cache["h3"])
    g_h3a, gw, gb = conv_bwd(cache["h3a"], W[3][0], g_h3)
    grads[3] = (gw, gb)
    g_h3a = relu_g(g_h3a, cache["h3a"])
    g_p2, gw, gb = conv_bwd(cache["p2"], W[2][0], g_h3a)
    grads[2] = (gw, gb)
    g_h2 = relu_g(pool_bwd(g_p2, cache["cnt2"], cache["h2"].shape[:3]) + g_s2, cache["h2"])
    g_p1, gw, gb = conv_bwd(cache["p1"], W[1][0], g_h2)
    grads[1] = (gw, gb)
    g_h1 = relu_g(pool_bwd(g_p1, cache["cnt1"], cache["h1"].shape[:3]) + g_s1, cache["h1"])
    _, gw, gb = conv_bwd(cache["x0"], W[0][0], g_h1)
    grads[0] = (gw, gb)
    return np.concatenate([np.concatenate([gw.ravel(), gb.ravel()]) for gw, gb in grads])


def init_params(seed: int = 0) -> np.ndarray:
    """He-normal weights (fan-in 27*cin), zero biases; the last layer scaled
    down so the untrained network starts near a small multiple of the input."""
    rng = np.random.default_rng(seed)
    out = []
    for li, (ci, co) in enumerate(CONVS):
        std = np.sqrt(2.0 / (27 * ci))
        if li == len(CONVS) - 1:
            std *= 0.1
        out.append(rng.standard_normal(27 * ci * co) * std)
        out.append(np.zeros(co))
    return np.concatenate(out).astype(np.float32)
