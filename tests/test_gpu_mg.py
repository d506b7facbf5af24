"""GPU tests of the V-cycle's first pass on one level (two damped-Jacobi
sweeps from zero, restriction of the residual): the single-pass kernel
against the two-kernel path and against a numpy restatement.  The V-cycle
belongs to the learned preconditioner, which has no counterpart in the
reference; its operator is the pressure stencil of python/boilingsim/
fluid.py:222-240 on class codes."""

from __future__ import annotations

import numpy as np
import pytest

from helpers import dev, host

pytestmark = pytest.mark.gpu

OMEGA = np.float32(0.8)


def _codes(mat):
    """0 air, 1 solid, 2 + solid neighbours for fluid (out of bounds = air)."""
    fl = mat == 0
    solid = (mat != 0) & (mat != 2)
    sp = np.pad(solid, 1)
    nx, ny, nz = mat.shape
    ns = sum(sp[1 + a:1 + a + nx, 1 + b:1 + b + ny, 1 + c:1 + c + nz].astype(np.int32)
             for a, b, c in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)))
    return np.where(fl, 2 + ns, np.where(solid, 1, 0)).astype(np.uint8)


def _apply(code, v):
    """A v with the stencil's neighbour rules, FP64."""
    v = v.astype(np.float64)
    vp = np.pad(v, 1)
    cp = np.pad(code, 1)  # 0 = air outside
    nx, ny, nz = code.shape
    s = np.zeros_like(v)
    for a, b, c in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
        cn = cp[1 + a:1 + a + nx, 1 + b:1 + b + ny, 1 + c:1 + c + nz]
        vn = vp[1 + a:1 + a + nx, 1 + b:1 + b + ny, 1 + c:1 + c + nz]
        s += np.where(cn >= 2, vn, np.where(cn == 1, v, 0.0))
    return np.where(code >= 2, 6.0 * v - s, 0.0)


def _reference(code, r):
    """numpy restatement (FP64) of the pass: z1 = w D^-1 r, z = z1 + w D^-1 (r - A z1),
    rc = 0.5 * sum over children of (r - A z)."""
    fl = code >= 2
    winv = np.where(fl, 0.8 / np.where(fl, 8.0 - code, 1.0), 0.0)
    r = np.where(fl, r.astype(np.float64), 0.0)
    z1 = winv * r
    z = np.where(fl, z1 + winv * (r - _apply(code, z1)), 0.0)
    res = np.where(fl, r - _apply(code, z), 0.0)
    nx, ny, nz = code.shape
    rc = np.zeros(((nx + 1) // 2, (ny + 1) // 2, (nz + 1) // 2))
    for a in range(2):
        for b in range(2):
            for c in range(2):
                part = res[a::2, b::2, c::2]
                rc[:part.shape[0], :part.shape[1], :part.shape[2]] += part
    return z, 0.5 * rc


def _mask(kind, shape, rng):
    from boilsim.fixtures import pot_mask

    if kind == "pot":
        return pot_mask(*shape)
    m = rng.choice([0, 1, 2, 3], size=shape, p=[0.7, 0.12, 0.12, 0.06]).astype(np.int32)
    return m


# (mask, shape): partial j tiles (12 rows) and k tiles (56 cells), several i
# segments with a short last one, fluid on the domain faces (random masks).
CASES = [
    ("pot", (64, 64, 48)),
    ("pot", (90, 30, 60)),
    ("random", (40, 26, 112)),
    ("random", (26, 14, 60)),
    ("random", (8, 12, 4)),
    ("random", (98, 24, 16)),
    ("pot", (128, 36, 64)),
]


@pytest.mark.parametrize("kind,shape", CASES)
def test_fused_pass_matches_two_kernels(simcore, torch_gpu, kind, shape):
    torch = torch_gpu
    rng = np.random.default_rng(sum(shape))
    mat = _mask(kind, shape, rng)
    code = _codes(mat)
    r = rng.standard_normal(shape).astype(np.float32)
    cshape = tuple((n + 1) // 2 for n in shape)
    cd, rd = dev(torch, code), dev(torch, r)
    out = {}
    for fused in (1, 0):
        z = torch.full(shape, np.nan, dtype=torch.float32, device="cuda")
        rc = torch.full(cshape, np.nan, dtype=torch.float32, device="cuda")
        simcore.mg_smooth_restrict(cd.data_ptr(), rd.data_ptr(), z.data_ptr(), rc.data_ptr(), *shape, fused)
        torch.cuda.synchronize()
        out[fused] = (host(z), host(rc))
    (zf, rcf), (zu, rcu) = out[1], out[0]
    print(f"[{kind} {shape}] fluid {int((code >= 2).sum())}, max|rc| {np.abs(rcu).max():.3e}, "
          f"max|rc_f - rc_u| {np.abs(rcf - rcu).max():.3e}")
    # every cell written, z bit-identical (the same float operations)
    assert np.isfinite(zf).all() and np.isfinite(rcf).all()
    assert np.array_equal(zf.view(np.uint32), zu.view(np.uint32))
    # the restriction sums the same FP32 terms in another order
    scale = np.abs(rcu).max()
    np.testing.assert_allclose(rcf, rcu, rtol=0, atol=2e-6 * scale)
    # both against the FP64 restatement
    z64, rc64 = _reference(code, r)
    np.testing.assert_allclose(zu, z64, rtol=0, atol=1e-6 * np.abs(z64).max())
    np.testing.assert_allclose(rcf, rc64, rtol=0, atol=2e-6 * np.abs(rc64).max())


def test_fused_pass_rejects_unsupported_shapes(simcore, torch_gpu):
    torch = torch_gpu
    for shape in ((9, 8, 8), (8, 7, 8), (8, 8, 6)):
        code = dev(torch, _codes(np.zeros(shape, np.int32)))
        r = torch.zeros(shape, dtype=torch.float32, device="cuda")
        z = torch.zeros_like(r)
        rc = torch.zeros(tuple((n + 1) // 2 for n in shape), dtype=torch.float32, device="cuda")
        with pytest.raises(RuntimeError):
            simcore.mg_smooth_restrict(code.data_ptr(), r.data_ptr(), z.data_ptr(), rc.data_ptr(), *shape, 1)
        simcore.mg_smooth_restrict(code.data_ptr(), r.data_ptr(), z.data_ptr(), rc.data_ptr(), *shape, 0)
    torch.cuda.synchronize()

