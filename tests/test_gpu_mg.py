"""GPU tests of the V-cycle's single-pass kernels against their two-kernel
paths and against numpy restatements: the first pass on one level (two
smoothing steps from zero, restriction of the residual), the Chebyshev steps,
and the fine level's last pass (prolongation, post-sweep, dot products).  The
V-cycle belongs to the learned preconditioner, which has no counterpart in the
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


def _dinv(code):
    """D^-1 on fluid cells; a fluid cell walled in by solid on all six sides
    has a zero row (its neighbours all mirror it), for which the kernels take
    D = 1."""
    dg = np.where(code >= 2, 8.0 - code, 0.0)
    return np.where(dg > 0, 1.0 / np.where(dg > 0, dg, 1.0), np.where(code >= 2, 1.0, 0.0))


def _reference(code, r, lo=0.0):
    """numpy restatement (FP64) of the pass: z = q(M) D^-1 r with M = D^-1 A and
    q the smoother's polynomial — damped Jacobi twice from zero, q = 2w - w^2 M
    (w = 0.8), or the two-step Chebyshev polynomial over [lo, 2] in closed form,
    q(l) = (1 - T2(y(l)) / T2(s)) / l, y = s - 2 l / (2 - lo), s = (2 + lo) / (2 - lo)
    (written out below, independent of the kernels' coefficients) — and
    rc = 0.5 * sum over children of (r - A z)."""
    fl = code >= 2
    dinv = _dinv(code)
    r = np.where(fl, r.astype(np.float64), 0.0)
    d = dinv * r
    md = dinv * _apply(code, d)
    if lo > 0:
        b, a = 2.0, lo
        s = (b + a) / (b - a)
        q0 = 8.0 * s / ((b - a) * (2 * s * s - 1))
        q1 = -8.0 / ((b - a) ** 2 * (2 * s * s - 1))
    else:
        q0, q1 = 2 * 0.8, -0.8 * 0.8
    z = np.where(fl, q0 * d + q1 * md, 0.0)
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


# (mask, shape): partial j tiles (12 rows; up to 60 on the taller tiles) and
# k tiles (56 cells), several i segments with a short last one, fluid on the
# domain faces (random masks).
CASES = [
    ("pot", (64, 64, 48)),
    ("pot", (90, 30, 60)),
    ("random", (40, 26, 112)),
    ("random", (26, 14, 60)),
    ("random", (8, 12, 4)),
    ("random", (98, 24, 16)),
    ("pot", (128, 36, 64)),
    ("random", (20, 130, 16)),  # three of the fine level's taller j tiles, the last partial
]


@pytest.mark.parametrize("lo", [0.0, 0.3])
@pytest.mark.parametrize("kind,shape", CASES)
def test_fused_pass_matches_two_kernels(simcore, torch_gpu, kind, shape, lo):
    """lo = 0: damped Jacobi; lo = 0.3: the V-cycle's Chebyshev pre-smoother."""
    torch = torch_gpu
    rng = np.random.default_rng(sum(shape))
    mat = _mask(kind, shape, rng)
    code = _codes(mat)
    r = rng.standard_normal(shape).astype(np.float32)
    cshape = tuple((n + 1) // 2 for n in shape)
    cd, rd = dev(torch, code), dev(torch, r)
    out = {}
    for fused in (2, 1, 0):
        z = torch.full(shape, np.nan, dtype=torch.float32, device="cuda")
        rc = torch.full(cshape, np.nan, dtype=torch.float32, device="cuda")
        simcore.mg_smooth_restrict(cd.data_ptr(), rd.data_ptr(), z.data_ptr(), rc.data_ptr(), *shape, fused, lo)
        torch.cuda.synchronize()
        out[fused] = (host(z), host(rc))
    (zu, rcu) = out[0]
    for fused in (1, 2):  # 256-thread tiles, the fine level's taller tiles
        zf, rcf = out[fused]
        print(f"[{kind} {shape} fused {fused}] fluid {int((code >= 2).sum())}, max|rc| {np.abs(rcu).max():.3e}, "
              f"max|rc_f - rc_u| {np.abs(rcf - rcu).max():.3e}")
        # every cell written, z bit-identical (the same float operations)
        assert np.isfinite(zf).all() and np.isfinite(rcf).all()
        assert np.array_equal(zf.view(np.uint32), zu.view(np.uint32))
        # the restriction sums the same FP32 terms in another order
        scale = np.abs(rcu).max()
        np.testing.assert_allclose(rcf, rcu, rtol=0, atol=2e-6 * scale)
    zf, rcf = out[2]
    # both against the FP64 restatement
    z64, rc64 = _reference(code, r, lo)
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



def _chebyshev_reference(code, r, x0, steps, lo):
    """x_m = x* + T_m(Y)(x0 - x*) / T_m(s) for A x* = r on the fluid cells, with
    Y = s I - (2 / (b - a)) D^-1 A, s = (b + a) / (b - a), [a, b] = [lo, 2]: the
    Chebyshev polynomial through its three-term recurrence, a formulation
    independent of the semi-iterative one the kernels use."""
    import scipy.sparse.linalg as spla
    from oracle import np_ref

    fl = (code >= 2).ravel()
    mat = np.where(code >= 2, 0, np.where(code == 1, 1, 2)).astype(np.int32)
    A = np_ref.laplacian_csr(mat).tocsr()
    idx = np.flatnonzero(fl)
    Af = A[idx][:, idx].tocsc()
    xs = np.zeros(code.size)
    xs[idx] = spla.spsolve(Af, r.ravel().astype(np.float64)[idx])
    dinv = _dinv(code).ravel()
    b, a = 2.0, lo
    s = (b + a) / (b - a)

    def Y(v):
        return s * v - (2.0 / (b - a)) * dinv * (A @ v)

    e0 = np.where(fl, x0.ravel().astype(np.float64) - xs, 0.0)
    v_prev, v = e0, Y(e0)
    t_prev, t = 1.0, s
    for _ in range(steps - 1):
        v_prev, v = v, 2.0 * Y(v) - v_prev
        t_prev, t = t, 2.0 * s * t - t_prev
    return (xs + v / t).reshape(code.shape)


@pytest.mark.parametrize("steps,lo", [(4, 0.02), (20, 0.02), (6, 0.3)])
@pytest.mark.parametrize("kind,shape", [("pot", (24, 24, 16)), ("random", (13, 10, 9))])
def test_chebyshev_matches_polynomial(simcore, torch_gpu, kind, shape, steps, lo):
    """The V-cycle's Chebyshev steps (post-smoother: 4 over [0.02, 2]; polish:
    20 over [0.02, 2]) against the Chebyshev error polynomial in FP64."""
    torch = torch_gpu
    rng = np.random.default_rng(3 + steps + sum(shape))
    mat = _mask(kind, shape, rng)
    code = _codes(mat)
    fl = code >= 2
    # A x* = r with a known smooth-plus-noise x*, started from noise
    xs = np.where(fl, rng.standard_normal(shape), 0.0)
    r = np.where(fl, _apply(code, xs), 0.0).astype(np.float32)
    x0 = np.where(fl, rng.standard_normal(shape), 0.0).astype(np.float32)
    x = dev(torch, x0.copy())
    tmp = torch.full(shape, np.nan, dtype=torch.float32, device="cuda")
    cd, rd = dev(torch, code), dev(torch, r)  # held: the kernels run after the call returns
    simcore.mg_chebyshev(cd.data_ptr(), rd.data_ptr(), x.data_ptr(), tmp.data_ptr(), *shape, steps, lo)
    torch.cuda.synchronize()
    got = host(x)
    ref = _chebyshev_reference(code, r, x0, steps, lo)
    err0 = np.abs(x0 - ref).max()
    print(f"[{kind} {shape} {steps} steps over [{lo}, 2]] max|x - ref| {np.abs(got - ref).max():.2e}, "
          f"max|x0 - ref| {err0:.2e}")
    assert not got[~fl].any()
    np.testing.assert_allclose(got, ref, rtol=0, atol=2e-5 * max(np.abs(ref).max(), 1.0))


def test_chebyshev_rejects_bad_arguments(simcore, torch_gpu):
    torch = torch_gpu
    shape = (8, 8, 8)
    code = dev(torch, _codes(np.zeros(shape, np.int32)))
    z = torch.zeros(shape, dtype=torch.float32, device="cuda")
    for steps, lo in ((3, 0.1), (-2, 0.1), (4, 0.0), (4, 2.0)):
        with pytest.raises(RuntimeError):
            simcore.mg_chebyshev(code.data_ptr(), z.data_ptr(), z.data_ptr(), z.data_ptr(), *shape, steps, lo)


def _prolong_reference(code, ec):
    """Cell-centred trilinear prolongation (weights 3/4 to the parent coarse
    cell, 1/4 to the neighbour on the child's side, clamped at the domain),
    FP64, on fluid cells only."""
    shape, cshape = code.shape, ec.shape

    def axis(n, N):
        i = np.arange(n)
        return i // 2, np.clip(i // 2 + np.where(i & 1, 1, -1), 0, N - 1)

    (I, I2), (J, J2), (K, K2) = (axis(n, N) for n, N in zip(shape, cshape))
    e = ec.astype(np.float64)
    out = np.zeros(shape)
    for ii, wa in ((I, .75), (I2, .25)):
        for jj, wb in ((J, .75), (J2, .25)):
            for kk, wc in ((K, .75), (K2, .25)):
                out += wa * wb * wc * e[ii][:, jj][:, :, kk]
    return np.where(code >= 2, out, 0.0)


# nz % 4 == 0 throughout (the single-pass kernel's condition): partial j and
# k tiles, odd nx and ny, one-row grids, several i segments with a short last.
PP_CASES = [
    ("pot", (64, 64, 48)),
    ("pot", (90, 30, 60)),
    ("random", (40, 26, 112)),
    ("random", (27, 15, 60)),
    ("random", (33, 29, 116)),
    ("random", (8, 12, 4)),
    ("random", (98, 24, 16)),
    ("random", (5, 1, 8)),
    ("random", (20, 130, 16)),  # three of the fine level's taller j tiles, the last partial
]


@pytest.mark.parametrize("with_old", [True, False])
@pytest.mark.parametrize("kind,shape", PP_CASES)
def test_prolong_post_matches_two_kernels(simcore, torch_gpu, kind, shape, with_old):
    """The fine level's last pass (prolongation, undamped post-sweep, r.z and
    r.z_old) in one kernel against the prolongation and sweep kernels."""
    torch = torch_gpu
    rng = np.random.default_rng(7 + sum(shape) + with_old)
    code = _codes(_mask(kind, shape, rng))
    fl = code >= 2
    cshape = tuple((n + 1) // 2 for n in shape)
    r = np.where(fl, rng.standard_normal(shape), 0.0).astype(np.float32)
    z_in = np.where(fl, rng.standard_normal(shape), 0.0).astype(np.float32)
    z_old = rng.standard_normal(shape).astype(np.float32)
    ec = rng.standard_normal(cshape).astype(np.float32)
    held = [dev(torch, a) for a in (code, z_in, ec, r, z_old)]  # held while the kernels run
    cd, zid, ecd, rd, zod = held
    n = int(np.prod(shape))
    out = {}
    for fused in (2, 1, 0):
        z = torch.full(shape, np.nan, dtype=torch.float32, device="cuda")
        tmp = torch.full(shape, np.nan, dtype=torch.float32, device="cuda")
        part = torch.full((2 * n + 64,), np.nan, dtype=torch.float64, device="cuda")
        g = simcore.mg_prolong_post(cd.data_ptr(), zid.data_ptr(), ecd.data_ptr(), rd.data_ptr(),
                                    zod.data_ptr() if with_old else 0, z.data_ptr(), tmp.data_ptr(), part.data_ptr(),
                                    part.numel(), *shape, fused, 1.0)
        torch.cuda.synchronize()
        p = host(part)
        out[fused] = (host(z), p[:g].sum(), p[g:2 * g].sum())
    (zf, af, bf), (zu, au, bu) = out[1], out[0]
    zt2, at2, bt2 = out[2]  # the fine level's taller tiles
    print(f"[{kind} {shape} z_old {with_old}] fluid {int(fl.sum())}, r.z {af:.6e} / {at2:.6e} / {au:.6e}, "
          f"r.z_old {bf:.6e} / {bt2:.6e} / {bu:.6e}")
    assert np.isfinite(zf).all() and np.isfinite(zt2).all()
    assert np.array_equal(zf.view(np.uint32), zu.view(np.uint32))
    assert np.array_equal(zt2.view(np.uint32), zu.view(np.uint32))
    # FP64 restatement: zt = z_in + P ec, one undamped Jacobi sweep from zt
    zt = z_in.astype(np.float64) + _prolong_reference(code, ec)
    z64 = np.where(fl, zt + _dinv(code) * (r - _apply(code, zt)), 0.0)
    np.testing.assert_allclose(zf, z64, rtol=0, atol=2e-6 * np.abs(z64).max())
    # the partials sum FP64 products of the same FP32 values in another order
    a64 = float((r.astype(np.float64) * zf.astype(np.float64)).sum())
    b64 = float((r.astype(np.float64) * z_old.astype(np.float64)).sum()) if with_old else 0.0
    scale_a = float(np.abs(r.astype(np.float64) * zf).sum())
    scale_b = float(np.abs(r.astype(np.float64) * z_old).sum())
    for got in (af, at2, au):
        assert abs(got - a64) <= 1e-11 * scale_a
    for got in (bf, bt2, bu):
        assert abs(got - b64) <= 1e-11 * scale_b


def test_prolong_post_rejects_unsupported(simcore, torch_gpu):
    torch = torch_gpu
    part = torch.zeros(4096, dtype=torch.float64, device="cuda")
    for shape, part_len in (((8, 8, 6), part.numel()), ((8, 8, 8), 1)):  # nz % 4 != 0; too few partial slots
        code = dev(torch, _codes(np.zeros(shape, np.int32)))
        z = torch.zeros(shape, dtype=torch.float32, device="cuda")
        z2 = torch.zeros(shape, dtype=torch.float32, device="cuda")
        ec = torch.zeros(tuple((n + 1) // 2 for n in shape), dtype=torch.float32, device="cuda")
        with pytest.raises(RuntimeError):
            simcore.mg_prolong_post(code.data_ptr(), z.data_ptr(), ec.data_ptr(), z.data_ptr(), 0, z2.data_ptr(),
                                    z2.data_ptr(), part.data_ptr(), part_len, *shape, 1, 1.0)
        torch.cuda.synchronize()
