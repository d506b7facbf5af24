"""GPU tests of the no-slip faces and the max-velocity reductions
(SimCore.enforce_no_slip / divergence_no_slip / divergence_max / max_velocity /
projection_step_jacobi against the oracle's restatement of fluid.py:111-162,
201-219, 612-734 and 796-833), and of boilsim.fluid.projection_step against
the reference's unfused sequence (pipeline.py:606-620) for every solver."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from helpers import bit_diff_report, dev, host, same_bits

pytestmark = pytest.mark.gpu

# nz % 4 == 0 (float4 path), nz % 4 == 2 (float2) and odd nz (scalar)
SHAPES = [(16, 16, 16), (24, 12, 20), (9, 7, 13), (33, 17, 42), (5, 40, 2), (1, 1, 1), (64, 64, 48), (3, 5, 1)]


def _faces(rng, shape, scale=1.0, specials=True):
    nx, ny, nz = shape
    f = [(rng.standard_normal(s) * scale).astype(np.float32)
         for s in ((nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1))]
    if specials:
        f[0].ravel()[::7] = -0.0
        f[1].ravel()[::5] = -0.0
        f[2].ravel()[::3] = -0.0
        f[2].ravel()[1::11] = 0.0
    return f


def _mat(rng, shape):
    """Random classes with fluid on the domain edges (no air ring)."""
    return rng.choice([0, 1, 2, 3], size=shape, p=[0.6, 0.15, 0.2, 0.05]).astype(np.int32)


def _bits(t):
    return host(t).view(np.float32)


@pytest.mark.parametrize("seed", [0, 1])
@pytest.mark.parametrize("shape", SHAPES)
def test_enforce_no_slip_bit_exact(simcore, torch_gpu, shape, seed):
    torch = torch_gpu
    rng = np.random.default_rng(seed)
    mat = _mat(rng, shape)
    ux, uy, uz = _faces(rng, shape)
    ref = [a.copy() for a in (ux, uy, uz)]
    oracle.enforce_no_slip(*ref, mat)
    d = [dev(torch, a) for a in (ux, uy, uz)]
    md = dev(torch, mat)
    simcore.enforce_no_slip(*(t.data_ptr() for t in d), md.data_ptr(), *shape)
    torch.cuda.synchronize()
    for g, r in zip(d, ref):
        assert same_bits(host(g), r), bit_diff_report(host(g), r)


def test_enforce_no_slip_misaligned(simcore, torch_gpu):
    """Face views one element off 16-byte alignment take the scalar path."""
    torch = torch_gpu
    shape = (12, 10, 16)
    rng = np.random.default_rng(3)
    mat = _mat(rng, shape)
    f = _faces(rng, shape)
    ref = [a.copy() for a in f]
    oracle.enforce_no_slip(*ref, mat)
    views, keep = [], []
    for a in f:
        buf = torch.zeros(a.size + 1, dtype=torch.float32, device="cuda")
        v = buf[1:].view(a.shape)
        v.copy_(torch.from_numpy(a))
        views.append(v)
        keep.append(buf)
    md = dev(torch, mat)
    simcore.enforce_no_slip(*(v.data_ptr() for v in views), md.data_ptr(), *shape)
    torch.cuda.synchronize()
    for g, r in zip(views, ref):
        assert same_bits(host(g), r), bit_diff_report(host(g), r)


@pytest.mark.parametrize("shape", SHAPES)
def test_divergence_no_slip_bit_exact(simcore, torch_gpu, shape):
    """enforce_no_slip then compute_divergence, fused: both the faces and the
    divergence match the two reference kernels run in order."""
    torch = torch_gpu
    rng = np.random.default_rng(11)
    mat = _mat(rng, shape)
    f = _faces(rng, shape)
    ref = [a.copy() for a in f]
    oracle.enforce_no_slip(*ref, mat)
    dx = 1.7e-3
    div_ref = oracle.divergence(*ref, mat, dx)
    d = [dev(torch, a) for a in f]
    md = dev(torch, mat)
    div = torch.full(shape, 7.0, dtype=torch.float32, device="cuda")
    simcore.divergence_no_slip(div.data_ptr(), *(t.data_ptr() for t in d), md.data_ptr(), *shape, dx)
    torch.cuda.synchronize()
    assert same_bits(host(div), div_ref), bit_diff_report(host(div), div_ref)
    for g, r in zip(d, ref):
        assert same_bits(host(g), r), bit_diff_report(host(g), r)


@pytest.mark.parametrize("shape", SHAPES)
def test_divergence_max_bit_exact(simcore, torch_gpu, shape):
    """compute_divergence with max|div| and max|u| from the same pass; the
    faces are not written."""
    torch = torch_gpu
    rng = np.random.default_rng(12)
    mat = _mat(rng, shape)
    f = _faces(rng, shape, specials=False)
    f[1].ravel()[-1] = -9.25  # the largest |u| on the last y face
    dx = 2.3e-3
    div_ref = oracle.divergence(*f, mat, dx)
    d = [dev(torch, a) for a in f]
    md = dev(torch, mat)
    div = torch.empty(shape, dtype=torch.float32, device="cuda")
    mb = torch.full((2,), -1, dtype=torch.int32, device="cuda")
    simcore.divergence_max(div.data_ptr(), *(t.data_ptr() for t in d), md.data_ptr(), *shape, dx, mb.data_ptr())
    torch.cuda.synchronize()
    assert same_bits(host(div), div_ref), bit_diff_report(host(div), div_ref)
    m = _bits(mb)
    assert m[0] == np.float32(np.abs(div_ref).max())
    assert m[1] == np.float32(oracle.max_velocity(*f)) == np.float32(9.25)
    for g, r in zip(d, f):
        assert same_bits(host(g), r)


@pytest.mark.parametrize("shape", [(16, 16, 16), (9, 7, 13), (64, 64, 48), (110, 110, 70), (5, 40, 2)])
def test_max_velocity(simcore, torch_gpu, shape):
    torch = torch_gpu
    rng = np.random.default_rng(13)
    f = _faces(rng, shape, scale=0.2, specials=False)
    d = [dev(torch, a) for a in f]
    out = torch.full((1,), -1, dtype=torch.int32, device="cuda")
    for q in range(3):
        f[q].ravel()[-1] = 3.0 + q  # the largest value at the end of each array in turn
        d[q].view(-1)[-1] = 3.0 + q
        simcore.max_velocity(*(t.data_ptr() for t in d), *shape, out.data_ptr())
        torch.cuda.synchronize()
        assert _bits(out)[0] == np.float32(oracle.max_velocity(*f)) == np.float32(3.0 + q)
    # misaligned uy (the scalar path of the reduction)
    buf = torch.zeros(f[1].size + 1, dtype=torch.float32, device="cuda")
    v = buf[1:]
    v.copy_(torch.from_numpy(f[1].ravel()))
    v[0] = -11.0
    simcore.max_velocity(d[0].data_ptr(), v.data_ptr(), d[2].data_ptr(), *shape, out.data_ptr())
    torch.cuda.synchronize()
    assert _bits(out)[0] == np.float32(11.0)


@pytest.mark.parametrize("n_iter", [0, 1, 2, 7, 200])
@pytest.mark.parametrize("geometry", ["pot", "random", "walled"])
def test_projection_step_jacobi_bit_exact(simcore, torch_gpu, geometry, n_iter):
    """no-slip, Jacobi projection, no-slip (pipeline.py:606-620) against the
    oracle's unfused sequence: p, p_tmp, div, the faces, max|div| and max|u|
    all identical."""
    torch = torch_gpu
    rng = np.random.default_rng(21)
    if geometry == "pot":
        from boilsim.fixtures import pot_mask

        mat = pot_mask(48, 48, 40)
    elif geometry == "random":
        mat = _mat(rng, (30, 26, 22))
    else:
        mat = np.full((20, 20, 20), 1, np.int32)
        mat[1:-1, 1:-1, 1:-1] = 0
    shape = mat.shape
    f = _faces(rng, shape, scale=0.05)
    dx, dt, rho = 2e-3, 0.01, 997.0
    ref = [a.copy() for a in f]
    p_r, pt_r, div_r, md_r, mu_r = oracle.projection_step(*ref, mat, dx, dt, rho, n_iter)
    d = [dev(torch, a) for a in f]
    md = dev(torch, mat)
    p = torch.full(shape, 5.0, dtype=torch.float32, device="cuda")
    pt = torch.full(shape, 6.0, dtype=torch.float32, device="cuda")
    div = torch.empty(shape, dtype=torch.float32, device="cuda")
    mb = torch.full((2,), -1, dtype=torch.int32, device="cuda")
    simcore.projection_step_jacobi(p.data_ptr(), pt.data_ptr(), div.data_ptr(), *(t.data_ptr() for t in d),
                                   md.data_ptr(), *shape, dx, dt, rho, n_iter, mb.data_ptr())
    torch.cuda.synchronize()
    for g, r in ((host(p), p_r), (host(pt), pt_r), (host(div), div_r), *((host(t), r) for t, r in zip(d, ref))):
        assert same_bits(g, r), bit_diff_report(g, r)
    m = _bits(mb)
    assert m[0] == np.float32(md_r) and m[1] == np.float32(mu_r)


@pytest.mark.parametrize("solver", ["jacobi", "cg-rust", "pcg-hip", "lpcg-hip"])
def test_projection_step_matches_unfused_sequence(simcore, torch_gpu, monkeypatch, solver):
    """boilsim.fluid.projection_step against enforce_no_slip,
    pressure_projection, enforce_no_slip, compute_max_velocity on the 64x64x48
    pot, for every solver: identical faces, pressure and max|div|, and
    ws.u_max equal to compute_max_velocity; compute_cfl_dt from it."""
    torch = torch_gpu
    from boilsim import fluid
    from boilsim.fixtures import CONFIGS, DT, RHO, pot_mask, seeded_velocities

    cfg = CONFIGS["small"]
    mat = pot_mask(*cfg.shape)
    ux, uy, uz = seeded_velocities(*cfg.shape, seed=5)
    monkeypatch.delenv("BOILINGSIM_USE_RUST_PRESSURE", raising=False)
    monkeypatch.setenv("BOILINGSIM_PRESSURE_SOLVER", solver)
    sc = fluid.ScenarioConfig()
    sc.solver.pressure_max_iter = 200 if solver == "jacobi" else 5000
    sc.solver.pressure_tol = 1e-6
    ga = fluid.make_grid(mat, ux, uy, uz, cfg.dx)
    wa = fluid.allocate_fluid_workspace(ga)
    fluid.enforce_no_slip(ga)
    m_a = fluid.pressure_projection(ga, wa, sc, DT, RHO)
    fluid.enforce_no_slip(ga)
    u_a = fluid.compute_max_velocity(ga, wa)
    gb = fluid.make_grid(mat, ux, uy, uz, cfg.dx)
    wb = fluid.allocate_fluid_workspace(gb)
    m_b = fluid.projection_step(gb, wb, sc, DT, RHO)
    print(f"{solver}: max|div| {m_b:.3e}, max|u| {wb.u_max:.4f}, {fluid.LAST_SOLVE_STATS}")
    for name in ("ux", "uy", "uz", "p"):
        a, b = host(getattr(ga, name)), host(getattr(gb, name))
        assert same_bits(a, b), f"{name}: " + bit_diff_report(a, b)
    assert m_a == m_b and u_a == wb.u_max > 0
    assert fluid.compute_cfl_dt(gb, sc, u_max=wb.u_max) == sc.solver.cfl_safety_factor * gb.dx / wb.u_max
    assert fluid.compute_cfl_dt(gb, sc, ws=wb) == fluid.compute_cfl_dt(gb, sc, u_max=wb.u_max)


def test_cfl_dt_at_rest(simcore, torch_gpu):
    """compute_cfl_dt (fluid.py:836-844): a fluid at rest gets max_dt_s."""
    torch = torch_gpu
    from boilsim import fluid

    mat = np.zeros((8, 8, 8), np.int32)
    z = [np.zeros(s, np.float32) for s in ((9, 8, 8), (8, 9, 8), (8, 8, 9))]
    g = fluid.make_grid(mat, *z, 1e-3)
    sc = fluid.ScenarioConfig()
    assert fluid.compute_max_velocity(g) == 0.0
    assert fluid.compute_cfl_dt(g, sc) == sc.solver.max_dt_s


def test_hydrostatic_closed_box(simcore, torch_gpu):
    """test_fluid.py:199-217 through projection_step: a closed box at rest
    stays at rest (max|div| < 1e-6, max|u| < 1e-6)."""
    torch = torch_gpu
    from boilsim import fluid

    mat = np.full((12, 12, 12), 1, np.int32)
    mat[1:-1, 1:-1, 1:-1] = 0
    z = [np.zeros(s, np.float32) for s in ((13, 12, 12), (12, 13, 12), (12, 12, 13))]
    g = fluid.make_grid(mat, *z, 2e-3)
    ws = fluid.allocate_fluid_workspace(g)
    m = fluid.projection_step(g, ws, fluid.ScenarioConfig(), 0.01, 997.0)
    assert m < 1e-6 and ws.u_max < 1e-6
