"""GPU tests of the semi-Lagrangian advection (SimCore.extend_temperature /
advect / advect_all against the oracle's restatement of
python/boilingsim/fluid.py:33-103, 338-492, 737-775), and of several steps of
advection + no-slip + projection against the oracle's unfused sequence."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from helpers import bit_diff_report, dev, host, same_bits

pytestmark = pytest.mark.gpu

ORIGIN = (-0.0123, -0.011, -0.004)


def _case(seed, shape, speed, p=(0.6, 0.15, 0.2, 0.05)):
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    mat = rng.choice([0, 1, 2, 3], size=shape, p=list(p)).astype(np.int32)
    f = [(rng.standard_normal(s) * speed).astype(np.float32)
         for s in ((nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1))]
    t = (290.0 + 100.0 * rng.random(shape)).astype(np.float32)
    return mat, f, t


def _advect_gpu(simcore, torch, f, t, mat, origin, dx, dt):
    d = [dev(torch, a) for a in (*f, t)]
    md = dev(torch, mat)
    shape = mat.shape
    t_ext = torch.full(shape, -1.0, dtype=torch.float32, device="cuda")
    outs = [torch.full(a.shape, 123.0, dtype=torch.float32, device="cuda") for a in d]
    simcore.extend_temperature(t_ext.data_ptr(), d[3].data_ptr(), md.data_ptr(), *shape)
    simcore.advect(*(o.data_ptr() for o in outs), *(a.data_ptr() for a in d), t_ext.data_ptr(), md.data_ptr(),
                   *shape, origin, dx, dt)
    torch.cuda.synchronize()
    return [host(o) for o in outs], host(t_ext)


@pytest.mark.parametrize("seed,shape,speed", [
    (0, (6, 7, 5), 0.3), (1, (9, 4, 11), 1.0), (2, (1, 1, 1), 0.5), (3, (70, 3, 4), 0.5),
    (4, (5, 66, 9), 3.0), (5, (12, 10, 80), 0.1), (6, (33, 17, 44), 0.4), (7, (64, 64, 48), 0.2),
    (8, (2, 3, 1), 2.0),
])
def test_advect_bit_exact(simcore, torch_gpu, seed, shape, speed):
    """Random classes with fluid on the domain edges, displacements up to
    several cells (clamps on every axis), a non-zero origin, an arbitrary dx,
    axes past 64 cells (the clamped upper index)."""
    torch = torch_gpu
    mat, f, t = _case(seed, shape, speed)
    dx, dt = 0.0017, 0.01
    want = oracle.advect(*f, t, mat, ORIGIN, dx, dt)
    got, t_ext = _advect_gpu(simcore, torch, f, t, mat, ORIGIN, dx, dt)
    assert same_bits(t_ext, want[4]), bit_diff_report(t_ext, want[4])
    for name, g, w in zip(("ux", "uy", "uz", "T"), got, want[:4]):
        assert same_bits(g, w), f"{name}: " + bit_diff_report(g, w)


@pytest.mark.parametrize("name", ["small", "dev"])
def test_advect_pot(simcore, torch_gpu, name):
    """The pot (64x64x48 and 110x110x70) with the seeded velocities plus the
    plume, a hot base, the geometry's origin and a CFL-limited dt."""
    torch = torch_gpu
    from boilsim.fixtures import CONFIGS, pot_mask, seeded_velocities

    cfg = CONFIGS[name]
    nx, ny, nz = cfg.shape
    mat = pot_mask(nx, ny, nz)
    f = list(seeded_velocities(nx, ny, nz, seed=3, plume=True))
    t = np.full(cfg.shape, 295.0, np.float32)
    t[(mat == 1) & (np.arange(nz)[None, None, :] < nz // 6)] = 480.0
    t[mat == 0] += np.random.default_rng(1).random(int((mat == 0).sum())).astype(np.float32)
    origin = (-nx * cfg.dx / 2, -ny * cfg.dx / 2, -4 * cfg.dx)
    umax = max(float(np.abs(a).max()) for a in f)
    dt = 0.4 * cfg.dx / umax
    want = oracle.advect(*f, t, mat, origin, cfg.dx, dt)
    got, _ = _advect_gpu(simcore, torch, f, t, mat, origin, cfg.dx, dt)
    for nm, g, w in zip(("ux", "uy", "uz", "T"), got, want[:4]):
        assert same_bits(g, w), f"{nm}: " + bit_diff_report(g, w)


def test_advect_all_in_place(simcore, torch_gpu):
    """boilsim.fluid.advect_all: the advected fields end in the grid, the
    workspace arrays hold the same values and T_ext the extension."""
    torch = torch_gpu
    from boilsim import fluid

    mat, f, t = _case(11, (20, 18, 24), 0.3)
    want = oracle.advect(*f, t, mat, ORIGIN, 0.002, 0.004)
    g = fluid.make_grid(mat, *f, 0.002, T=t, origin=ORIGIN)
    ws = fluid.allocate_fluid_workspace(g)
    fluid.advect_all(g, ws, 0.004)
    torch.cuda.synchronize()
    for nm, w in zip(("ux", "uy", "uz", "T"), want[:4]):
        got = host(getattr(g, nm))
        assert same_bits(got, w), f"{nm}: " + bit_diff_report(got, w)
    assert same_bits(host(ws.T_tmp), want[3]) and same_bits(host(ws.T_ext), want[4])


def test_grid_too_large_is_an_error(simcore, torch_gpu):
    """Face arrays beyond 2^31 elements are refused before any launch."""
    with pytest.raises(RuntimeError):
        simcore.advect(*([0] * 10), 1300, 1300, 1300, ORIGIN, 1e-3, 1e-3)


@pytest.mark.parametrize("solver", ["jacobi", "pcg-hip"])
def test_steps_match_oracle(simcore, torch_gpu, monkeypatch, solver):
    """Eight steps of advect_all, then no-slip + projection + no-slip, on the
    64x64x48 pot with the plume: every field bit-identical to the oracle's
    unfused sequence after every step for the Jacobi solver; for the PCG
    (no oracle of its own here) the GPU's unfused and fused sequences agree
    bit for bit."""
    torch = torch_gpu
    from boilsim import fluid
    from boilsim.fixtures import CONFIGS, DT, RHO, pot_mask, seeded_velocities

    cfg = CONFIGS["small"]
    nx, ny, nz = cfg.shape
    mat = pot_mask(nx, ny, nz)
    f = list(seeded_velocities(nx, ny, nz, seed=9, plume=True))
    t = np.full(cfg.shape, 300.0, np.float32)
    t[mat == 0] += 5.0 * np.random.default_rng(2).random(int((mat == 0).sum())).astype(np.float32)
    origin = (-nx * cfg.dx / 2, -ny * cfg.dx / 2, -4 * cfg.dx)
    dt = 0.002
    monkeypatch.delenv("BOILINGSIM_USE_RUST_PRESSURE", raising=False)
    monkeypatch.setenv("BOILINGSIM_PRESSURE_SOLVER", solver)
    sc = fluid.ScenarioConfig()
    sc.solver.pressure_max_iter = 60 if solver == "jacobi" else 2000
    sc.solver.pressure_tol = 1e-6
    ga = fluid.make_grid(mat, *f, cfg.dx, T=t, origin=origin)
    wa = fluid.allocate_fluid_workspace(ga)
    gb = fluid.make_grid(mat, *f, cfg.dx, T=t, origin=origin)
    wb = fluid.allocate_fluid_workspace(gb)
    ref = [a.copy() for a in (*f, t)]
    for step in range(8):
        fluid.advect_all(ga, wa, dt)
        fluid.enforce_no_slip(ga)
        m_a = fluid.pressure_projection(ga, wa, sc, dt, RHO)
        fluid.enforce_no_slip(ga)
        fluid.advect_all(gb, wb, dt)
        m_b = fluid.projection_step(gb, wb, sc, dt, RHO)
        assert m_a == m_b
        for nm in ("ux", "uy", "uz", "T", "p"):
            a, b = host(getattr(ga, nm)), host(getattr(gb, nm))
            assert same_bits(a, b), f"step {step} {nm}: " + bit_diff_report(a, b)
        if solver == "jacobi":
            ux, uy, uz, tt, _ = oracle.advect(*ref, mat, origin, cfg.dx, dt)
            p_r, _, _, md_r, mu_r = oracle.projection_step(ux, uy, uz, mat, cfg.dx, dt, RHO, 60)
            ref = [ux, uy, uz, tt]
            for nm, r in zip(("ux", "uy", "uz", "T"), ref):
                g = host(getattr(gb, nm))
                assert same_bits(g, r), f"step {step} {nm} vs oracle: " + bit_diff_report(g, r)
            assert same_bits(host(gb.p), p_r) and m_b == np.float32(md_r) and wb.u_max == np.float32(mu_r)
