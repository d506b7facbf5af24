"""GPU tests of the mixed-precision iterative refinement (SimCore.refine_solve):
an FP64 outer loop around a Jacobi-PCG whose vectors are stored in FP32, FP16
or BF16.  The FP64 true residual it reports is recomputed by the oracle from
the returned FP64 iterate against the same FP32-rounded b, and the solution
is compared with the oracle's FP64 PCG solve."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from helpers import dev, host

pytestmark = pytest.mark.gpu

DX, DT, RHO = 2e-3, 1e-3, 958.0


def b32(div, mat, dx=DX, dt=DT, rho=RHO):
    """b = bscale * div rounded as the solver forms it in FP32:
    bscale = ((-rho * dx) * dx) / dt (pressure_solve_pcg.cu:137-138)."""
    f = np.float32
    bscale = ((f(-rho) * f(dx)) * f(dx)) / f(dt)
    b = (bscale * div.astype(np.float32)).astype(np.float32)
    return np.where(mat == 0, b, np.float32(0)).astype(np.float64)


def refine(core, torch, mat_d, div_d, shape, storage, tol, warm=None, inner_tol=None, want_x32=False):
    ws = torch.empty(core.refine_workspace_bytes(*shape), dtype=torch.uint8, device="cuda")
    x = torch.zeros(shape, dtype=torch.float64, device="cuda") if warm is None else warm.clone()
    x32 = torch.full(shape, 7.0, dtype=torch.float32, device="cuda") if want_x32 else None
    st = core.refine_solve(x.data_ptr(), div_d.data_ptr(), mat_d.data_ptr(), *shape, DX, DT, RHO, tol,
                           ws.data_ptr(), ws.numel(), storage=storage, inner_tol=inner_tol,
                           warm_start=warm is not None, x32_ptr=x32.data_ptr() if want_x32 else 0)
    torch.cuda.synchronize()
    return x, x32, st


@pytest.mark.parametrize("storage", ["fp32", "fp16", "bf16"])
def test_refine_divergence_rhs(simcore, torch_gpu, storage):
    """64x64x48 pot, divergence of seeded velocities (seed 7): converged to
    1e-8 (beyond what an FP32 iterate reaches), the reported FP64 residual
    equals the oracle's to 1e-9 relative, and the error against the FP64
    oracle solution is at the level the conditioning allows."""
    torch = torch_gpu
    from boilsim.fixtures import pot_mask, seeded_velocities

    shape = (64, 64, 48)
    mat = pot_mask(*shape)
    ux, uy, uz = seeded_velocities(*shape, seed=7)
    div = oracle.divergence(ux, uy, uz, mat, DX)
    mat_d, div_d = dev(torch, mat), dev(torch, div)
    tol = 1e-8
    x, x32, st = refine(simcore, torch, mat_d, div_d, shape, storage, tol, want_x32=True)
    xh = host(x)
    b = b32(div, mat)
    rel2, _ = oracle.residual_f64(xh, b, mat)
    x64, _ = oracle.pcg_f64(b, mat, tol=1e-13)
    fl = mat == 0
    err = np.linalg.norm((xh - x64)[fl]) / np.linalg.norm(x64[fl])
    print(f"[{storage}] {st}; oracle relres {rel2:.3e}; error vs FP64 solve {err:.3e}")
    assert st["converged"] == 1 and st["outer"] >= 1
    assert rel2 <= tol * 1.0001
    assert abs(rel2 - st["relres_true"]) <= 1e-9 * rel2 + 1e-20
    assert err < 1e-5
    # The FP32 output is the FP64 iterate rounded.
    assert np.array_equal(host(x32), xh.astype(np.float32))
    # Non-fluid cells stay exactly zero.
    assert not xh[~fl].any()


@pytest.mark.parametrize("storage", ["fp16", "bf16"])
def test_refine_16bit_production_grid(simcore, torch_gpu, storage):
    """440x440x280 pot, divergence of the seed-11 faces (bench.py's
    mixed-precision leg): the refinement around 16-bit inner vectors reaches a
    true 1e-6 by the oracle's FP64 residual of the returned FP64 iterate, and
    the residual it reports is the oracle's to 1e-9 relative."""
    torch = torch_gpu
    from boilsim.fixtures import CONFIGS, DT as DT_P, RHO as RHO_P, pot_mask, seeded_velocities

    cfg = CONFIGS["prod"]
    shape = cfg.shape
    print(f"[refine 440] {storage}", flush=True)
    mat = pot_mask(*shape)
    ux, uy, uz = seeded_velocities(*shape, seed=11, plume=False)
    div = oracle.divergence(ux, uy, uz, mat, cfg.dx)
    del ux, uy, uz
    mat_d, div_d = dev(torch, mat), dev(torch, div)
    ws = torch.empty(simcore.refine_workspace_bytes(*shape), dtype=torch.uint8, device="cuda")
    x = torch.zeros(shape, dtype=torch.float64, device="cuda")
    st = simcore.refine_solve(x.data_ptr(), div_d.data_ptr(), mat_d.data_ptr(), *shape, cfg.dx, DT_P, RHO_P, 1e-6,
                              ws.data_ptr(), ws.numel(), storage=storage)
    torch.cuda.synchronize()
    xh = host(x)
    rel2, _ = oracle.residual_f64(xh, b32(div, mat, cfg.dx, DT_P, RHO_P), mat)
    print(f"  {st}; oracle relres {rel2:.3e}", flush=True)
    assert st["converged"] == 1 and rel2 <= 1.0001e-6
    assert abs(rel2 - st["relres_true"]) <= 1e-9 * rel2
    assert not xh[mat != 0].any()


def test_refine_smooth_rhs_beats_fp32_floor(simcore, torch_gpu):
    """110x110x70 pot with a smooth right-hand side (Gaussian random field,
    correlation length 12 cells): the FP32 true-residual PCG stalls above
    1e-7 at its FP32 iterate, the refinement reaches it in FP64 with FP16
    inner vectors."""
    torch = torch_gpu
    from boilsim.fixtures import pot_mask
    from boilsim.learned import grf_torch

    shape = (110, 110, 70)
    mat = pot_mask(*shape)
    mat_d = dev(torch, mat)
    div_d, _ = grf_torch(torch, mat_d == 0, seed=5, ell=12.0)
    div = host(div_d)
    b = b32(div, mat)
    tol = 1e-7
    x, _, st = refine(simcore, torch, mat_d, div_d, shape, "fp16", tol)
    rel2, _ = oracle.residual_f64(host(x), b, mat)
    # FP32 iterate of the plain true-residual PCG on the same problem.
    wsp = torch.empty(simcore.pcg_workspace_bytes(*shape), dtype=torch.uint8, device="cuda")
    x32 = torch.zeros(shape, dtype=torch.float32, device="cuda")
    sp = simcore.pcg_solve(x32.data_ptr(), div_d.data_ptr(), mat_d.data_ptr(), *shape, DX, DT, RHO, tol, 5000,
                           wsp.data_ptr(), wsp.numel())
    torch.cuda.synchronize()
    rel32, _ = oracle.residual_f64(host(x32).astype(np.float64), b, mat)
    print(f"refine fp16: {st}; oracle relres {rel2:.3e}; FP32 PCG: {sp}; its oracle relres {rel32:.3e}")
    assert st["converged"] == 1 and rel2 <= tol * 1.0001
    assert rel32 > tol  # the FP32 iterate cannot represent a 1e-7 solution here


def test_refine_warm_start_and_zero_rhs(simcore, torch_gpu):
    """A warm start from a converged FP64 solution takes no refinement step;
    b = 0 returns x = 0 at once."""
    torch = torch_gpu
    from boilsim.fixtures import pot_mask, seeded_velocities

    shape = (40, 40, 30)
    mat = pot_mask(*shape)
    ux, uy, uz = seeded_velocities(*shape, seed=3)
    div = oracle.divergence(ux, uy, uz, mat, DX)
    mat_d, div_d = dev(torch, mat), dev(torch, div)
    x, _, st = refine(simcore, torch, mat_d, div_d, shape, "bf16", 1e-7)
    assert st["converged"] == 1
    x2, _, st2 = refine(simcore, torch, mat_d, div_d, shape, "bf16", 1e-7, warm=x)
    assert st2["converged"] == 1 and st2["outer"] == 0
    assert torch.equal(x, x2)
    zero = torch.zeros_like(div_d)
    x3, _, st3 = refine(simcore, torch, mat_d, zero, shape, "fp16", 1e-7)
    assert st3["converged"] == 1 and st3["outer"] == 0 and not host(x3).any()


def test_learned_refine_smooth_rhs(simcore, torch_gpu):
    """The FP64 refinement around the learned flexible CG (V-cycle + trained
    U-Net) reaches 1e-7 on the smooth right-hand side above, in far fewer
    inner iterations than the FP32-storage Jacobi-PCG refinement, and the
    oracle's FP64 residual of its iterate agrees."""
    torch = torch_gpu
    from boilsim.fixtures import pot_mask
    from boilsim.learned import grf_torch
    from boilsim.unet import UNet

    shape = (110, 110, 70)
    mat = pot_mask(*shape)
    mat_d = dev(torch, mat)
    div_d, _ = grf_torch(torch, mat_d == 0, seed=5, ell=12.0)
    b = b32(host(div_d), mat)
    tol = 1e-7
    net = UNet.load()
    ws = torch.empty(simcore.fcg_refine_workspace_bytes(*shape, 2), dtype=torch.uint8, device="cuda")
    x = torch.zeros(shape, dtype=torch.float64, device="cuda")
    st = simcore.fcg_refine_solve(x.data_ptr(), div_d.data_ptr(), mat_d.data_ptr(), *shape, DX, DT, RHO, tol, 2,
                                  net.params.data_ptr(), net.packed.data_ptr(), ws.data_ptr(), ws.numel())
    torch.cuda.synchronize()
    rel2, _ = oracle.residual_f64(host(x), b, mat)
    _, _, sp = refine(simcore, torch, mat_d, div_d, shape, "fp32", tol)
    print(f"learned refine: {st}; oracle relres {rel2:.3e}; Jacobi-PCG refine: {sp}")
    assert st["converged"] == 1 and rel2 <= tol * 1.0001
    assert abs(rel2 - st["relres_true"]) <= 1e-9 * rel2
    assert st["inner_iterations"] * 4 < sp["inner_iterations"]
