"""GPU test of bs_pcg_solve_noalloc, the PCG entry point that allocates
nothing (its workspace is the library's fixed per-device scratch array).

Called through the C ABI (ctypes) and checked against bs_pcg_solve with a
caller workspace on the same inputs: the same kernels run in the same order,
so the iterate must be bit-identical and the statistics equal.  Two calls in a
row, the second on a smaller grid, so stale scratch contents from the first
call would show; the oracle's FP64 residual judges both solutions."""

from __future__ import annotations

import ctypes

import numpy as np
import pytest

import oracle
from helpers import bit_diff_report, dev, host, same_bits

pytestmark = pytest.mark.gpu

TOL = 1e-6


def _case(name, seed):
    from boilsim.fixtures import CONFIGS, DT, RHO, pot_mask, seeded_velocities

    cfg = CONFIGS[name]
    mat = pot_mask(*cfg.shape)
    ux, uy, uz = seeded_velocities(*cfg.shape, seed=seed, plume=False)
    div = oracle.divergence(ux, uy, uz, mat, cfg.dx)
    return cfg.shape, cfg.dx, DT, RHO, mat, div


def _solve(lib, torch, fn, shape, dx, dt, rho, mat_d, div_d, ws=None, warm=0, x0=None):
    from sim_core import _lib

    x = torch.full(shape, 5.0, dtype=torch.float32, device="cuda") if x0 is None else x0.clone()
    st = _lib.SolveStats()
    args = [x.data_ptr(), div_d.data_ptr(), mat_d.data_ptr(), *shape, dx, dt, rho, 0, 2, TOL, 20000, warm, 4]
    if ws is not None:
        args += [ws.data_ptr(), ws.numel()]
    rc = getattr(lib, fn)(*args, None, ctypes.byref(st))
    torch.cuda.synchronize()
    assert rc == 0, f"{fn} returned hipError_t {rc}"
    return x, st.as_dict()


def test_noalloc_matches_workspace_solver_two_calls(simcore, torch_gpu):
    from sim_core import _lib

    torch = torch_gpu
    lib = _lib.load()
    for name, seed in (("dev", 11), ("small", 21)):  # 110x110x70, then 64x64x48
        shape, dx, dt, rho, mat, div = _case(name, seed)
        print(f"[no-alloc] {name} {shape} seed {seed}", flush=True)
        mat_d, div_d = dev(torch, mat), dev(torch, div)
        nbytes = lib.bs_pcg_workspace_bytes(*shape)
        assert nbytes <= lib.bs_pcg_noalloc_scratch_bytes()
        ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        x_ref, st_ref = _solve(lib, torch, "bs_pcg_solve", shape, dx, dt, rho, mat_d, div_d, ws=ws)
        x_na, st_na = _solve(lib, torch, "bs_pcg_solve_noalloc", shape, dx, dt, rho, mat_d, div_d)
        print(f"  workspace: {st_ref}\n  no-alloc:  {st_na}", flush=True)
        assert same_bits(host(x_na), host(x_ref)), bit_diff_report(host(x_na), host(x_ref))
        assert st_na == st_ref
        assert st_na["converged"] == 1
        b = oracle.rhs_f64(div, mat, dx, dt, rho)
        rel2, _ = oracle.residual_f64(host(x_na), b, mat)
        assert rel2 <= 1.05 * TOL
        # Warm start from a perturbed solution: both entry points again agree.
        x0 = x_ref * 0.5
        xw_ref, stw_ref = _solve(lib, torch, "bs_pcg_solve", shape, dx, dt, rho, mat_d, div_d, ws=ws, warm=1, x0=x0)
        xw_na, stw_na = _solve(lib, torch, "bs_pcg_solve_noalloc", shape, dx, dt, rho, mat_d, div_d, warm=1, x0=x0)
        assert same_bits(host(xw_na), host(xw_ref)), bit_diff_report(host(xw_na), host(xw_ref))
        assert stw_na == stw_ref and stw_na["converged"] == 1


def test_noalloc_rejects_grids_beyond_the_scratch(simcore, torch_gpu):
    from sim_core import _lib

    lib = _lib.load()
    st = _lib.SolveStats()
    # 440x440x280 needs ~0.92 GB of workspace: refused before any work is queued
    # (the pointers are never dereferenced).
    rc = lib.bs_pcg_solve_noalloc(None, None, None, 440, 440, 280, 5e-4, 1e-3, 958.0, 0, 2, TOL, 10, 0, 4, None,
                                  ctypes.byref(st))
    assert rc == 1  # hipErrorInvalidValue
    with pytest.raises(ValueError, match="fixed scratch array"):
        simcore.pcg_solve_noalloc(0, 0, 0, 440, 440, 280, 5e-4, 1e-3, 958.0, TOL, 10)
