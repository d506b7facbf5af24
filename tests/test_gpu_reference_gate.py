"""The reference's own 200-step gate case on the device
(python/tests/test_pressure_parity.py:159-224): single_carrot.yaml voxelised
at dx = 4 mm by build_pot_geometry (restated in oracle/pot_geometry.py),
N(0, 0.05) face velocities from default_rng(7), pressure_projection with
200 Jacobi iterations, dt = 0.01, rho = 997, through both of the reference's
Jacobi selections (the Warp default and BOILINGSIM_USE_RUST_PRESSURE=1).
The reference's gate is max_rel_diff < 1e-4 over fluid cells with scale
max(|p|, 1e-6) (test_pressure_parity.py:208-224); the device must match the
oracle bit for bit, and the committed digests of the oracle's outputs."""

from __future__ import annotations

import hashlib
import json
import os

import numpy as np
import pytest

import oracle
from helpers import bit_diff_report, host, same_bits
from oracle import pot_geometry as pg

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("rust_env", [False, True])
def test_reference_200_step_gate(simcore, torch_gpu, monkeypatch, rust_env):
    from boilsim import fluid

    gold = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                       "config_digests.json")))["reference_gate"]
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).view(np.uint8).tobytes()).hexdigest()
    mat, origin = pg.build_pot_mask(pg.GATE_DX)
    ux, uy, uz = pg.gate_velocities(mat.shape)
    assert sha(mat) == gold["sha256"]["mat"] and sha(uz) == gold["sha256"]["uz"]
    print(f"[gate] {mat.shape} rust_env={rust_env}", flush=True)
    monkeypatch.delenv("BOILINGSIM_PRESSURE_SOLVER", raising=False)
    if rust_env:
        monkeypatch.setenv("BOILINGSIM_USE_RUST_PRESSURE", "1")
    else:
        monkeypatch.delenv("BOILINGSIM_USE_RUST_PRESSURE", raising=False)
    grid = fluid.make_grid(mat, ux, uy, uz, pg.GATE_DX, origin=origin)
    ws = fluid.allocate_fluid_workspace(grid)
    cfg = fluid.ScenarioConfig()
    cfg.solver.pressure_max_iter = pg.GATE_ITERS
    max_div = fluid.pressure_projection(grid, ws, cfg, dt=pg.GATE_DT, rho=pg.GATE_RHO)
    p = host(grid.p)
    fx, fy, fz = ux.copy(), uy.copy(), uz.copy()
    p_ref, _, _, max_ref = oracle.projection(fx, fy, fz, mat, pg.GATE_DX, pg.GATE_DT, pg.GATE_RHO, pg.GATE_ITERS)
    fl = mat == pg.MAT_FLUID
    rel = np.abs(p - p_ref)[fl] / np.maximum(np.abs(p_ref[fl]), 1e-6)
    print(f"  max_rel_diff {rel.max():.3e} (gate 1e-4); max|div| {max_div:.6e} vs {max_ref:.6e}")
    assert rel.max() < 1e-4
    assert same_bits(p, p_ref), bit_diff_report(p, p_ref)
    assert sha(p) == gold["sha256"]["p"]
    for k, t in (("ux_after", grid.ux), ("uy_after", grid.uy), ("uz_after", grid.uz)):
        assert sha(host(t)) == gold["sha256"][k], k
    assert max_div == np.float32(max_ref) == np.float32(gold["max_div_after"])
