"""GPU parity of the Jacobi sweep and the fused n_iter solve against the
oracle (the reference's Warp arithmetic restated in C).

The reference gates one sweep at rtol 1e-5 / atol 1e-7
(python/tests/test_pressure_parity.py:150-156) and 200 sweeps at
max_rel_diff < 1e-4 (test_pressure_parity.py:208-224).  The HIP kernels are
built without contraction and with IEEE division, so these tests demand
bit-identical results, which implies both gates."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from helpers import (PARITY_CASES, bit_diff_report, dev, fluid_everywhere_edge_case, host,
                     parity_case, same_bits)

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("seed,shape", PARITY_CASES)
def test_one_step_bit_exact(simcore, torch_gpu, seed, shape):
    torch = torch_gpu
    mat, p_old, div_u, dx, dt, rho = parity_case(seed, shape)
    want = oracle.jacobi_step(p_old, div_u, mat, dx, dt, rho)
    p_new = torch.zeros(shape, dtype=torch.float32, device="cuda")
    p_old_d, div_d, mat_d = dev(torch, p_old), dev(torch, div_u), dev(torch, mat)
    nx, ny, nz = shape
    simcore.jacobi_pressure_step(p_new.data_ptr(), p_old_d.data_ptr(), div_d.data_ptr(),
                                 mat_d.data_ptr(), nx, ny, nz, dx, dt, rho, 0, 2)
    torch.cuda.synchronize()
    got = host(p_new)
    np.testing.assert_allclose(got, want, rtol=1e-5, atol=1e-7)
    assert same_bits(got, want), bit_diff_report(got, want)


@pytest.mark.parametrize("shape", [(9, 7, 13), (6, 5, 10), (5, 6, 8), (4, 4, 4)])
def test_one_step_fluid_on_domain_edge(simcore, torch_gpu, shape):
    """Odd nz (scalar path), nz % 4 == 2 (pairs) and nz % 4 == 0 (quads), with
    fluid cells on every face of the domain: out of bounds is air."""
    torch = torch_gpu
    mat, p_old, div_u = fluid_everywhere_edge_case(7, shape)
    dx, dt, rho = 2e-3, 1e-3, 958.0
    want = oracle.jacobi_step(p_old, div_u, mat, dx, dt, rho)
    p_new = torch.zeros(shape, dtype=torch.float32, device="cuda")
    p_old_d, div_d, mat_d = dev(torch, p_old), dev(torch, div_u), dev(torch, mat)
    simcore.jacobi_pressure_step(p_new.data_ptr(), p_old_d.data_ptr(), div_d.data_ptr(),
                                 mat_d.data_ptr(), *shape, dx, dt, rho)
    torch.cuda.synchronize()
    got = host(p_new)
    assert same_bits(got, want), bit_diff_report(got, want)


def test_one_step_misaligned_views(simcore, torch_gpu):
    """Buffers offset by one float (not 16-byte aligned) take the narrower
    vector path and still agree bit for bit."""
    torch = torch_gpu
    shape = (12, 10, 16)
    mat, p_old, div_u, dx, dt, rho = parity_case(5, shape)
    n = mat.size
    want = oracle.jacobi_step(p_old, div_u, mat, dx, dt, rho)
    big_p = torch.zeros(n + 1, dtype=torch.float32, device="cuda")
    big_o = torch.zeros(n + 1, dtype=torch.float32, device="cuda")
    big_d = torch.zeros(n + 1, dtype=torch.float32, device="cuda")
    big_m = torch.zeros(n + 1, dtype=torch.int32, device="cuda")
    big_o[1:] = dev(torch, p_old.ravel())
    big_d[1:] = dev(torch, div_u.ravel())
    big_m[1:] = dev(torch, mat.ravel())
    simcore.jacobi_pressure_step(big_p[1:].data_ptr(), big_o[1:].data_ptr(), big_d[1:].data_ptr(),
                                 big_m[1:].data_ptr(), *shape, dx, dt, rho)
    torch.cuda.synchronize()
    got = host(big_p[1:]).reshape(shape)
    assert same_bits(got, want), bit_diff_report(got, want)


@pytest.mark.parametrize("n_iter", [0, 1, 2, 3, 4, 7, 50])
def test_fused_solve_matches_ping_pong(simcore, torch_gpu, n_iter):
    """pressure_solve leaves p (and p_tmp) exactly as the reference's ping-pong
    driver does (fluid.py:688-704, jacobi_pressure.cu:228-270), for even and
    odd counts, starting from a non-zero guess with garbage at non-fluid cells."""
    torch = torch_gpu
    shape = (20, 18, 24)
    mat, p0, div_u, dx, dt, rho = parity_case(11, shape)
    p_ref = p0.copy()
    tmp_ref = np.full(shape, 3.25, np.float32)
    oracle.pressure_solve(p_ref, tmp_ref, div_u, mat, dx, dt, rho, n_iter)
    p = dev(torch, p0)
    tmp = dev(torch, np.full(shape, 3.25, np.float32))
    div_d, mat_d = dev(torch, div_u), dev(torch, mat)
    simcore.pressure_solve(p.data_ptr(), tmp.data_ptr(), div_d.data_ptr(), mat_d.data_ptr(),
                           *shape, dx, dt, rho, n_iter)
    torch.cuda.synchronize()
    assert same_bits(host(p), p_ref), bit_diff_report(host(p), p_ref)
    assert same_bits(host(tmp), tmp_ref), bit_diff_report(host(tmp), tmp_ref)


def _pot_case(name):
    from boilsim.fixtures import CONFIGS, DT, RHO, pot_mask, seeded_velocities

    cfg = CONFIGS[name]
    mat = pot_mask(*cfg.shape)
    ux, uy, uz = seeded_velocities(*cfg.shape, seed=7)
    div = oracle.divergence(ux, uy, uz, mat, cfg.dx)
    return cfg, mat, div, DT, RHO


@pytest.mark.parametrize("name", ["small", "dev"])
def test_200_sweeps_pot(simcore, torch_gpu, name):
    """200 sweeps from p = 0 on the pot (configs 64x64x48 and 110x110x70):
    bit-identical to the oracle, hence within the reference's
    max_rel_diff < 1e-4 gate; the error against the FP64 solution of the same
    system is reported."""
    torch = torch_gpu
    cfg, mat, div, dt, rho = _pot_case(name)
    p_ref = np.zeros(cfg.shape, np.float32)
    tmp_ref = np.zeros(cfg.shape, np.float32)
    oracle.pressure_solve(p_ref, tmp_ref, div, mat, cfg.dx, dt, rho, 200)
    p = torch.zeros(cfg.shape, dtype=torch.float32, device="cuda")
    tmp = torch.zeros(cfg.shape, dtype=torch.float32, device="cuda")
    div_d, mat_d = dev(torch, div), dev(torch, mat)
    simcore.pressure_solve(p.data_ptr(), tmp.data_ptr(), div_d.data_ptr(), mat_d.data_ptr(),
                           *cfg.shape, cfg.dx, dt, rho, 200)
    torch.cuda.synchronize()
    got = host(p)
    fluid = mat == 0
    scale = np.maximum(np.abs(p_ref[fluid]), 1e-6)
    max_rel = float((np.abs(got - p_ref)[fluid] / scale).max())
    assert max_rel < 1e-4
    assert same_bits(got, p_ref), bit_diff_report(got, p_ref)
    # Error statement against the FP64 solve of A p = b (b = -rho dx^2 div/dt).
    b = oracle.rhs_f64(div, mat, cfg.dx, dt, rho)
    x64, _ = oracle.pcg_f64(b, mat, tol=1e-12)
    err = np.linalg.norm((got - x64)[fluid]) / np.linalg.norm(x64[fluid])
    print(f"[{name}] 200 Jacobi sweeps: rel L2 error vs FP64 solution = {err:.4e} "
          f"(200 sweeps are far from converged; this is the method's error, equal for both)")


@pytest.mark.parametrize("shape", [(97, 83, 68), (40, 30, 56), (33, 29, 116), (9, 7, 4), (64, 440, 280)])
@pytest.mark.parametrize("n_iter", [5, 9, 30])
@pytest.mark.parametrize("edge", [False, True])
@pytest.mark.parametrize("pairs", ["1", "0"])
def test_two_sweep_pass_bit_exact(simcore, torch_gpu, monkeypatch, shape, n_iter, edge, pairs):
    """The fused driver runs its middle sweeps two per pass (nz % 4 == 0) when
    a sweep's working set exceeds the Infinity Cache; these grids are smaller,
    so BOILSIM_JACOBI_PAIR=1 forces that kernel on (=0 off): either way p and
    p_tmp end bit-identical to the reference's one-sweep ping-pong, on grids
    with partial (j, k) tiles, several i segments, and fluid on the domain
    edges (out-of-bounds neighbours are air)."""
    torch = torch_gpu
    monkeypatch.setenv("BOILSIM_JACOBI_PAIR", pairs)
    if edge:
        mat, p0, div_u = fluid_everywhere_edge_case(5, shape)
        dx, dt, rho = 0.002, 0.001, 997.0
    else:
        mat, p0, div_u, dx, dt, rho = parity_case(23, shape)
    p_ref = p0.copy()
    tmp_ref = np.full(shape, -1.5, np.float32)
    oracle.pressure_solve(p_ref, tmp_ref, div_u, mat, dx, dt, rho, n_iter)
    p = dev(torch, p0)
    tmp = dev(torch, np.full(shape, -1.5, np.float32))
    div_d, mat_d = dev(torch, div_u), dev(torch, mat)
    simcore.pressure_solve(p.data_ptr(), tmp.data_ptr(), div_d.data_ptr(), mat_d.data_ptr(),
                           *shape, dx, dt, rho, n_iter)
    torch.cuda.synchronize()
    assert same_bits(host(p), p_ref), bit_diff_report(host(p), p_ref)
    assert same_bits(host(tmp), tmp_ref), bit_diff_report(host(tmp), tmp_ref)


@pytest.mark.parametrize("n_iter", [1, 2, 3])
def test_projection_zero_divergence_cells_keep_positive_zero(simcore, torch_gpu, n_iter):
    """The projection's cold start (p = 0, first sweep without reading p) on
    fluid cells whose divergence is exactly +0 or -0: (s - rhs)/6 with s = +0
    is +0 for both, bit for bit as the oracle, not -0."""
    torch = torch_gpu
    rng = np.random.default_rng(4)
    shape = (12, 10, 16)
    mat = rng.choice([0, 1, 2], size=shape, p=[0.7, 0.2, 0.1]).astype(np.int32)
    nx, ny, nz = shape
    f = [(rng.standard_normal(s) * 0.05).astype(np.float32)
         for s in ((nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1))]
    # A block of zero faces, -0 on odd and +0 on even face indices along each
    # face's axis: cells there get div = (-0 - +0)/dx + ... = -0 (even cells)
    # or +0.
    for q, a in enumerate(f):
        blk = a[3:9, 2:9, 4:13]
        idx = np.indices(blk.shape)[q] + (3, 2, 4)[q]
        blk[...] = np.where(idx % 2 == 1, np.float32(-0.0), np.float32(0.0))
    div0 = oracle.divergence(*f, mat, 2e-3)
    zero = (mat == 0) & (div0 == 0)
    assert zero.sum() > 20 and (zero & np.signbit(div0)).sum() > 5 and (zero & ~np.signbit(div0)).sum() > 5
    ref = [a.copy() for a in f]
    p_r, pt_r, div_r, m_r = oracle.projection(*ref, mat, 2e-3, 0.01, 997.0, n_iter)
    d = [dev(torch, a) for a in f]
    md = dev(torch, mat)
    p = torch.full(shape, 3.0, dtype=torch.float32, device="cuda")
    pt = torch.full(shape, 4.0, dtype=torch.float32, device="cuda")
    div = torch.empty(shape, dtype=torch.float32, device="cuda")
    mb = torch.zeros(1, dtype=torch.int32, device="cuda")
    simcore.projection_jacobi(p.data_ptr(), pt.data_ptr(), div.data_ptr(), *(t.data_ptr() for t in d),
                              md.data_ptr(), *shape, 2e-3, 0.01, 997.0, n_iter, mb.data_ptr())
    torch.cuda.synchronize()
    for g, r in ((host(p), p_r), (host(pt), pt_r), (host(div), div_r), *((host(t), r) for t, r in zip(d, ref))):
        assert same_bits(g, r), bit_diff_report(g, r)


def test_config1_matches_committed_digests(simcore, torch_gpu):
    """Config 1 on the device (seed-7 velocities with the plume, 64x64x48):
    the divergence, 200 sweeps (p and p_tmp) and the fused projection hash to
    the committed digests of the oracle's outputs, bit for bit."""
    import hashlib
    import json
    import os

    torch = torch_gpu
    from boilsim.fixtures import CONFIGS, DT, RHO, pot_mask, seeded_velocities

    gold = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                       "config_digests.json")))["config1"]
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).view(np.uint8).tobytes()).hexdigest()
    cfg = CONFIGS["small"]
    mat = pot_mask(*cfg.shape)
    ux, uy, uz = seeded_velocities(*cfg.shape, seed=gold["seed"], plume=True)
    assert sha(mat) == gold["sha256"]["mat"] and sha(uz) == gold["sha256"]["uz"]
    d = [dev(torch, a) for a in (ux, uy, uz)]
    md = dev(torch, mat)
    div = torch.empty(cfg.shape, device="cuda")
    simcore.compute_divergence(div.data_ptr(), *(t.data_ptr() for t in d), md.data_ptr(), *cfg.shape, cfg.dx)
    p = torch.zeros(cfg.shape, device="cuda")
    pt = torch.zeros(cfg.shape, device="cuda")
    simcore.pressure_solve(p.data_ptr(), pt.data_ptr(), div.data_ptr(), md.data_ptr(), *cfg.shape, cfg.dx, DT, RHO,
                           gold["sweeps"])
    torch.cuda.synchronize()
    assert sha(host(div)) == gold["sha256"]["div"]
    assert sha(host(p)) == gold["sha256"]["p"] and sha(host(pt)) == gold["sha256"]["p_tmp"]
    p2 = torch.zeros(cfg.shape, device="cuda")
    pt2 = torch.zeros(cfg.shape, device="cuda")
    mb = torch.zeros(1, dtype=torch.int32, device="cuda")
    simcore.projection_jacobi(p2.data_ptr(), pt2.data_ptr(), div.data_ptr(), *(t.data_ptr() for t in d),
                              md.data_ptr(), *cfg.shape, cfg.dx, DT, RHO, gold["sweeps"], mb.data_ptr())
    torch.cuda.synchronize()
    for k, t in zip(("ux_after", "uy_after", "uz_after"), d):
        assert sha(host(t)) == gold["sha256"][k], k
    assert sha(host(div)) == gold["sha256"]["div_after"]
    assert float(host(mb).view(np.float32)[0]) == gold["max_div_after"]


@pytest.mark.parametrize("name", ["dev", "prod"])
def test_config2_matches_committed_digests(simcore, torch_gpu, name):
    """Config 2 on the device at 110x110x70 and 440x440x280 (seed-7 velocities
    with the plume): the divergence and 200 sweeps (p and p_tmp; at
    440x440x280 the two-sweeps-per-pass kernel, selected by size) hash to the
    committed digests of the oracle's outputs, bit for bit."""
    import hashlib
    import json
    import os

    torch = torch_gpu
    from boilsim.fixtures import CONFIGS, DT, RHO, pot_mask, seeded_velocities

    gold = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                                       "config_digests.json")))["config2"][name]
    sha = lambda a: hashlib.sha256(np.ascontiguousarray(a).view(np.uint8).tobytes()).hexdigest()
    cfg = CONFIGS[name]
    print(f"[config2] {name} {cfg.shape}", flush=True)
    mat = pot_mask(*cfg.shape)
    ux, uy, uz = seeded_velocities(*cfg.shape, seed=gold["seed"], plume=True)
    assert sha(uz) == gold["sha256"]["uz"]
    d = [dev(torch, a) for a in (ux, uy, uz)]
    del ux, uy, uz
    md = dev(torch, mat)
    div = torch.empty(cfg.shape, device="cuda")
    simcore.compute_divergence(div.data_ptr(), *(t.data_ptr() for t in d), md.data_ptr(), *cfg.shape, cfg.dx)
    p = torch.zeros(cfg.shape, device="cuda")
    pt = torch.zeros(cfg.shape, device="cuda")
    simcore.pressure_solve(p.data_ptr(), pt.data_ptr(), div.data_ptr(), md.data_ptr(), *cfg.shape, cfg.dx, DT, RHO,
                           gold["sweeps"])
    torch.cuda.synchronize()
    assert sha(host(div)) == gold["sha256"]["div"]
    assert sha(host(p)) == gold["sha256"]["p"]
    assert sha(host(pt)) == gold["sha256"]["p_tmp"]


def test_div6_exhaustive_on_device(simcore, torch_gpu):
    """The exact division by 6 of the Jacobi kernels equals the device's IEEE
    x / 6.0f for all 2^32 float bit patterns (bs_div6_check); the one-term
    control x * RN(1/6) differs somewhere, so the check can fail."""
    from sim_core import _lib

    torch = torch_gpu
    counts = torch.full((3,), 7, dtype=torch.int64, device="cuda")
    _lib.call("bs_div6_check", counts.data_ptr(), None)
    torch.cuda.synchronize()
    mism, fma_path, one_term = (int(v) for v in host(counts))
    print(f"div6 on device: {mism} mismatches, {fma_path} on the fma path, one-term form differs on {one_term}")
    assert mism == 0
    assert fma_path == 7 * 2**29
    assert one_term > 0
