"""GPU tests of the PCG drop-in (SimCore.pressure_solve_pcg), the true-residual
solver (SimCore.pcg_solve) and the reference PCG's building blocks.

Gates restated from python/tests/test_pressure_cg.py (sign, bsq, tolerance
control) plus parity with the oracle's reference-semantics PCG (same control
flow, FP32 vectors, FP64 scalars) and error statements against the FP64
oracle solve."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from helpers import bit_diff_report, dev, host, same_bits

pytestmark = pytest.mark.gpu


class PcgWs:
    """The reference's PCG workspace (python/boilingsim/fluid.py:557-582)."""

    def __init__(self, torch, shape):
        f = lambda *s: torch.zeros(s, dtype=torch.float32, device="cuda")
        self.b, self.r, self.z, self.p_search, self.Ap = (f(*shape) for _ in range(5))
        self.dot_ws = f(1024)
        self.scal = {k: f(1) for k in ("alpha", "beta", "rzold", "rznew", "bsq", "rsq", "pAp")}

    def args(self):
        s = self.scal
        return [self.b.data_ptr(), self.r.data_ptr(), self.z.data_ptr(), self.p_search.data_ptr(),
                self.Ap.data_ptr(), self.dot_ws.data_ptr(), s["alpha"].data_ptr(), s["beta"].data_ptr(),
                s["rzold"].data_ptr(), s["rznew"].data_ptr(), s["bsq"].data_ptr(), s["rsq"].data_ptr(),
                s["pAp"].data_ptr()]


def run_pcg(core, torch, mat, div, tol, max_iter, dx=0.002, dt=0.001, rho=997.0):
    shape = mat.shape
    ws = PcgWs(torch, shape)
    p = torch.full(shape, 7.0, dtype=torch.float32, device="cuda")  # ignored (cold start)
    mat_d, div_d = dev(torch, mat), dev(torch, div)
    iters = core.pressure_solve_pcg(p.data_ptr(), div_d.data_ptr(), mat_d.data_ptr(), *shape, dx, dt, rho,
                                    tol, max_iter, *ws.args(), 0, 2)
    torch.cuda.synchronize()
    return host(p), iters, ws


def test_gate1_sign_and_scale(simcore, torch_gpu):
    """div < 0 on a slab drives p > 0 there (test_pressure_cg.py:49-156)."""
    torch = torch_gpu
    nx, ny, nz = 8, 8, 16
    mat = np.zeros((nx, ny, nz), np.int32)
    mat[:, :, 0] = 2
    mat[:, :, -1] = 2
    div = np.zeros((nx, ny, nz), np.float32)
    div[:, :, 4:12] = -1.0e-3
    p, iters, _ = run_pcg(simcore, torch, mat, div, 1e-5, 200)
    assert (p[:, :, 4:12] > 0).all()
    typical_b = 997.0 * 0.002 * 0.002 * 1e-3 / 0.001
    assert p[:, :, 4:12].max() < 1e4 * typical_b
    assert 0 < iters <= 200


def test_gate2_bsq_matches_numpy(simcore, torch_gpu):
    """max_iter=0 still forms b.b (test_pressure_cg.py:166-232); FP64 here,
    so the gate's rel < 1e-4 holds with orders of magnitude to spare."""
    torch = torch_gpu
    nx, ny, nz = 12, 8, 10
    mat = np.zeros((nx, ny, nz), np.int32)
    div = (np.random.default_rng(1).standard_normal((nx, ny, nz)).astype(np.float32) * 1e-2)
    p, iters, ws = run_pcg(simcore, torch, mat, div, 1e-12, 0)
    scale = np.float32(-997.0) * np.float32(0.002) * np.float32(0.002) / np.float32(0.001)
    b = (div * scale).astype(np.float32).astype(np.float64)
    want = float((b * b).sum())
    got = float(host(ws.scal["bsq"])[0])
    assert iters == 0
    assert abs(got - want) / want < 1e-6
    assert np.all(p == 0.0)


def test_gate2b_zero_rhs_returns_zero(simcore, torch_gpu):
    """bsq == 0 -> p = 0, 0 iterations (pressure_solve_pcg.cu:151-155)."""
    torch = torch_gpu
    mat = np.zeros((6, 6, 6), np.int32)
    div = np.zeros((6, 6, 6), np.float32)
    p, iters, _ = run_pcg(simcore, torch, mat, div, 1e-5, 50)
    assert iters == 0 and np.all(p == 0.0)


def test_gate3_tol_controls_residual(simcore, torch_gpu):
    """Tighter tol -> smaller true residual (test_pressure_cg.py:275-343)."""
    torch = torch_gpu
    nx, ny, nz = 16, 16, 24
    mat = np.zeros((nx, ny, nz), np.int32)
    mat[:, :, 0] = 2
    mat[:, :, -1] = 2
    div = np.random.default_rng(42).standard_normal((nx, ny, nz)).astype(np.float32) * 5e-3
    p_loose, _, _ = run_pcg(simcore, torch, mat, div, 1e-2, 200)
    p_tight, _, _ = run_pcg(simcore, torch, mat, div, 1e-6, 200)
    b = oracle.rhs_f64(div, mat, 0.002, 0.001, 997.0)
    _, r_loose = oracle.residual_f64(p_loose, b, mat)
    _, r_tight = oracle.residual_f64(p_tight, b, mat)
    assert r_tight < 0.5 * r_loose
    assert r_tight < 1e-3


def _pot(name, seed=11):
    from boilsim.fixtures import CONFIGS, DT, RHO, pot_mask, seeded_velocities

    cfg = CONFIGS[name]
    mat = pot_mask(*cfg.shape)
    ux, uy, uz = seeded_velocities(*cfg.shape, seed=seed, plume=False)
    div = oracle.divergence(ux, uy, uz, mat, cfg.dx)
    return cfg, mat, div, DT, RHO


@pytest.mark.parametrize("name,tol", [("small", 1e-6), ("small", 1e-3), ("dev", 1e-6)])
def test_dropin_pcg_matches_oracle(simcore, torch_gpu, name, tol):
    """Same control flow and the same FP32 vector arithmetic as the oracle's
    reference-semantics PCG (FP64 scalars in both): the same iteration count
    and a bit-identical solution; both errors against the FP64 solution are
    stated."""
    torch = torch_gpu
    cfg, mat, div, dt, rho = _pot(name)
    p, iters, _ = run_pcg(simcore, torch, mat, div, tol, 5000, cfg.dx, dt, rho)
    p_ref, it_ref, _ = oracle.pcg_ref(div, mat, cfg.dx, dt, rho, tol, 5000)
    b = oracle.rhs_f64(div, mat, cfg.dx, dt, rho)
    x64, _ = oracle.pcg_f64(b, mat, tol=1e-12)
    fl = mat == 0
    nrm = np.linalg.norm(x64[fl])
    d_ref = np.linalg.norm((p - p_ref)[fl]) / nrm
    e_gpu = np.linalg.norm((p - x64)[fl]) / nrm
    e_ref = np.linalg.norm((p_ref - x64)[fl]) / nrm
    print(f"[{name} tol={tol}] iters gpu={iters} oracle={it_ref}; |gpu-oracle|/|x64|={d_ref:.3e}; "
          f"err vs FP64: gpu={e_gpu:.3e} oracle={e_ref:.3e}")
    assert iters == it_ref
    assert same_bits(p, p_ref), bit_diff_report(p, p_ref)


@pytest.mark.parametrize("name", ["small", "dev"])
def test_pcg_solve_true_residual(simcore, torch_gpu, name):
    torch = torch_gpu
    cfg, mat, div, dt, rho = _pot(name)
    shape = cfg.shape
    nbytes = simcore.pcg_workspace_bytes(*shape)
    ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
    x = torch.full(shape, 3.0, dtype=torch.float32, device="cuda")
    mat_d, div_d = dev(torch, mat), dev(torch, div)
    st = simcore.pcg_solve(x.data_ptr(), div_d.data_ptr(), mat_d.data_ptr(), *shape, cfg.dx, dt, rho,
                           1e-6, 20000, ws.data_ptr(), nbytes)
    torch.cuda.synchronize()
    xs = host(x)
    b = oracle.rhs_f64(div, mat, cfg.dx, dt, rho)
    rel2, _ = oracle.residual_f64(xs, b, mat)
    x64, it64 = oracle.pcg_f64(b, mat, tol=1e-12)
    fl = mat == 0
    err = np.linalg.norm((xs - x64)[fl]) / np.linalg.norm(x64[fl])
    print(f"[{name}] pcg_solve: {st}; oracle true relres {rel2:.3e}; err vs FP64 {err:.3e}")
    assert st["converged"] == 1
    assert st["relres_true"] <= 1e-6
    assert rel2 <= 1.05e-6
    assert abs(rel2 - st["relres_true"]) <= 1e-3 * rel2 + 1e-9
    # Warm start from the converged answer: nothing left to do.
    st2 = simcore.pcg_solve(x.data_ptr(), div_d.data_ptr(), mat_d.data_ptr(), *shape, cfg.dx, dt, rho,
                            1e-6, 20000, ws.data_ptr(), nbytes, warm_start=True)
    assert st2["iterations"] == 0 and st2["converged"] == 1


def test_building_blocks(simcore, torch_gpu):
    """The reference PCG's kernels one at a time: SpMV known answers
    (p = i^2 -> Ap = -2 and p linear -> Ap = 0 on interior fluid,
    test_pressure_cg.py:62-70), the diagonal preconditioner, FP64 dot, axpys."""
    torch = torch_gpu
    from sim_core import _lib

    nx, ny, nz = 10, 9, 12
    mat = np.zeros((nx, ny, nz), np.int32)
    ii = np.arange(nx, dtype=np.float32)[:, None, None] * np.ones((nx, ny, nz), np.float32)
    mat_d = dev(torch, mat)
    out = torch.zeros((nx, ny, nz), dtype=torch.float32, device="cuda")
    for p_np, want in ((ii * ii, -2.0), (ii + 2 * np.arange(ny)[None, :, None] + 3 * np.arange(nz), 0.0)):
        p_np = np.ascontiguousarray(p_np, dtype=np.float32)
        p_d = dev(torch, p_np)
        _lib.call("bs_laplacian_spmv_launch", out.data_ptr(), p_d.data_ptr(), mat_d.data_ptr(), nx, ny, nz, 0, 2)
        torch.cuda.synchronize()
        got = host(out)[1:-1, 1:-1, 1:-1]
        assert np.all(got == np.float32(want))
        # Whole field against the FP64 oracle operator.
        ref = oracle.laplacian_f64(p_np, mat)
        np.testing.assert_allclose(host(out), ref, rtol=1e-6, atol=1e-4)
    # Diagonal preconditioner on a random mask (out of bounds = air).
    rng = np.random.default_rng(3)
    m2 = rng.choice([0, 1, 2, 3], size=(nx, ny, nz), p=[0.6, 0.2, 0.1, 0.1]).astype(np.int32)
    r = rng.standard_normal((nx, ny, nz)).astype(np.float32)
    d64 = oracle.diag(m2)
    d32 = np.where(d64 > 0, d64, 1).astype(np.float32)
    want = np.where(d64 == 0, np.float32(0), np.where(d64 < 0, r, (r / d32).astype(np.float32))).astype(np.float32)
    r_d, m2_d = dev(torch, r), dev(torch, m2)
    _lib.call("bs_diag_inverse_apply_launch", out.data_ptr(), r_d.data_ptr(), m2_d.data_ptr(), nx, ny, nz, 0, 2)
    torch.cuda.synchronize()
    assert np.array_equal(host(out), want)
    # FP64 dot.
    x = rng.standard_normal(100003).astype(np.float32)
    y = rng.standard_normal(100003).astype(np.float32)
    wsd = torch.zeros(1024, dtype=torch.float32, device="cuda")
    res = torch.zeros(1, dtype=torch.float32, device="cuda")
    res64 = torch.zeros(1, dtype=torch.float64, device="cuda")
    xd, yd = dev(torch, x), dev(torch, y)
    _lib.call("bs_dot_launch_f64", xd.data_ptr(), yd.data_ptr(), x.size, wsd.data_ptr(), res.data_ptr(),
              res64.data_ptr())
    torch.cuda.synchronize()
    want64 = float(np.dot(x.astype(np.float64), y.astype(np.float64)))
    assert abs(float(host(res64)[0]) - want64) <= 1e-12 * np.abs(x.astype(np.float64) * y).sum()
    assert float(host(res)[0]) == np.float32(float(host(res64)[0]))
    # axpy with a device scalar, negated; scaled_axpy in place; divide; zero.
    a = torch.tensor([0.75], dtype=torch.float32, device="cuda")
    yy = yd.clone()
    _lib.call("bs_axpy_device_launch", yy.data_ptr(), xd.data_ptr(), a.data_ptr(), x.size, 1)
    torch.cuda.synchronize()
    assert np.array_equal(host(yy), (np.float32(-0.75) * x + y).astype(np.float32))
    _lib.call("bs_scaled_axpy_launch", yy.data_ptr(), xd.data_ptr(), yy.data_ptr(), a.data_ptr(), x.size)
    torch.cuda.synchronize()
    want_s = (x + np.float32(0.75) * (np.float32(-0.75) * x + y).astype(np.float32)).astype(np.float32)
    assert np.array_equal(host(yy), want_s)
    num = torch.tensor([3.0], dtype=torch.float32, device="cuda")
    den = torch.tensor([0.0], dtype=torch.float32, device="cuda")
    _lib.call("bs_divide_scalars_launch", num.data_ptr(), den.data_ptr(), a.data_ptr())
    torch.cuda.synchronize()
    assert float(host(a)[0]) == 0.0
    _lib.call("bs_zero_buffer_launch", yy.data_ptr(), x.size)
    torch.cuda.synchronize()
    assert not host(yy).any()


@pytest.mark.parametrize("name", ["small", "dev"])
def test_pcg_solve_matches_oracle_true_residual_pcg(simcore, torch_gpu, name):
    """bs_pcg_solve against the oracle's restatement of the same algorithm
    (oracle.pcg_true: every-iteration test, FP64 true residual, residual
    replacement): the same iteration and restart counts, and solutions whose
    difference is far below their error against the FP64 solution."""
    torch = torch_gpu
    cfg, mat, div, dt, rho = _pot(name)
    shape = cfg.shape
    ws = torch.empty(simcore.pcg_workspace_bytes(*shape), dtype=torch.uint8, device="cuda")
    x = torch.zeros(shape, dtype=torch.float32, device="cuda")
    mat_d, div_d = dev(torch, mat), dev(torch, div)
    st = simcore.pcg_solve(x.data_ptr(), div_d.data_ptr(), mat_d.data_ptr(), *shape, cfg.dx, dt, rho, 1e-6, 20000,
                           ws.data_ptr(), ws.numel())
    torch.cuda.synchronize()
    xo, it, rst, rel = oracle.pcg_true(div, mat, cfg.dx, dt, rho, 1e-6, 20000)
    xg = host(x)
    b = oracle.rhs_f64(div, mat, cfg.dx, dt, rho)
    x64, _ = oracle.pcg_f64(b, mat, tol=1e-12)
    fl = mat == 0
    nrm = np.linalg.norm(x64[fl])
    d = np.linalg.norm((xg - xo)[fl]) / nrm
    e = np.linalg.norm((xg - x64)[fl]) / nrm
    print(f"[{name}] gpu {st}; oracle {it} its {rst} restarts relres {rel:.3e}; |gpu-oracle| {d:.3e}, "
          f"bits equal {same_bits(xg, xo)}; error vs FP64 {e:.3e}")
    assert st["iterations"] == it and st["restarts"] == rst
    assert abs(st["relres_true"] - rel) <= 1e-3 * rel
    assert d < 1e-3 * e
