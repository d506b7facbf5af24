"""Pins the CPU oracle (oracle/cpu_ref.c) before it is trusted as the checker.

The reference commits no golden vectors for this path and Warp cannot be
imported here, so every oracle function is pinned by checks a plausible
mistake would fail: an independent numpy restatement (bit-exact), closed forms
and known answers from the reference's own tests, brute-force loops on tiny
grids, scipy solves of the assembled matrix, and the reference's PCG gates."""

from __future__ import annotations

import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle
from helpers import PARITY_CASES, fluid_everywhere_edge_case, parity_case, same_bits
from oracle import np_ref


# ---------------------------------------------------------------- Jacobi sweep
@pytest.mark.parametrize("seed,shape", PARITY_CASES)
def test_jacobi_c_matches_numpy_bit_exact(seed, shape):
    mat, p, d, dx, dt, rho = parity_case(seed, shape)
    a = oracle.jacobi_step(p, d, mat, dx, dt, rho)
    b = np_ref.jacobi_step(p, d, mat, dx, dt, rho)
    assert same_bits(a, b)


def test_jacobi_edge_fluid_is_air_bounded():
    mat, p, d = fluid_everywhere_edge_case(5, (7, 6, 9))
    a = oracle.jacobi_step(p, d, mat, 2e-3, 1e-3, 958.0)
    b = np_ref.jacobi_step(p, d, mat, 2e-3, 1e-3, 958.0)
    assert same_bits(a, b)


def test_jacobi_brute_force_tiny():
    """Pure-Python loop over fluid.py:243-276 on a 5x4x6 grid, float32 scalars."""
    rng = np.random.default_rng(9)
    shape = (5, 4, 6)
    mat = rng.choice([0, 1, 2, 3], size=shape, p=[0.6, 0.15, 0.15, 0.1]).astype(np.int32)
    p = rng.standard_normal(shape).astype(np.float32)
    d = (rng.standard_normal(shape) * 0.1).astype(np.float32)
    dx, dt, rho = np.float32(0.002), np.float32(0.001), np.float32(958.0)
    want = np.zeros(shape, np.float32)
    nx, ny, nz = shape
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                if mat[i, j, k] != 0:
                    continue
                ps = p[i, j, k]
                vals = []
                for di, dj, dk in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
                    a, b, c = i + di, j + dj, k + dk
                    if not (0 <= a < nx and 0 <= b < ny and 0 <= c < nz) or mat[a, b, c] == 2:
                        vals.append(np.float32(0.0))
                    elif mat[a, b, c] == 0:
                        vals.append(p[a, b, c])
                    else:
                        vals.append(ps)
                s = np.float32(vals[0] + vals[1])
                for v in vals[2:]:
                    s = np.float32(s + v)
                rhs = np.float32(np.float32(np.float32(rho * dx) * dx) * d[i, j, k])
                rhs = np.float32(rhs / dt)
                want[i, j, k] = np.float32(np.float32(s - rhs) / np.float32(6.0))
    got = oracle.jacobi_step(p, d, mat, float(dx), float(dt), float(rho))
    assert same_bits(got, want)


def test_jacobi_known_answers():
    """All-fluid interior, div = 0: p = i^2 gives one sweep p + 2/6 (A p = -2),
    p linear is a fixed point (A p = 0) — test_pressure_cg.py:62-70."""
    shape = (8, 7, 9)
    mat = np.full(shape, 2, np.int32)
    mat[1:-1, 1:-1, 1:-1] = 0
    ii = np.arange(shape[0], dtype=np.float32)[:, None, None] * np.ones(shape, np.float32)
    zero = np.zeros(shape, np.float32)
    out = oracle.jacobi_step(ii * ii, zero, mat, 1e-3, 1e-3, 1000.0)
    inner = (slice(2, -2), slice(2, -2), slice(2, -2))  # neighbours all fluid
    np.testing.assert_allclose(out[inner], (ii * ii)[inner] + np.float32(2.0 / 6.0), rtol=1e-6)
    lin = ii + 2 * np.arange(shape[1], dtype=np.float32)[None, :, None] + 3 * np.arange(shape[2], dtype=np.float32)
    lin = np.ascontiguousarray(lin, dtype=np.float32)
    out = oracle.jacobi_step(lin, zero, mat, 1e-3, 1e-3, 1000.0)
    assert np.array_equal(out[inner], lin[inner])
    # Non-fluid cells always 0; a wrong sign on rhs flips this mean.
    d = np.full(shape, -1e-3, np.float32)
    out = oracle.jacobi_step(zero, d, mat, 2e-3, 1e-3, 958.0)
    assert np.all(out[mat != 0] == 0) and np.all(out[mat == 0] > 0)


def test_jacobi_fixed_point_solves_system():
    """A converged sweep is a fixed point of A p = b with b = -rho dx^2 div/dt:
    the scipy FP64 solution changes by only FP32 rounding under one sweep."""
    from boilsim.fixtures import pot_mask

    mat = pot_mask(20, 20, 16)
    rng = np.random.default_rng(4)
    d = (rng.standard_normal(mat.shape) * 0.01).astype(np.float32)
    dx, dt, rho = 2e-3, 1e-3, 958.0
    A = np_ref.laplacian_csr(mat)
    b = oracle.rhs_f64(d, mat, dx, dt, rho)
    fl = (mat == 0).ravel()
    x = np.zeros(mat.size)
    x[fl] = spla.spsolve(A[fl][:, fl].tocsc(), b.ravel()[fl])
    x = x.reshape(mat.shape)
    out = oracle.jacobi_step(x.astype(np.float32), d, mat, dx, dt, rho)
    scale = np.abs(x).max()
    assert np.abs(out - x).max() < 1e-5 * scale


def test_ping_pong_odd_even():
    """N and N+1 sweeps differ by exactly one sweep; the result is in p for
    either parity (fluid.py:688-704)."""
    mat, p0, d, dx, dt, rho = parity_case(1, (16, 16, 16))
    for n in (1, 2, 5, 6):
        p, q = p0.copy(), np.zeros_like(p0)
        oracle.pressure_solve(p, q, d, mat, dx, dt, rho, n)
        p1, q1 = p0.copy(), np.zeros_like(p0)
        oracle.pressure_solve(p1, q1, d, mat, dx, dt, rho, n + 1)
        assert same_bits(p1, oracle.jacobi_step(p, d, mat, dx, dt, rho))
        ref = p0.copy()
        for _ in range(n):
            ref = np_ref.jacobi_step(ref, d, mat, dx, dt, rho)
        assert same_bits(p, ref)


# ------------------------------------------------- divergence and gradient
def _faces(rng, shape, scale=0.01):
    nx, ny, nz = shape
    return (rng.standard_normal((nx + 1, ny, nz)).astype(np.float32) * np.float32(scale),
            rng.standard_normal((nx, ny + 1, nz)).astype(np.float32) * np.float32(scale),
            rng.standard_normal((nx, ny, nz + 1)).astype(np.float32) * np.float32(scale))


def test_divergence_c_matches_numpy_and_closed_forms():
    rng = np.random.default_rng(2)
    shape = (9, 8, 11)
    mat = rng.choice([0, 1, 2], size=shape, p=[0.7, 0.2, 0.1]).astype(np.int32)
    ux, uy, uz = _faces(rng, shape)
    assert same_bits(oracle.divergence(ux, uy, uz, mat, 2e-3), np_ref.divergence(ux, uy, uz, mat, 2e-3))
    # Linear field u = (a x, b y, c z) on faces -> div = a + b + c on fluid, 0 elsewhere.
    nx, ny, nz = shape
    dx = 0.5
    ux = np.broadcast_to((np.arange(nx + 1) * dx * 2.0)[:, None, None], (nx + 1, ny, nz)).astype(np.float32)
    uy = np.broadcast_to((np.arange(ny + 1) * dx * 3.0)[None, :, None], (nx, ny + 1, nz)).astype(np.float32)
    uz = np.broadcast_to((np.arange(nz + 1) * dx * -1.5)[None, None, :], (nx, ny, nz + 1)).astype(np.float32)
    div = oracle.divergence(ux, uy, uz, mat, dx)
    assert np.all(div[mat == 0] == np.float32(3.5)) and np.all(div[mat != 0] == 0)


def test_gradient_c_matches_numpy_and_skip_rules():
    rng = np.random.default_rng(3)
    shape = (8, 9, 10)
    mat = rng.choice([0, 1, 2], size=shape, p=[0.7, 0.2, 0.1]).astype(np.int32)
    ux, uy, uz = _faces(rng, shape)
    p = rng.standard_normal(shape).astype(np.float32)
    a = [u.copy() for u in (ux, uy, uz)]
    oracle.gradient(*a, p, mat, 2e-3, 1e-3, 958.0)
    b = np_ref.gradient(ux, uy, uz, p, mat, 2e-3, 1e-3, 958.0)
    for x, y in zip(a, b):
        assert same_bits(x, y)
    # Boundary faces and faces touching a non-fluid cell are untouched.
    assert same_bits(a[0][0], ux[0]) and same_bits(a[0][-1], ux[-1])
    fl = mat == 0
    untouched = ~(fl[1:] & fl[:-1])
    assert same_bits(a[0][1:-1][untouched], ux[1:-1][untouched])
    # p linear in i -> constant correction (dt/rho)*slope/dx on fluid-fluid x faces.
    p_lin = np.ascontiguousarray(np.broadcast_to(np.arange(shape[0], dtype=np.float32)[:, None, None] * 4.0, shape))
    ux0 = np.zeros((shape[0] + 1,) + shape[1:], np.float32)
    a = [ux0.copy(), np.zeros_like(uy), np.zeros_like(uz)]
    oracle.gradient(*a, p_lin, mat, 0.5, 0.25, 2.0)
    want = -np.float32(np.float32(np.float32(0.25) / np.float32(2.0)) * np.float32(4.0)) / np.float32(0.5)
    inner = a[0][1:-1][fl[1:] & fl[:-1]]
    assert np.all(inner == np.float32(want))


def test_projection_physics_gates():
    """fluid.py:612-734 through the oracle: walled box with no-slip (the
    reference's test_fluid.py:118-144 drives 1500 sweeps on 16^3 and wants a
    >= 1e3 reduction of max|div|) and a closed box at rest (max_div < 1e-6,
    test_fluid.py:218-235)."""
    shape = (16, 16, 16)
    mat = np.full(shape, 1, np.int32)
    mat[1:-1, 1:-1, 1:-1] = 0
    rng = np.random.default_rng(42)
    ux, uy, uz = _faces(rng, shape, 0.1)
    # enforce_no_slip (fluid.py:111-162): zero every face touching a non-fluid cell.
    fl = mat == 0
    ux[1:-1][~(fl[1:] & fl[:-1])] = 0
    ux[0] = ux[-1] = 0
    uy[:, 1:-1][~(fl[:, 1:] & fl[:, :-1])] = 0
    uy[:, 0] = uy[:, -1] = 0
    uz[:, :, 1:-1][~(fl[:, :, 1:] & fl[:, :, :-1])] = 0
    uz[:, :, 0] = uz[:, :, -1] = 0
    before = float(np.abs(oracle.divergence(ux, uy, uz, mat, 0.002)).max())
    _, _, _, after = oracle.projection(ux, uy, uz, mat, 0.002, 0.01, 997.0, 1500)
    assert before > 1.0 and after < 1e-3 and after < before * 1e-3
    z = [np.zeros_like(u) for u in (ux, uy, uz)]
    _, _, _, m = oracle.projection(*z, mat, 0.002, 0.01, 997.0, 200)
    assert m < 1e-6 and max(np.abs(u).max() for u in z) < 1e-6


# ------------------------------------------------------------- linear algebra
def test_laplacian_matches_assembled_matrix():
    rng = np.random.default_rng(5)
    shape = (7, 8, 9)
    mat = rng.choice([0, 1, 2, 3], size=shape, p=[0.6, 0.15, 0.15, 0.1]).astype(np.int32)
    x = rng.standard_normal(shape)
    y = oracle.laplacian_f64(x, mat)
    A = np_ref.laplacian_csr(mat)
    np.testing.assert_allclose(y.ravel(), A @ x.ravel(), rtol=1e-13, atol=1e-13)
    # Diagonal of the assembled matrix = oracle's Jacobi diagonal on fluid rows.
    d = oracle.diag(mat).ravel()
    fl = (mat == 0).ravel()
    dA = A.diagonal()
    assert np.all(np.where(d[fl] > 0, d[fl], 0) == np.where(dA[fl] > 0, dA[fl], 0))


def test_pcg_f64_matches_scipy():
    from boilsim.fixtures import pot_mask

    mat = pot_mask(24, 24, 18)
    rng = np.random.default_rng(6)
    b = np.where(mat == 0, rng.standard_normal(mat.shape), 0.0)
    x, it = oracle.pcg_f64(b, mat, tol=1e-13)
    A = np_ref.laplacian_csr(mat)
    fl = (mat == 0).ravel()
    xs = spla.spsolve(A[fl][:, fl].tocsc(), b.ravel()[fl])
    np.testing.assert_allclose(x.ravel()[fl], xs, rtol=1e-9, atol=1e-9 * np.abs(xs).max())
    r2, _ = oracle.residual_f64(x, b, mat)
    assert r2 < 1e-12 and it > 10


def test_pcg_ref_gates():
    """The reference's PCG gates against the oracle PCG: sign
    (test_pressure_cg.py:49-156), b.b (166-232), tolerance control with the
    CPU residual oracle (275-343)."""
    mat = np.zeros((8, 8, 16), np.int32)
    mat[:, :, 0] = 2
    mat[:, :, -1] = 2
    div = np.zeros(mat.shape, np.float32)
    div[:, :, 4:12] = -1e-3
    p, it, _ = oracle.pcg_ref(div, mat, 0.002, 0.001, 997.0, 1e-5, 200)
    assert (p[:, :, 4:12] > 0).all() and 0 < it <= 200
    mat = np.zeros((12, 8, 10), np.int32)
    div = (np.random.default_rng(1).standard_normal(mat.shape).astype(np.float32) * 1e-2)
    _, it, bsq = oracle.pcg_ref(div, mat, 0.002, 0.001, 997.0, 1e-12, 0)
    b = (div * np.float32(-997.0 * 0.002 * 0.002 / 0.001)).astype(np.float64)
    assert it == 0 and abs(bsq - (b * b).sum()) / (b * b).sum() < 1e-6
    mat = np.zeros((16, 16, 24), np.int32)
    mat[:, :, 0] = 2
    mat[:, :, -1] = 2
    div = np.random.default_rng(42).standard_normal(mat.shape).astype(np.float32) * 5e-3
    bb = oracle.rhs_f64(div, mat, 0.002, 0.001, 997.0)
    pl, _, _ = oracle.pcg_ref(div, mat, 0.002, 0.001, 997.0, 1e-2, 200)
    pt, it_t, _ = oracle.pcg_ref(div, mat, 0.002, 0.001, 997.0, 1e-6, 200)
    _, rl = oracle.residual_f64(pl, bb, mat)
    _, rt = oracle.residual_f64(pt, bb, mat)
    assert rt < 0.5 * rl and rt < 1e-3
    # Converged count follows the reference cadence: iter >= 8, iter % 5 == 0, +1.
    assert it_t % 5 == 1 and it_t >= 9


def test_residual_oracle_matches_reference_restatement():
    """residual_f64's inf-norm equals the reference's pure-Python CPU residual
    (test_pressure_cg.py:295-325) restated as loops."""
    rng = np.random.default_rng(8)
    shape = (5, 6, 4)
    mat = rng.choice([0, 1, 2], size=shape, p=[0.7, 0.15, 0.15]).astype(np.int32)
    div = rng.standard_normal(shape).astype(np.float32) * 5e-3
    p = rng.standard_normal(shape).astype(np.float32)
    dx, dt, rho = 0.002, 0.001, 997.0
    b = -rho * dx * dx * div.astype(np.float64) / dt
    b[mat != 0] = 0
    Ap = np.zeros(shape)
    nx, ny, nz = shape
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                if mat[i, j, k] != 0:
                    continue
                s = 0.0
                for di, dj, dk in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
                    a, bq, c = i + di, j + dj, k + dk
                    if 0 <= a < nx and 0 <= bq < ny and 0 <= c < nz:
                        m = mat[a, bq, c]
                        s += p[a, bq, c] if m == 0 else (0.0 if m == 2 else p[i, j, k])
                Ap[i, j, k] = 6 * p[i, j, k] - s
    want = np.abs(Ap - b).max() / np.abs(b).max()
    _, got = oracle.residual_f64(p, oracle.rhs_f64(div, mat, dx, dt, rho), mat)
    assert abs(got - want) <= 1e-6 * want


def test_pcg_true_residual_semantics():
    """oracle.pcg_true (the semantics of bs_pcg_solve): k iterations with no
    test reached are the reference-semantics PCG's k iterations bit for bit
    (the same passes, only the stopping test differs); to tol it stops at a
    true FP64 relative residual <= tol (the oracle's own residual), restarting
    from b - A x when the recursive residual has drifted; the FP64 PCG's
    solution of the same system agrees to what that residual allows."""
    from boilsim.fixtures import CONFIGS, DT, RHO, pot_mask, seeded_velocities

    cfg = CONFIGS["small"]
    mat = pot_mask(*cfg.shape)
    ux, uy, uz = seeded_velocities(*cfg.shape, seed=11, plume=False)
    div = oracle.divergence(ux, uy, uz, mat, cfg.dx)
    for k in (1, 7, 40):
        xt, it, rst, _ = oracle.pcg_true(div, mat, cfg.dx, DT, RHO, 0.0, k, max_restarts=0)
        xr, itr, _ = oracle.pcg_ref(div, mat, cfg.dx, DT, RHO, 0.0, k)
        assert it == itr == k and rst == 0
        assert np.array_equal(xt.view(np.uint32), xr.view(np.uint32))
    # b as the solver forms it: (-rho dx dx / dt) div rounded to FP32 per cell
    f = np.float32
    bscale = ((f(-RHO) * f(cfg.dx)) * f(cfg.dx)) / f(DT)
    b = np.where(mat == 0, (bscale * div).astype(f), f(0)).astype(np.float64)
    for tol in (1e-4, 1e-6, 1e-7):
        x, it, rst, rel = oracle.pcg_true(div, mat, cfg.dx, DT, RHO, tol, 20000)
        rel2, _ = oracle.residual_f64(x, b, mat)
        print(f"tol {tol:g}: {it} iterations, {rst} restarts, relres {rel:.3e} (oracle {rel2:.3e})")
        assert rel <= tol and rel2 <= 1.01 * tol
        assert abs(rel2 - rel) <= 1e-3 * rel2 + 1e-12
    x64, _ = oracle.pcg_f64(b, mat, tol=1e-12)
    fl = mat == 0
    assert np.linalg.norm((x - x64)[fl]) / np.linalg.norm(x64[fl]) < 1e-3
