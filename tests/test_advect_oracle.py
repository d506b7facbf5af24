"""CPU pins of the semi-Lagrangian advection oracle (oracle/advect_ref.c,
restating python/boilingsim/fluid.py:33-103, 338-492 and 737-775): the
independent numpy restatement, brute-force loops, closed forms (identity,
whole-cell shifts, linear fields) and the reference's temperature-drift gate
(test_fluid.py:147-184)."""

from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from oracle import np_ref

MAT_FLUID, MAT_POT_WALL, MAT_AIR = 0, 1, 2


def bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def case(seed, shape, speed=0.3, p=(0.6, 0.15, 0.2, 0.05)):
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    mat = rng.choice([0, 1, 2, 3], size=shape, p=list(p)).astype(np.int32)
    f = [(rng.standard_normal(s) * speed).astype(np.float32)
         for s in ((nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1))]
    t = (290.0 + 100.0 * rng.random(shape)).astype(np.float32)
    return mat, f, t


@pytest.mark.parametrize("seed,shape,speed", [(0, (6, 7, 5), 0.3), (1, (9, 4, 11), 1.0), (2, (1, 1, 1), 0.5),
                                              (3, (70, 3, 4), 0.5), (4, (5, 66, 9), 3.0), (5, (12, 10, 80), 0.1)])
def test_c_matches_numpy(seed, shape, speed):
    """Bit for bit, with displacements up to several cells (the clamps), a
    non-zero origin, an arbitrary dx, and axes of 65+ cells where the upper
    clamp reaches n-1."""
    mat, (ux, uy, uz), t = case(seed, shape, speed)
    origin, dx, dt = (-0.0123, -0.011, -0.004), 0.0017, 0.01
    got = oracle.advect(ux, uy, uz, t, mat, origin, dx, dt)
    want = np_ref.advect(ux, uy, uz, t, mat, origin, dx, dt)
    for g, w in zip(got[:4], want):
        assert np.array_equal(bits(g), bits(w))


def test_extend_brute_force():
    mat, _, t = case(7, (6, 5, 7), p=(0.3, 0.3, 0.3, 0.1))
    nx, ny, nz = mat.shape
    want = t.copy()
    for i in range(nx):
        for j in range(ny):
            for k in range(nz):
                if mat[i, j, k] == MAT_FLUID:
                    continue
                for (a, b, c) in ((i - 1, j, k), (i + 1, j, k), (i, j - 1, k), (i, j + 1, k), (i, j, k - 1),
                                  (i, j, k + 1)):
                    if 0 <= a < nx and 0 <= b < ny and 0 <= c < nz and mat[a, b, c] == MAT_FLUID:
                        want[i, j, k] = t[a, b, c]
                        break
    assert np.array_equal(oracle.extend_temperature(t, mat), want)
    assert np.array_equal(np_ref.extend_temperature(t, mat), want)


def test_zero_velocity_is_identity():
    """dx a power of two and the origin a multiple of it: every sample point
    is an exact lattice point, so zero velocity leaves u and T unchanged
    (edge faces aside, which advect_u* sets to 0).  The exception is the
    reference's upper clamp float(n-1) - 1e-6: on these short axes it stays
    below n-1, so a fluid cell on a high edge samples 1e-6 of its lower
    neighbour."""
    mat, (ux, uy, uz), t = case(8, (10, 9, 12))
    z = [np.zeros_like(a) for a in (ux, uy, uz)]
    dx = 2.0 ** -9
    origin = (-5 * dx, -4.5 * dx, -3 * dx)
    # velocities used for sampling are zero; the advected fields are t only
    _, _, _, t_new, _ = oracle.advect(*z, t, mat, origin, dx, 0.01)
    fl = mat == MAT_FLUID
    inner = np.zeros_like(fl)
    inner[:-1, :-1, :-1] = True
    assert np.array_equal(t_new[fl & inner], t[fl & inner])
    assert np.array_equal(t_new[~fl], t[~fl])
    edge = fl & ~inner
    t_ext = oracle.extend_temperature(t, mat)
    i, j, k = np.argwhere(edge & (np.indices(fl.shape)[0] == 9) & (np.indices(fl.shape)[1] < 8)
                          & (np.indices(fl.shape)[2] < 11))[0]
    w = np.float32(np.float32(9.0) - np.float32(1e-6)) - np.float32(8.0)  # tx at the clamp
    want = t_ext[8, j, k] * (np.float32(1.0) - w) + t_ext[9, j, k] * w
    assert t_new[i, j, k] == want
    assert (t_new[edge] != t[edge]).any()
    # a velocity field advected by itself at dt = 0 is also unchanged (off the
    # high edges of the two cell-centred axes, where the clamp acts as above)
    ux_n, uy_n, uz_n, _, _ = oracle.advect(ux, uy, uz, t, mat, origin, dx, 0.0)
    assert np.array_equal(ux_n[1:-1, :-1, :-1], ux[1:-1, :-1, :-1]) and (ux_n[[0, -1]] == 0).all()
    assert np.array_equal(uy_n[:-1, 1:-1, :-1], uy[:-1, 1:-1, :-1]) and (uy_n[:, [0, -1]] == 0).all()
    assert np.array_equal(uz_n[:-1, :-1, 1:-1], uz[:-1, :-1, 1:-1]) and (uz_n[:, :, [0, -1]] == 0).all()


@pytest.mark.parametrize("m", [1, 2, -3])
def test_whole_cell_shift(m):
    """All fluid, uniform ux = m dx/dt (powers of two): T moves by exactly m
    cells along x wherever the departure point is inside the grid."""
    nx, ny, nz = 16, 6, 5
    mat = np.zeros((nx, ny, nz), np.int32)
    dx, dt = 2.0 ** -8, 2.0 ** -6
    U = m * dx / dt
    ux = np.full((nx + 1, ny, nz), U, np.float32)
    uy = np.zeros((nx, ny + 1, nz), np.float32)
    uz = np.zeros((nx, ny, nz + 1), np.float32)
    t = np.random.default_rng(9).random((nx, ny, nz)).astype(np.float32)
    ux_n, _, _, t_new, _ = oracle.advect(ux, uy, uz, t, mat, (0.0, 0.0, 0.0), dx, dt)
    # departure cells i - m in [0, nx - 2] (nx - 1 is below the upper clamp);
    # the last j and k rows sit on the upper clamp of those short axes
    lo, hi = max(m, 0), min(nx, nx - 1 + m)
    assert np.array_equal(t_new[lo:hi, :-1, :-1], t[lo - m:hi - m, :-1, :-1])
    assert np.array_equal(ux_n[1:-1], ux[1:-1])  # a uniform field advects to itself


def test_linear_field_sub_cell():
    """Trilinear interpolation is exact on a linear field, so a uniform
    sub-cell velocity carries T = a + b.x to a + b.(x - u dt) up to float
    rounding, away from the domain edge."""
    nx, ny, nz = 14, 12, 10
    mat = np.zeros((nx, ny, nz), np.int32)
    dx, dt = 1.3e-3, 0.02
    u = np.array([0.023, -0.017, 0.031])
    ux = np.full((nx + 1, ny, nz), u[0], np.float32)
    uy = np.full((nx, ny + 1, nz), u[1], np.float32)
    uz = np.full((nx, ny, nz + 1), u[2], np.float32)
    I, J, K = np.indices((nx, ny, nz)).astype(np.float64)
    b = np.array([0.7, -1.1, 0.4])
    t = (300.0 + b[0] * I + b[1] * J + b[2] * K).astype(np.float32)
    _, _, _, t_new, _ = oracle.advect(ux, uy, uz, t, mat, (0.01, -0.02, 0.0), dx, dt)
    s = u * dt / dx  # displacement in cells (0.35, -0.26, 0.48)
    want = 300.0 + b[0] * (I - s[0]) + b[1] * (J - s[1]) + b[2] * (K - s[2])
    inner = (slice(2, -2),) * 3
    assert np.abs(t_new[inner] - want[inner]).max() < 1e-4
    # the departure point really is off the lattice: the answer is not a shift
    assert np.abs(t_new[inner] - t[inner]).max() > 0.1


def test_temperature_drift_gate():
    """test_fluid.py:147-184: a walled 12^3 grid with a cosine T bump, a
    random velocity made divergence-free (no-slip, 1500 Jacobi sweeps), then
    20 advection steps of u and T at dt = 0.01: the fluid's sum of T drifts
    by less than 2 %."""
    nx = ny = nz = 12
    mat = np.full((nx, ny, nz), MAT_POT_WALL, np.int32)
    mat[1:-1, 1:-1, 1:-1] = MAT_FLUID
    t = np.full((nx, ny, nz), 293.15, np.float32)
    for i in range(1, nx - 1):
        t[i] = 293.15 + 30.0 * math.cos(math.pi * (i - nx / 2) / nx)
    rng = np.random.default_rng(0)
    ux = (rng.standard_normal((nx + 1, ny, nz)) * 0.05).astype(np.float32)
    uy = (rng.standard_normal((nx, ny + 1, nz)) * 0.05).astype(np.float32)
    uz = (rng.standard_normal((nx, ny, nz + 1)) * 0.05).astype(np.float32)
    oracle.enforce_no_slip(ux, uy, uz, mat)
    oracle.projection(ux, uy, uz, mat, 0.002, 0.01, 997.0, 1500)
    fl = mat == MAT_FLUID
    before = float(t[fl].sum())
    for _ in range(20):
        ux, uy, uz, t, _ = oracle.advect(ux, uy, uz, t, mat, (0.0, 0.0, 0.0), 0.002, 0.01)
    drift = abs(float(t[fl].sum()) - before) / before
    assert drift < 0.02
