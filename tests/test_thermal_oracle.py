"""CPU pins of the backward-Euler conduction oracle (oracle/thermal_ref.c,
restating python/boilingsim/thermal.py:259-388 and 631-665): closed forms,
an independently assembled FP64 backward-Euler system solved by scipy, energy
conservation and the discrete decay of a cosine mode — the implicit-scheme
counterparts of test_thermal.py:135-296."""

from __future__ import annotations

import math

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle
from boilsim.geometry import MAT_AIR, MAT_CARROT, MAT_FLUID, MAT_POT_WALL
from boilsim.thermal import MaterialProps

PROPS = MaterialProps.from_pot("steel_304")
K, RHO, CP = PROPS.k, PROPS.rho, PROPS.c_p
H_CONV, T_AMB = 10.0, 295.15


def solve(t0, mat, dx, dt, n_iter, h=H_CONV, t_amb=T_AMB):
    t = np.ascontiguousarray(t0, dtype=np.float32).copy()
    t_alt = np.zeros_like(t)
    t_old = np.zeros_like(t)
    oracle.conduction_solve(t, t_alt, t_old, mat, K, RHO, CP, dx, dt, h, t_amb, n_iter)
    return t


def be_system(t_old, mat, dx, dt, h=H_CONV, t_amb=T_AMB):
    """FP64 backward-Euler system over the non-air cells, assembled from the
    face rules: harmonic k at material faces, h*dx at air faces, none at the
    domain edge."""
    nx, ny, nz = mat.shape
    idx = -np.ones(mat.shape, np.int64)
    cells = np.argwhere(mat != MAT_AIR)
    idx[tuple(cells.T)] = np.arange(len(cells))
    rows, cols, vals = [], [], []
    b = np.zeros(len(cells))
    for q, (i, j, k) in enumerate(cells):
        m = mat[i, j, k]
        rc = float(RHO[m]) * float(CP[m])
        diag = 1.0
        b[q] = float(t_old[i, j, k])
        for di, dj, dk in ((-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1)):
            a, c, d = i + di, j + dj, k + dk
            if not (0 <= a < nx and 0 <= c < ny and 0 <= d < nz):
                continue
            mn = mat[a, c, d]
            if mn == MAT_AIR:
                g = dt * h / (rc * dx)
                b[q] += g * t_amb
            else:
                kf = 2.0 * float(K[m]) * float(K[mn]) / (float(K[m]) + float(K[mn]))
                g = dt * kf / (rc * dx * dx)
                rows.append(q)
                cols.append(idx[a, c, d])
                vals.append(-g)
            diag += g
        rows.append(q)
        cols.append(q)
        vals.append(diag)
    A = sp.csr_matrix((vals, (rows, cols)), shape=(len(cells), len(cells)))
    return A, b, cells


def test_newton_cooling_single_cell():
    """A steel cell between two air cells: one sweep is the exact solve
    T = (T_old + 2 g T_amb)/(1 + 2 g), g = dt h / (rho c_p dx)."""
    mat = np.array([MAT_AIR, MAT_POT_WALL, MAT_AIR], np.int32).reshape(3, 1, 1)
    t0 = np.array([200.0, 400.0, 250.0], np.float32).reshape(3, 1, 1)
    dx, dt = 1e-3, 5.0
    t = solve(t0, mat, dx, dt, 1)
    g = dt * H_CONV / (float(RHO[1]) * float(CP[1]) * dx)
    expect = (400.0 + 2 * g * T_AMB) / (1 + 2 * g)
    assert abs(t[1, 0, 0] - expect) < 2e-5 * expect
    # air cells are left alone: the odd count copies the other buffer, whose
    # air cells were never written (zeros here)
    assert t[0, 0, 0] == 0.0 and t[2, 0, 0] == 0.0


def test_harmonic_face_two_cells():
    """Water and steel cells: the converged sweeps solve the 2x2 BE system
    with the harmonic face conductivity (an arithmetic mean fails this)."""
    mat = np.array([MAT_FLUID, MAT_POT_WALL], np.int32).reshape(2, 1, 1)
    t0 = np.array([300.0, 400.0], np.float32).reshape(2, 1, 1)
    dx, dt = 1e-3, 0.5
    t = solve(t0, mat, dx, dt, 400)
    kw, ks = float(K[0]), float(K[1])
    kf = 2 * kw * ks / (kw + ks)
    gw = dt * kf / (float(RHO[0]) * float(CP[0]) * dx * dx)
    gs = dt * kf / (float(RHO[1]) * float(CP[1]) * dx * dx)
    M = np.array([[1 + gw, -gw], [-gs, 1 + gs]])
    x = np.linalg.solve(M, [300.0, 400.0])
    assert np.allclose(t.ravel(), x, rtol=2e-6), (t.ravel(), x)
    ka = 0.5 * (kw + ks)
    M2 = np.array([[1 + gw * ka / kf, -gw * ka / kf], [-gs * ka / kf, 1 + gs * ka / kf]])
    assert not np.allclose(t.ravel(), np.linalg.solve(M2, [300.0, 400.0]), rtol=1e-3)


def test_uniform_field_is_fixed():
    """T = T_amb everywhere is a fixed point of the sweep (to float rounding)."""
    rng = np.random.default_rng(3)
    mat = rng.choice([0, 1, 2, 3], size=(6, 5, 7), p=[0.5, 0.2, 0.2, 0.1]).astype(np.int32)
    t0 = np.full(mat.shape, T_AMB, np.float32)
    t = solve(t0, mat, 2e-3, 1.0, 3)
    live = mat != MAT_AIR
    assert np.abs(t[live] - T_AMB).max() <= 2e-4


def test_converged_sweeps_solve_the_be_system():
    """Random materials (air included): many sweeps converge to the FP64 BE
    solution assembled independently and solved by scipy."""
    rng = np.random.default_rng(11)
    mat = rng.choice([0, 1, 2, 3], size=(8, 7, 6), p=[0.5, 0.2, 0.2, 0.1]).astype(np.int32)
    t0 = (290.0 + 80.0 * rng.random(mat.shape)).astype(np.float32)
    dx, dt = 2e-3, 2.0
    A, b, cells = be_system(t0, mat, dx, dt)
    x = spla.spsolve(A.tocsc(), b)
    t = solve(t0, mat, dx, dt, 601)
    got = t[tuple(cells.T)]
    assert np.abs(got - x).max() < 2e-5 * np.abs(x).max()
    # the same check fails for a sweep count far from convergence
    t5 = solve(t0, mat, dx, dt, 5)
    assert np.abs(t5[tuple(cells.T)] - x).max() > 1e-3


def test_energy_conserved_in_closed_box():
    """No air: the converged BE step conserves sum(rho c_p T) exactly (the
    face coefficients are symmetric) across water, steel and carrot."""
    rng = np.random.default_rng(5)
    mat = rng.choice([MAT_FLUID, MAT_POT_WALL, MAT_CARROT], size=(10, 9, 8)).astype(np.int32)
    t0 = (300.0 + 50.0 * rng.random(mat.shape)).astype(np.float32)
    rc = (RHO.astype(np.float64) * CP.astype(np.float64))[mat]
    e0 = float((rc * t0).sum())
    t = solve(t0, mat, 1e-3, 0.5, 801)
    e1 = float((rc * t.astype(np.float64)).sum())
    assert abs(e1 - e0) / e0 < 1e-6


def test_cosine_mode_decay_factor():
    """Steel bar, adiabatic ends: the first cosine mode is an eigenvector of
    the discrete operator, so one converged BE step multiplies its amplitude
    by 1 / (1 + 2 r (1 - cos(pi/N))), r = alpha dt / dx^2, and keeps the mean."""
    n, dx, dt = 40, 5e-4, 0.5
    x = (np.arange(n) + 0.5) / n
    t0 = (293.15 + 50.0 * np.cos(math.pi * x)).astype(np.float32).reshape(n, 1, 1)
    mat = np.full((n, 1, 1), MAT_POT_WALL, np.int32)
    t = solve(t0, mat, dx, dt, 2001).ravel().astype(np.float64)
    alpha = float(K[1]) / (float(RHO[1]) * float(CP[1]))
    r = alpha * dt / dx ** 2
    amp = 50.0 / (1 + 2 * r * (1 - math.cos(math.pi / n)))
    fit = 2 * np.dot(t - t.mean(), np.cos(math.pi * x)) / n
    assert abs(fit - amp) < 1e-4 * 50.0
    assert abs(t.mean() - float(t0.astype(np.float64).mean())) < 1e-4
