"""CPU pins of the no-slip and max-velocity oracle (oracle/cpu_ref.c
orc_enforce_no_slip / orc_max_velocity, restating python/boilingsim/
fluid.py:111-162 and 796-833), and of the property the fused projection step
rests on: after enforce_no_slip and a projection, the second enforce_no_slip
of pipeline.py:618-620 changes no bit."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from oracle import np_ref

MAT_FLUID, MAT_POT_WALL, MAT_AIR = 0, 1, 2


def walled(nx, ny, nz):
    """_make_walled_grid (test_fluid.py:53-72): fluid inside a one-cell wall shell."""
    mat = np.full((nx, ny, nz), MAT_POT_WALL, np.int32)
    mat[1:-1, 1:-1, 1:-1] = MAT_FLUID
    return mat


def faces(rng, shape, scale=1.0):
    nx, ny, nz = shape
    return tuple((rng.standard_normal(s) * scale).astype(np.float32)
                 for s in ((nx + 1, ny, nz), (nx, ny + 1, nz), (nx, ny, nz + 1)))


def random_mat(rng, shape, p=(0.55, 0.2, 0.2, 0.05)):
    """Random classes with fluid on the domain edges too (no air ring)."""
    return rng.choice([0, 1, 2, 3], size=shape, p=list(p)).astype(np.int32)


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def brute_no_slip(ux, uy, uz, mat, mf=MAT_FLUID):
    """Per-face loops with the reference kernels' branches (fluid.py:111-150)."""
    nx, ny, nz = mat.shape
    ux, uy, uz = ux.copy(), uy.copy(), uz.copy()
    for i in range(nx + 1):
        for j in range(ny):
            for k in range(nz):
                if i == 0 or i == nx or mat[i - 1, j, k] != mf or mat[i, j, k] != mf:
                    ux[i, j, k] = 0.0
    for i in range(nx):
        for j in range(ny + 1):
            for k in range(nz):
                if j == 0 or j == ny or mat[i, j - 1, k] != mf or mat[i, j, k] != mf:
                    uy[i, j, k] = 0.0
    for i in range(nx):
        for j in range(ny):
            for k in range(nz + 1):
                if k == 0 or k == nz or mat[i, j, k - 1] != mf or mat[i, j, k] != mf:
                    uz[i, j, k] = 0.0
    return ux, uy, uz


@pytest.mark.parametrize("seed,shape", [(0, (4, 5, 6)), (1, (7, 3, 5)), (2, (1, 1, 1)), (3, (2, 6, 3))])
def test_no_slip_matches_brute_force(seed, shape):
    rng = np.random.default_rng(seed)
    mat = random_mat(rng, shape)
    ux, uy, uz = faces(rng, shape)
    want = brute_no_slip(ux, uy, uz, mat)
    oracle.enforce_no_slip(ux, uy, uz, mat)
    for g, w in zip((ux, uy, uz), want):
        assert np.array_equal(bits(g), bits(w))


@pytest.mark.parametrize("seed,shape", [(5, (16, 16, 16)), (6, (33, 17, 44)), (7, (9, 40, 2))])
def test_no_slip_matches_numpy_restatement(seed, shape):
    rng = np.random.default_rng(seed)
    mat = random_mat(rng, shape)
    ux, uy, uz = faces(rng, shape)
    # negative zeros, NaN and inf among the inputs
    ux.ravel()[::7] = -0.0
    uy.ravel()[::11] = np.nan
    uz.ravel()[::13] = -np.inf
    want = np_ref.enforce_no_slip(ux, uy, uz, mat)
    oracle.enforce_no_slip(ux, uy, uz, mat)
    for g, w in zip((ux, uy, uz), want):
        assert np.array_equal(bits(g), bits(w))


def test_reference_walled_grid_check():
    """test_fluid.py:94-117: random velocities on a walled 8^3 grid; every x
    face touching a wall cell is zero afterwards.  The y and z faces are
    checked the same way, and the fluid-fluid faces keep their values."""
    nx = ny = nz = 8
    mat = walled(nx, ny, nz)
    ux = np.random.default_rng(0).standard_normal((nx + 1, ny, nz)).astype(np.float32)
    uy = np.random.default_rng(1).standard_normal((nx, ny + 1, nz)).astype(np.float32)
    uz = np.random.default_rng(2).standard_normal((nx, ny, nz + 1)).astype(np.float32)
    u0 = (ux.copy(), uy.copy(), uz.copy())
    oracle.enforce_no_slip(ux, uy, uz, mat)
    for i in range(1, nx):
        for j in range(ny):
            for k in range(nz):
                both = mat[i - 1, j, k] == MAT_FLUID and mat[i, j, k] == MAT_FLUID
                assert ux[i, j, k] == (u0[0][i, j, k] if both else 0.0)
    assert (ux[0] == 0).all() and (ux[nx] == 0).all()
    inner = mat[:, 1:, :] == MAT_FLUID
    inner &= mat[:, :-1, :] == MAT_FLUID
    assert np.array_equal(uy[:, 1:-1, :][inner], u0[1][:, 1:-1, :][inner])
    assert (uy[:, 1:-1, :][~inner] == 0).all()
    innz = (mat[:, :, 1:] == MAT_FLUID) & (mat[:, :, :-1] == MAT_FLUID)
    assert np.array_equal(uz[:, :, 1:-1][innz], u0[2][:, :, 1:-1][innz])
    assert (uz[:, :, 1:-1][~innz] == 0).all()


def test_no_slip_idempotent_and_positive_zero():
    rng = np.random.default_rng(9)
    shape = (10, 9, 8)
    mat = random_mat(rng, shape)
    ux, uy, uz = faces(rng, shape)
    for a in (ux, uy, uz):
        a.ravel()[::3] = -0.0
    oracle.enforce_no_slip(ux, uy, uz, mat)
    once = [a.copy() for a in (ux, uy, uz)]
    oracle.enforce_no_slip(ux, uy, uz, mat)
    for a, b in zip((ux, uy, uz), once):
        assert np.array_equal(bits(a), bits(b))
    # the zeroed faces hold +0.0 (the store of 0.0), kept -0.0 faces stay -0.0
    keep = np_ref.enforce_no_slip(np.ones_like(ux), np.ones_like(uy), np.ones_like(uz), mat)
    for a, k in zip((ux, uy, uz), keep):
        assert (bits(a)[k == 0] == 0).all()
        kept_neg_zero = (k == 1) & (bits(a) == 0x80000000)
        assert kept_neg_zero.sum() > 0


def test_max_velocity():
    rng = np.random.default_rng(4)
    ux, uy, uz = faces(rng, (12, 7, 9), scale=0.3)
    uy[3, 2, 1] = -5.5
    assert oracle.max_velocity(ux, uy, uz) == 5.5
    m = max(np.abs(ux).max(), np.abs(uy).max(), np.abs(uz).max())
    assert oracle.max_velocity(ux, uy, uz) == float(m)
    z = [np.zeros_like(a) for a in (ux, uy, uz)]
    assert oracle.max_velocity(*z) == 0.0


@pytest.mark.parametrize("geometry", ["walled", "random", "pot"])
def test_post_no_slip_changes_nothing(geometry):
    """After enforce_no_slip and a Jacobi projection, the faces the second
    enforce_no_slip would zero are already +0.0: the gradient subtraction
    writes only interior faces between two fluid cells."""
    rng = np.random.default_rng(17)
    if geometry == "walled":
        mat = walled(14, 12, 10)
    elif geometry == "random":
        mat = random_mat(rng, (13, 11, 16))
    else:
        from boilsim.fixtures import pot_mask

        mat = pot_mask(24, 24, 20)
    ux, uy, uz = faces(rng, mat.shape, 0.05)
    oracle.enforce_no_slip(ux, uy, uz, mat)
    oracle.projection(ux, uy, uz, mat, 2e-3, 0.01, 997.0, 30)
    before = [a.copy() for a in (ux, uy, uz)]
    oracle.enforce_no_slip(ux, uy, uz, mat)
    for a, b in zip((ux, uy, uz), before):
        assert np.array_equal(bits(a), bits(b))
    # without the first no-slip the same second pass does change faces
    ux, uy, uz = faces(rng, mat.shape, 0.05)
    oracle.projection(ux, uy, uz, mat, 2e-3, 0.01, 997.0, 30)
    before = [a.copy() for a in (ux, uy, uz)]
    oracle.enforce_no_slip(ux, uy, uz, mat)
    assert any(not np.array_equal(bits(a), bits(b)) for a, b in zip((ux, uy, uz), before))


def test_hydrostatic_closed_box():
    """test_fluid.py:199-217: a closed box at rest stays at rest through
    no-slip, projection, no-slip."""
    mat = walled(12, 12, 12)
    z = [np.zeros(s, np.float32) for s in ((13, 12, 12), (12, 13, 12), (12, 12, 13))]
    _, _, _, max_div, max_u = oracle.projection_step(*z, mat, 2e-3, 0.01, 997.0, 200)
    assert max_div < 1e-6 and max_u < 1e-6


def test_divergence_free_after_projection():
    """test_fluid.py:120-146: random 0.1 N(0,1) velocities on a walled 16^3
    grid, no-slip, 1500 Jacobi sweeps: max|div| below 1e-3 and 1e-3 of its
    value before."""
    mat = walled(16, 16, 16)
    rng = np.random.default_rng(42)
    ux, uy, uz = faces(rng, mat.shape, 0.1)
    oracle.enforce_no_slip(ux, uy, uz, mat)
    before = float(np.abs(oracle.divergence(ux, uy, uz, mat, 0.002)).max())
    _, _, _, after = oracle.projection(ux, uy, uz, mat, 0.002, 0.01, 997.0, 1500)
    assert before > 1.0
    assert after < 1e-3 and after < before * 1e-3
