"""The oracle's restatement of the reference's pot + carrot voxelisation
(oracle/pot_geometry.py; python/boilingsim/geometry.py:51-214, 317-339) for
the reference's 200-step gate case (single_carrot.yaml at dx = 4 mm,
python/tests/test_pressure_parity.py:159-224), pinned by: the grid size the
reference's own report quotes for dx = 2 mm; at 4 mm the same float32 mask
with and without FMA contraction of the cell centres, and a float64
evaluation that differs only on the one layer whose centres lie exactly on
the carrot's base (shown in exact arithmetic); a brute-force per-cell loop;
and closed forms for the carrot and the water column."""

from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import pot_geometry as pg


def test_grid_dims():
    # The dx = 2 mm pot is 110x110x68 (benchmarks/phase5_phase6_implementation_report.md
    # quotes this grid for the published timings); 4 mm gives 58x58x38.
    assert pg.compute_grid_dims(0.002, 0.20, 0.12)[:3] == (110, 110, 68)
    nx, ny, nz, origin = pg.compute_grid_dims(0.004, 0.20, 0.12)
    assert (nx, ny, nz) == (58, 58, 38)
    assert origin == pytest.approx((-0.116, -0.116, -0.016))


def test_float32_mask_and_its_knife_edge():
    """At dx = 4 mm Warp's float32 mask does not depend on whether the
    compiler contracts origin + (i + 0.5) dx into an FMA.  A float64
    evaluation differs on exactly one layer: the carrot's base plane
    z = 0.03 m, on which the k = 11 cell centres lie exactly (exact decimal
    arithmetic), so the strict test z - 0.03 > 0 is decided by rounding;
    float32 rounds those centres up (inside the carrot) either way."""
    from fractions import Fraction as Fr

    m32, origin = pg.build_pot_mask(0.004)
    assert np.array_equal(m32, pg.build_pot_mask(0.004, fused=True)[0])
    m64, _ = pg.build_pot_mask(0.004, dtype=np.float64)
    diff = np.argwhere(m32 != m64)
    assert len(diff) == 32 and set(diff[:, 2]) == {11}
    assert np.all(m32[m32 != m64] == pg.MAT_CARROT) and np.all(m64[m32 != m64] == pg.MAT_FLUID)
    assert Fr("-0.016") + (11 + Fr(1, 2)) * Fr("0.004") == Fr("0.03")


def test_two_mm_mask_hinges_on_contraction():
    """Recorded for the reference's dx = 2 mm grid (not the gate case): the
    cavity floor z = 0.005 m lies exactly on the k = 6 centre layer, and
    whether that layer is fluid or air depends on the compiler's rounding."""
    m32, _ = pg.build_pot_mask(0.002)
    mf, _ = pg.build_pot_mask(0.002, fused=True)
    diff = np.argwhere(m32 != mf)
    assert len(diff) > 0 and set(diff[:, 2]) == {6}


def test_mask_brute_force():
    """A per-cell Python loop with the geometry's plain inequalities in
    float64 equals the vectorised float64 evaluation."""
    dx = 0.004
    mat, origin = pg.build_pot_mask(dx, dtype=np.float64)
    nx, ny, nz = mat.shape
    s = pg.SINGLE_CARROT
    ro, ri = 0.10, 0.10 - 0.003
    base, h = 0.005, 0.12
    wl = base + 0.75 * (h - base)
    cz = s["carrot_position"][2]
    want = np.empty_like(mat)
    for i in range(nx):
        x = origin[0] + (i + 0.5) * dx
        for j in range(ny):
            y = origin[1] + (j + 0.5) * dx
            r = math.sqrt(x * x + y * y)
            for k in range(nz):
                z = origin[2] + (k + 0.5) * dx
                in_outer = r < ro and 0.0 < z < h
                in_cavity = r < ri and base < z < base + (h - base)
                if in_outer and not in_cavity:
                    m = pg.MAT_POT_WALL
                elif r < ri and base < z < wl:
                    m = pg.MAT_CARROT if (r < 0.0125 and 0.0 < z - cz < 0.05) else pg.MAT_FLUID
                else:
                    m = pg.MAT_AIR
                want[i, j, k] = m
    assert np.array_equal(mat, want)


def test_closed_forms():
    mat, origin = pg.build_pot_mask(0.004)
    # Carrot: 13 layers (z = 0.030 .. 0.078 m) of the cells whose centres,
    # odd multiples of 2 mm from the axis, lie within 12.5 mm: 8 per quadrant.
    assert (mat == pg.MAT_CARROT).sum() == 13 * 32
    assert set(np.unique(np.nonzero(mat == pg.MAT_CARROT)[2])) == set(range(11, 24))
    # Water column: z in (0.005, 0.09125) is k = 5..26; every fluid column is
    # capped by air (the free surface), so A is SPD.
    fl = mat == pg.MAT_FLUID
    ks = np.nonzero(fl)[2]
    assert ks.min() == 5 and ks.max() == 26
    assert np.all(mat[:, :, 27][fl[:, :, 26]] == pg.MAT_AIR)
    # The grid edge is air (4-cell pad).
    for face in (mat[0], mat[-1], mat[:, 0], mat[:, -1], mat[:, :, 0], mat[:, :, -1]):
        assert np.all(face == pg.MAT_AIR)


def test_gate_velocities_draw_order():
    ux, uy, uz = pg.gate_velocities((58, 58, 38))
    rng = np.random.default_rng(7)
    a = rng.standard_normal((59, 58, 38)).astype(np.float32) * 0.05
    assert ux.dtype == np.float32 and np.array_equal(ux, a)
    assert uy.shape == (58, 59, 38) and uz.shape == (58, 58, 39)
