"""ORACLE (test infrastructure only): the reference's pot + carrot
voxelisation, restated with numpy for the reference's own 200-step gate case.

Restates python/boilingsim/geometry.py:
  - compute_grid_dims (geometry.py:317-339): grid bounding the pot with a
    4-cell pad, nx and ny rounded up to even, origin (-nx dx/2, -ny dx/2, -pad);
  - build_pot_sdf / _sdf_cylinder (geometry.py:51-83): the pot wall is where
    max(d_outer, -d_inner) < 0, d_inner the cavity cylinder lifted by the
    base thickness;
  - init_water_volume_fraction (geometry.py:86-110): liquid where the cell
    centre is inside the cavity radius, above the base and below the water
    line base + fill * (height - base);
  - populate_base_materials (geometry.py:118-148): wall, else fluid where
    alpha > 0.5, else air;
  - overlay_ingredient_voxels, axis z (geometry.py:151-214), with
    auto_place_carrots' count-1 placement (config.py:973-1003, the anchor
    itself): fluid cells with r_perp < radius and 0 < z - base_z < length
    become carrot;
  - the scenario numbers of configs/scenarios/single_carrot.yaml:14-35.

Each kernel's arithmetic is float32 in the kernel's written order (cell
centre origin + (i + 0.5) dx, squared radii, square roots), as Warp runs it.
The classification is a set of strict inequalities; tests check that the
same mask comes out of a float64 evaluation and of an FMA-contracted
float32 one, so it does not hinge on the compiler's choices."""

from __future__ import annotations

import math

import numpy as np

MAT_FLUID, MAT_POT_WALL, MAT_AIR, MAT_CARROT = 0, 1, 2, 3

# configs/scenarios/single_carrot.yaml:14-35
SINGLE_CARROT = {
    "pot_diameter_m": 0.20,
    "pot_height_m": 0.12,
    "wall_thickness_m": 0.003,
    "base_thickness_m": 0.005,
    "fill_fraction": 0.75,
    "carrot_diameter_m": 0.025,
    "carrot_length_m": 0.05,
    "carrot_position": (0.0, 0.0, 0.03),
}


def compute_grid_dims(dx: float, pot_diameter_m: float, pot_height_m: float, pad_cells: int = 4):
    """geometry.py:317-339 (Python float64, as the reference)."""
    r_outer = pot_diameter_m / 2
    pad = pad_cells * dx
    nx = int(math.ceil((2 * r_outer + 2 * pad) / dx))
    ny = int(math.ceil((2 * r_outer + 2 * pad) / dx))
    nz = int(math.ceil((pot_height_m + 2 * pad) / dx))
    nx += nx % 2
    ny += ny % 2
    return nx, ny, nz, (-nx * dx / 2, -ny * dx / 2, -pad)


def _centres(n, o, dx, dtype, fused):
    """Cell-centre coordinates o + (float(i) + 0.5) * dx along one axis."""
    t = np.arange(n, dtype=dtype) + dtype(0.5)
    if fused:  # one rounding of the multiply-add (what a contracting compiler may emit)
        return (t.astype(np.float64) * float(dtype(dx)) + float(dtype(o))).astype(dtype)
    return dtype(o) + t * dtype(dx)


def _sdf_cylinder(x, y, z, r, h, dtype):
    """geometry.py:51-62 in `dtype`: distance to a capped cylinder, z in [0, h]."""
    c = dtype
    d_xy = np.sqrt(x * x + y * y) - c(r)
    d_z = np.abs(z - c(h) * c(0.5)) - c(h) * c(0.5)
    a = np.maximum(d_xy, c(0.0))
    b = np.maximum(d_z, c(0.0))
    out = np.sqrt(a * a + b * b)
    ins = np.minimum(np.maximum(d_xy, d_z), c(0.0))
    return out + ins


def build_pot_mask(dx: float, scenario: dict = SINGLE_CARROT, dtype=np.float32, fused: bool = False):
    """int32 class ids (nx, ny, nz) of build_pot_geometry for a one-carrot
    scenario, and the grid origin.  dtype/fused select the arithmetic (the
    default is Warp's float32 in written order)."""
    s = scenario
    nx, ny, nz, origin = compute_grid_dims(dx, s["pot_diameter_m"], s["pot_height_m"])
    r_outer = s["pot_diameter_m"] / 2
    r_inner = r_outer - s["wall_thickness_m"]
    h_outer = s["pot_height_m"]
    h_inner = h_outer - s["base_thickness_m"]
    base = s["base_thickness_m"]
    water_line_z = base + s["fill_fraction"] * h_inner
    c = dtype
    x = _centres(nx, origin[0], dx, c, fused)[:, None, None]
    y = _centres(ny, origin[1], dx, c, fused)[None, :, None]
    z = _centres(nz, origin[2], dx, c, fused)[None, None, :]
    # build_pot_sdf: max(d_outer, -d_inner), inner cylinder offset by the base
    d_outer = _sdf_cylinder(x, y, z, r_outer, h_outer, c)
    d_inner = _sdf_cylinder(x, y, z - c(base), r_inner, h_inner, c)
    wall = np.maximum(d_outer, -d_inner) < c(0.0)
    # init_water_volume_fraction
    r_xy = np.sqrt(x * x + y * y)
    liquid = (r_xy < c(r_inner)) & (z > c(base)) & (z < c(water_line_z))
    mat = np.where(wall, MAT_POT_WALL, np.where(liquid, MAT_FLUID, MAT_AIR)).astype(np.int32)
    # overlay_ingredient_voxels, axis z, one instance at the anchor (its base)
    cx, cy, cz = (c(v) for v in s["carrot_position"])
    rx, ry, rz = x - cx, y - cy, z - cz
    r_perp = np.sqrt(rx * rx + ry * ry)
    carrot = (r_perp < c(s["carrot_diameter_m"] / 2.0)) & (rz > c(0.0)) & (rz < c(s["carrot_length_m"]))
    mat = np.where((mat == MAT_FLUID) & carrot, MAT_CARROT, mat).astype(np.int32)
    return np.ascontiguousarray(mat), origin


def gate_velocities(shape, seed: int = 7, sigma: float = 0.05):
    """The gate's face velocities (python/tests/test_pressure_parity.py:184-187):
    one default_rng(seed), ux, uy, uz drawn in that order as
    standard_normal(shape).astype(float32) * sigma."""
    nx, ny, nz = shape
    rng = np.random.default_rng(seed)
    ux = rng.standard_normal((nx + 1, ny, nz)).astype(np.float32) * np.float32(sigma)
    uy = rng.standard_normal((nx, ny + 1, nz)).astype(np.float32) * np.float32(sigma)
    uz = rng.standard_normal((nx, ny, nz + 1)).astype(np.float32) * np.float32(sigma)
    return ux, uy, uz


# The gate case of test_pressure_parity.py:159-224: single_carrot.yaml at
# dx = 4 mm, 200 Jacobi iterations, dt = 0.01 s, rho = 997 kg/m^3.
GATE_DX, GATE_ITERS, GATE_DT, GATE_RHO = 0.004, 200, 0.01, 997.0
