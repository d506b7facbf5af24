"""Deterministic synthetic inputs for the pressure projection.

The reference builds its pot by SDF voxelisation (python/boilingsim/geometry.py:
341-546) and seeds velocities with ``np.random.default_rng(seed)``
(python/tests/test_pressure_parity.py:184-187).  These generators restate that
set-up as closed-form, frozen definitions so every configuration can be rebuilt
bit-for-bit anywhere:

* :func:`pot_mask` — a cylindrical pot (2-cell steel wall, 2-cell base, air
  pad below and around), water up to a water line, air above (the free surface,
  Dirichlet p = 0) and a solid carrot cylinder on the axis.  Lengths are given
  on a 64x64x48 grid and scale with nx/64 (x, y) and nz/48 (z).
* :func:`seeded_velocities` — face velocities ux (nx+1,ny,nz), uy (nx,ny+1,nz),
  uz (nx,ny,nz+1) drawn in that order as float32 N(0,1)*sigma from one
  ``default_rng(seed)``, plus an optional Gaussian buoyant plume on uz.
* :func:`gaussian_random_field` — masked, unit-variance Gaussian random fields
  (white noise smoothed by a Gaussian of correlation length ell cells), the
  right-hand sides the learned preconditioner is trained and evaluated on.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .geometry import MAT_AIR, MAT_CARROT, MAT_FLUID, MAT_POT_WALL

# Physical constants of every configuration (dt in s, rho in kg/m^3).
DT = 1.0e-3
RHO = 958.0


@dataclass(frozen=True)
class GridConfig:
    name: str
    nx: int
    ny: int
    nz: int
    dx: float  # m

    @property
    def shape(self) -> tuple[int, int, int]:
        return (self.nx, self.ny, self.nz)

    @property
    def n(self) -> int:
        return self.nx * self.ny * self.nz


CONFIGS = {
    "small": GridConfig("small", 64, 64, 48, 2.0e-3),
    "dev": GridConfig("dev", 110, 110, 70, 2.0e-3),
    "prod": GridConfig("prod", 440, 440, 280, 0.5e-3),
}


def pot_mask(nx: int, ny: int, nz: int) -> np.ndarray:
    """int32 class ids (0 fluid, 1 pot wall, 2 air, 3 carrot), C-contiguous.

    Cell centres: X = i+0.5-nx/2, Y = j+0.5-ny/2, Z = k+0.5, R = sqrt(X^2+Y^2).
    On the 64x64x48 grid: pot outer radius 28, wall 2 (inner radius 26), base
    Z in [2,4), rim Z < 46, water line 40, carrot R < 6 for Z in [8,33).
    """
    sx = nx / 64.0
    sz = nz / 48.0
    X = (np.arange(nx, dtype=np.float64) + 0.5 - nx / 2.0)[:, None, None]
    Y = (np.arange(ny, dtype=np.float64) + 0.5 - ny / 2.0)[None, :, None]
    Z = (np.arange(nz, dtype=np.float64) + 0.5)[None, None, :]
    R = np.sqrt(X * X + Y * Y)
    r_out, r_in, r_car = 28.0 * sx, 26.0 * sx, 6.0 * sx
    z_base0, z_base1, z_rim, z_wl = 2.0 * sz, 4.0 * sz, 46.0 * sz, 40.0 * sz
    z_car0, z_car1 = 8.0 * sz, 33.0 * sz
    mat = np.full((nx, ny, nz), MAT_AIR, dtype=np.int32)
    wall = (R < r_out) & (Z >= z_base0) & (Z < z_rim) & ((R >= r_in) | (Z < z_base1))
    fluid = (R < r_in) & (Z >= z_base1) & (Z < z_wl)
    carrot = fluid & (R < r_car) & (Z >= z_car0) & (Z < z_car1)
    mat[np.broadcast_to(wall, mat.shape)] = MAT_POT_WALL
    mat[np.broadcast_to(fluid, mat.shape)] = MAT_FLUID
    mat[np.broadcast_to(carrot, mat.shape)] = MAT_CARROT
    return mat


def seeded_velocities(nx: int, ny: int, nz: int, seed: int, sigma: float = 0.01,
                      plume: bool = True) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
    """Face velocities (m/s): N(0, sigma) float32 drawn ux, uy, uz in that order,
    plus, if ``plume``, w += 0.05*exp(-R^2/(2*(8*nx/64)^2)) on z-faces at or
    above Z = 4*nz/48, R measured from the column centre of the face."""
    rng = np.random.default_rng(seed)
    ux = rng.standard_normal((nx + 1, ny, nz)).astype(np.float32) * np.float32(sigma)
    uy = rng.standard_normal((nx, ny + 1, nz)).astype(np.float32) * np.float32(sigma)
    uz = rng.standard_normal((nx, ny, nz + 1)).astype(np.float32) * np.float32(sigma)
    if plume:
        X = (np.arange(nx, dtype=np.float64) + 0.5 - nx / 2.0)[:, None]
        Y = (np.arange(ny, dtype=np.float64) + 0.5 - ny / 2.0)[None, :]
        w = 8.0 * nx / 64.0
        prof = (0.05 * np.exp(-(X * X + Y * Y) / (2.0 * w * w))).astype(np.float32)
        kz = np.arange(nz + 1, dtype=np.float64)
        above = kz >= 4.0 * nz / 48.0
        uz[:, :, above] += prof[:, :, None]
    return ux, uy, uz


def gaussian_random_field(mat: np.ndarray, seed: int, ell: float | None = None,
                          ell_range: tuple[float, float] = (4.0, 16.0),
                          mat_fluid: int = MAT_FLUID) -> tuple[np.ndarray, float]:
    """Unit-variance (over fluid cells) float32 GRF, zero off fluid.

    White noise from ``default_rng(seed)`` is smoothed with a Gaussian of
    standard deviation ``ell`` cells (FFT, periodic); ``ell`` defaults to a
    draw from U(ell_range) taken from the same generator before the noise.
    Returns (field, ell)."""
    rng = np.random.default_rng(seed)
    if ell is None:
        ell = float(rng.uniform(*ell_range))
    shape = mat.shape
    noise = rng.standard_normal(shape).astype(np.float32)
    f = np.fft.rfftn(noise)
    kx = np.fft.fftfreq(shape[0])[:, None, None]
    ky = np.fft.fftfreq(shape[1])[None, :, None]
    kz = np.fft.rfftfreq(shape[2])[None, None, :]
    k2 = (kx * kx + ky * ky + kz * kz) * (2.0 * np.pi) ** 2
    f *= np.exp(-0.5 * k2 * ell * ell)
    g = np.fft.irfftn(f, s=shape).astype(np.float64)
    fluid = mat == mat_fluid
    vals = g[fluid]
    vals = (vals - vals.mean()) / max(vals.std(), 1e-30)
    out = np.zeros(shape, dtype=np.float32)
    out[fluid] = vals.astype(np.float32)
    return out, ell


def ar1_sequence_step(prev: np.ndarray, innov: np.ndarray, phi: float = 0.95) -> np.ndarray:
    """f_{t+1} = phi*f_t + sqrt(1-phi^2)*xi_t (unit variance preserved)."""
    return (phi * prev + np.sqrt(1.0 - phi * phi) * innov).astype(np.float32)
