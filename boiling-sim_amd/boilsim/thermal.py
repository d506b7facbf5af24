"""Backward-Euler Jacobi heat conduction (the implicit branch of the
reference's conduct_one_step, python/boilingsim/thermal.py:631-665) over the
HIP library.

Material properties are indexed by material id [fluid, pot wall, air,
carrot] as in MaterialProps.from_scenario (thermal.py:82-125): water at its
25 C reference density, the pot's metal, dry air at 25 C, carrot tissue; the
values restate data/materials.json:2-63 (float32, as the reference's arrays).
The boundary-source kernels that follow the conduction in conduct_one_step
(base heat flux, evaporative cooling, thermal.py:667-720) are not part of
this path.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .geometry import MAT_AIR, MAT_FLUID

# (rho kg/m^3, c_p J/(kg K), k W/(m K)), data/materials.json
WATER = (997.0, 4186.0, 0.606)
POT_METALS = {
    "steel_304": (8000.0, 500.0, 16.2),
    "cast_iron": (7200.0, 460.0, 55.0),
    "aluminum": (2700.0, 900.0, 237.0),
    "copper": (8960.0, 385.0, 401.0),
}
AIR = (1.184, 1005.0, 0.0262)  # thermal.py:103-106
CARROT = (1040.0, 3900.0, 0.605)


@dataclass
class MaterialProps:
    rho: np.ndarray
    c_p: np.ndarray
    k: np.ndarray

    @classmethod
    def from_pot(cls, pot_material: str = "steel_304") -> "MaterialProps":
        rows = [WATER, POT_METALS[pot_material], AIR, CARROT]
        arr = np.array(rows, dtype=np.float32)
        return cls(rho=np.ascontiguousarray(arr[:, 0]), c_p=np.ascontiguousarray(arr[:, 1]),
                   k=np.ascontiguousarray(arr[:, 2]))

    def device(self, torch):
        """(k, rho, c_p) as float32 device tensors (the reference's k_wp,
        rho_wp, cp_wp)."""
        f = lambda a: torch.from_numpy(a).cuda()
        return f(self.k), f(self.rho), f(self.c_p)


@dataclass
class ThermalWorkspace:
    """The reference's implicit-conduction scratch (thermal.py:563-580): the
    ping-pong partner of T and the T_old snapshot."""

    T_iter_alt: object
    T_old_snapshot: object


def allocate_thermal_workspace(torch, shape) -> ThermalWorkspace:
    z = lambda: torch.zeros(shape, dtype=torch.float32, device="cuda")
    return ThermalWorkspace(T_iter_alt=z(), T_old_snapshot=z())


def conduct_implicit(core, T, mat, props_dev, ws: ThermalWorkspace, dx, dt, n_iter, h_conv=10.0,
                     ambient_temp_c=22.0, stream=0):
    """The backward-Euler branch of conduct_one_step (thermal.py:648-665) on
    device tensors: T_old = T, n_iter Jacobi sweeps ping-ponging between T and
    T_iter_alt, the result in T.  Defaults are the reference's
    diffusion_max_iter (15 in config.py:184), h_conv_outer_w_per_m2_k (10)
    and ambient_temp_c (22)."""
    k, rho, cp = props_dev
    nx, ny, nz = T.shape
    core.conduction_solve(T.data_ptr(), ws.T_iter_alt.data_ptr(), ws.T_old_snapshot.data_ptr(), mat.data_ptr(),
                          k.data_ptr(), rho.data_ptr(), cp.data_ptr(), k.numel(), nx, ny, nz, dx, dt, h_conv,
                          ambient_temp_c + 273.15, n_iter, MAT_AIR, MAT_FLUID, stream=stream)


__all__ = ["MaterialProps", "ThermalWorkspace", "allocate_thermal_workspace", "conduct_implicit"]
