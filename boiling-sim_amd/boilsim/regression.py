"""Multi-step regression runs over the rows built here (the developer guide's
50-step protocol, multiphysics_boiling_developer_guide.md §5.4: run N steps,
dump the state, compare per field against a reference run).

One step is the part of Simulation.step (python/boilingsim/pipeline.py:
521-640) that this library covers, in the reference's order:

    advect_all (pipeline.py:528-529)          semi-Lagrangian u and T
    conduct_one_step, implicit branch (550-551) n_cond backward-Euler sweeps
    enforce_no_slip, pressure_projection,     fused as projection_step
    enforce_no_slip (606-620)

(buoyancy, bubbles, the conduction's boundary sources and the nutrient
scalars are not part of it).  ``run`` drives N steps on the device and dumps
the state with boilsim.statefile; ``compare_files`` applies the protocol's
per-field limits to two dumps.

    python -m boilsim.regression --grid 48 48 40 --steps 50 --out gpu_state.bsst
"""

from __future__ import annotations

import argparse
from dataclasses import dataclass, field

import numpy as np

from . import fluid, statefile
from .fixtures import pot_mask, seeded_velocities
from .thermal import MaterialProps, allocate_thermal_workspace, conduct_implicit


@dataclass
class StepConfig:
    dt: float = 2.0e-3  # s
    rho: float = 958.0  # kg/m^3
    n_cond: int = 15  # diffusion_max_iter (config.py:184)
    h_conv: float = 10.0  # W/(m^2 K)
    ambient_c: float = 22.0
    pot_material: str = "steel_304"
    solver: fluid.SolverConfig = field(default_factory=fluid.SolverConfig)  # pressure_max_iter, pressure_tol


def initial_state(nx: int, ny: int, nz: int, dx: float, seed: int = 9):
    """The regression's starting state: the pot geometry, N(0, 0.01) face
    velocities with the buoyant plume (seed `seed`), water at 300 K plus up
    to 5 K of seeded noise, a 480 K pot base, air at ambient; origin as the
    reference's compute_grid_dims (geometry.py:320-338) with a 4-cell pad."""
    mat = pot_mask(nx, ny, nz)
    ux, uy, uz = seeded_velocities(nx, ny, nz, seed=seed, plume=True)
    t = np.full((nx, ny, nz), 295.15, np.float32)
    water = mat == 0
    t[water] = 300.0 + 5.0 * np.random.default_rng(seed + 1).random(int(water.sum())).astype(np.float32)
    t[(mat == 1) & (np.arange(nz)[None, None, :] < max(1, nz // 6))] = 480.0
    origin = (-nx * dx / 2.0, -ny * dx / 2.0, -4.0 * dx)
    return {"mat": mat, "ux": ux, "uy": uy, "uz": uz, "T": t, "origin": origin, "dx": dx}


class Run:
    """A device-resident run: grid, workspaces and material tables."""

    def __init__(self, state: dict, cfg: StepConfig):
        import torch

        self.cfg = cfg
        self.core = fluid._get_sim_core()
        self.grid = fluid.make_grid(state["mat"], state["ux"], state["uy"], state["uz"], state["dx"],
                                    T=state["T"], origin=state["origin"])
        self.ws = fluid.allocate_fluid_workspace(self.grid)
        self.tws = allocate_thermal_workspace(torch, self.grid.shape)
        self.props = MaterialProps.from_pot(cfg.pot_material).device(torch)
        self.steps = 0
        self.last_max_div = float("nan")

    def step(self) -> float:
        c, g = self.cfg, self.grid
        fluid.advect_all(g, self.ws, c.dt)
        conduct_implicit(self.core, g.T, g.mat, self.props, self.tws, g.dx, c.dt, c.n_cond, c.h_conv, c.ambient_c)
        self.last_max_div = fluid.projection_step(g, self.ws, fluid.ScenarioConfig(solver=c.solver), c.dt, c.rho)
        self.steps += 1
        return self.last_max_div

    def dump(self, path, **meta) -> None:
        statefile.dump_grid(path, self.grid, self.ws, step=self.steps, dt=self.cfg.dt,
                            max_div=self.last_max_div, u_max=self.ws.u_max, **meta)


def run(state: dict, cfg: StepConfig, steps: int, out=None, every: int = 0) -> Run:
    """`steps` steps from `state`; dumps to `out` at the end (and to
    `out`.stepNNNN every `every` steps if every > 0)."""
    r = Run(state, cfg)
    for s in range(steps):
        r.step()
        if out and every and (s + 1) % every == 0 and s + 1 < steps:
            r.dump(f"{out}.step{s + 1:04d}")
    if out:
        r.dump(out)
    return r


def compare_files(path, ref_path, limits=None) -> dict:
    """The protocol's report between two dumps (see statefile.compare)."""
    a, _ = statefile.load_state(path)
    b, _ = statefile.load_state(ref_path)
    return statefile.compare(a, b, limits=limits)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--grid", type=int, nargs=3, default=[48, 48, 40])
    ap.add_argument("--dx", type=float, default=2.0e-3)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--dt", type=float, default=2.0e-3)
    ap.add_argument("--pressure-iters", type=int, default=200)
    ap.add_argument("--out", required=True)
    ap.add_argument("--every", type=int, default=0)
    ap.add_argument("--compare", help="a reference dump: print the protocol's report against it")
    a = ap.parse_args(argv)
    cfg = StepConfig(dt=a.dt)
    cfg.solver.pressure_max_iter = a.pressure_iters
    r = run(initial_state(*a.grid, a.dx), cfg, a.steps, a.out, a.every)
    print(f"{a.steps} steps on {tuple(a.grid)}: max|div| {r.last_max_div:.3e}, max|u| {r.ws.u_max:.4f} -> {a.out}")
    if a.compare:
        rep = compare_files(a.out, a.compare)
        for k, v in rep.items():
            if isinstance(v, dict):
                print(f"  {k:6s} max_abs {v['max_abs']:.3e} max_rel {v['max_rel']:.3e} bits_equal {v['bits_equal']}"
                      f" limit {v['limit']} {'ok' if v['passes'] else 'FAIL'}")
        print("passed" if rep["passed"] else "FAILED")
        return 0 if rep["passed"] else 1
    return 0


if __name__ == "__main__":
    raise SystemExit(main())
