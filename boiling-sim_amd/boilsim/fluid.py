"""pressure_projection on MI355X — the host driver of the reference's
python/boilingsim/fluid.py:500-734, over torch device tensors and the HIP
library behind ``sim_core.cuda.SimCore``.

Solver selection keeps the reference's environment contract
(fluid.py:585-609): ``BOILINGSIM_USE_RUST_PRESSURE=1`` selects the fused
Jacobi solve, ``BOILINGSIM_PRESSURE_SOLVER`` is ``jacobi`` (default) or
``cg-rust`` (the reference-semantics PCG), and setting both raises
``ValueError``.  Two values are added: ``pcg-hip`` (Jacobi-PCG to a true
relative residual ``pressure_tol``, FP64 residual check) and ``lpcg-hip``
(flexible PCG with the learned preconditioner, falling back to ``pcg-hip``).
The Warp kernels of the reference (divergence, gradient subtraction) run as
HIP kernels here, and the returned max|div| is reduced on the device instead of
copying the whole field to the host (fluid.py:733-734).

The faces around the projection: ``enforce_no_slip`` (fluid.py:111-162) is one
launch, ``compute_max_velocity`` / ``compute_cfl_dt`` (fluid.py:796-844) reduce
on the device, and ``projection_step`` runs a step's no-slip, projection,
no-slip (pipeline.py:606-620) with the first no-slip fused into the
divergence pass, the second one omitted because it would change no bit, and
max|u| for the next step's CFL time step taken from the diagnostic divergence
pass and read back with max|div|.
"""

from __future__ import annotations

import os
from dataclasses import dataclass, field

import numpy as np

from .geometry import MAT_AIR, MAT_FLUID

VALID_SOLVERS = ("jacobi", "cg-rust", "pcg-hip", "lpcg-hip", "mgpcg-hip")


@dataclass
class SolverConfig:
    """The solver knobs the projection and the CFL time step read
    (python/boilingsim/config.py:161-190)."""

    pressure_tol: float = 1e-5
    pressure_max_iter: int = 200
    cfl_safety_factor: float = 0.4  # config.py:164
    max_dt_s: float = 0.1  # config.py:165


@dataclass
class ScenarioConfig:
    solver: SolverConfig = field(default_factory=SolverConfig)


@dataclass
class Grid:
    """Cell fields (nx,ny,nz) and MAC faces, as torch tensors on one device
    (python/boilingsim/geometry.py:254-315)."""

    nx: int
    ny: int
    nz: int
    dx: float
    p: object
    mat: object
    ux: object
    uy: object
    uz: object
    T: object = None  # cell temperature (K), advected with the velocity
    origin: tuple = (0.0, 0.0, 0.0)  # world position of cell corner (0,0,0), m

    @property
    def shape(self) -> tuple[int, int, int]:
        return (self.nx, self.ny, self.nz)


@dataclass
class FluidWorkspace:
    """Projection scratch (fluid.py:500-525, the fields the projection uses)."""

    div_u: object
    p_tmp: object
    max_bits: object  # int32[2]: max|div| and max|u| as float bits
    u_max: float = 0.0  # max|u| after the last projection_step
    # advect_all's scratch (fluid.py:505-524), allocated on first use
    ux_tmp: object = None
    uy_tmp: object = None
    uz_tmp: object = None
    T_tmp: object = None
    T_ext: object = None


def make_grid(mat: np.ndarray, ux: np.ndarray, uy: np.ndarray, uz: np.ndarray, dx: float,
              device: str = "cuda", T: np.ndarray | None = None, origin=(0.0, 0.0, 0.0)) -> Grid:
    import torch

    nx, ny, nz = mat.shape
    t = lambda a, dt: torch.from_numpy(np.ascontiguousarray(a, dtype=dt)).to(device)
    return Grid(nx, ny, nz, float(dx), torch.zeros((nx, ny, nz), dtype=torch.float32, device=device),
                t(mat, np.int32), t(ux, np.float32), t(uy, np.float32), t(uz, np.float32),
                None if T is None else t(T, np.float32), tuple(float(c) for c in origin))


def allocate_fluid_workspace(grid: Grid) -> FluidWorkspace:
    import torch

    dev = grid.p.device
    return FluidWorkspace(
        div_u=torch.zeros(grid.shape, dtype=torch.float32, device=dev),
        p_tmp=torch.zeros(grid.shape, dtype=torch.float32, device=dev),
        max_bits=torch.zeros(2, dtype=torch.int32, device=dev),
    )


_SIM_CORE = None


def _get_sim_core():
    """Process-global SimCore(0), as _get_rust_sim (fluid.py:533-547)."""
    global _SIM_CORE
    if _SIM_CORE is None:
        import sim_core

        if sim_core.cuda is None:
            raise RuntimeError(
                "The HIP pressure path needs sim_core.cuda, which is unavailable (no HIP device, "
                "or libboilsim_hip.so not built). Build it with __graft_entry__.build()."
            )
        _SIM_CORE = sim_core.cuda.SimCore(0)
    return _SIM_CORE


_PCG_WORKSPACE = {"shape": None, "buffers": None}


def _get_pcg_workspace(shape, device="cuda"):
    """The reference's PCG workspace (fluid.py:557-582): 5 n-vectors, a
    1024-float dot workspace and 7 one-float scalars, cached by shape."""
    import torch

    if _PCG_WORKSPACE["shape"] == tuple(shape) and _PCG_WORKSPACE["buffers"] is not None:
        return _PCG_WORKSPACE["buffers"]
    z = lambda *s: torch.zeros(s, dtype=torch.float32, device=device)
    bufs = {k: z(*shape) for k in ("b", "r", "z", "p_search", "Ap")}
    bufs["dot_ws"] = z(1024)
    for k in ("alpha", "beta", "rzold", "rznew", "bsq", "rsq", "pAp"):
        bufs[k] = z(1)
    _PCG_WORKSPACE["shape"] = tuple(shape)
    _PCG_WORKSPACE["buffers"] = bufs
    return bufs


_SOLVE_WORKSPACE = {"shape": None, "buf": None}


def _get_solve_workspace(core, shape, device="cuda"):
    import torch

    if _SOLVE_WORKSPACE["shape"] == tuple(shape):
        return _SOLVE_WORKSPACE["buf"]
    nb = core.pcg_workspace_bytes(*shape)
    buf = torch.empty(nb, dtype=torch.uint8, device=device)
    _SOLVE_WORKSPACE["shape"] = tuple(shape)
    _SOLVE_WORKSPACE["buf"] = buf
    return buf


def _select_pressure_solver() -> str:
    """"jacobi", "rust-jacobi", "cg-rust", "pcg-hip", "lpcg-hip" or
    "mgpcg-hip".  The reference's contract (fluid.py:585-609): an unknown
    BOILINGSIM_PRESSURE_SOLVER, or BOILINGSIM_USE_RUST_PRESSURE=1 together
    with a solver other than jacobi, raises ValueError."""
    use_rust_jacobi = os.environ.get("BOILINGSIM_USE_RUST_PRESSURE") == "1"
    solver = os.environ.get("BOILINGSIM_PRESSURE_SOLVER", "jacobi").lower()
    if solver not in VALID_SOLVERS:
        raise ValueError(
            f"unknown pressure solver BOILINGSIM_PRESSURE_SOLVER={solver!r}; expected one of "
            + ", ".join(repr(v) for v in VALID_SOLVERS) + " ('jacobi' when unset)"
        )
    if use_rust_jacobi and solver != "jacobi":
        raise ValueError(
            f"conflicting pressure solver settings: BOILINGSIM_USE_RUST_PRESSURE=1 selects the fused "
            f"Jacobi solve, BOILINGSIM_PRESSURE_SOLVER={solver} another solver; set only one of them"
        )
    if solver != "jacobi":
        return solver
    if use_rust_jacobi:
        return "rust-jacobi"
    return "jacobi"


LAST_SOLVE_STATS: dict = {}


def _solve(core, solver, grid: Grid, ws: FluidWorkspace, cfg: ScenarioConfig, dt, rho, preconditioner):
    """The non-Jacobi solvers of pressure_projection (fluid.py:643-687): p from
    ws.div_u."""
    nx, ny, nz = grid.shape
    dx = grid.dx
    mi = int(cfg.solver.pressure_max_iter)
    if solver == "cg-rust":
        w = _get_pcg_workspace(grid.shape, grid.p.device)
        iters = core.pressure_solve_pcg(
            grid.p.data_ptr(), ws.div_u.data_ptr(), grid.mat.data_ptr(), nx, ny, nz, dx, dt, rho,
            float(cfg.solver.pressure_tol), mi,
            *(w[k].data_ptr() for k in ("b", "r", "z", "p_search", "Ap", "dot_ws", "alpha", "beta",
                                        "rzold", "rznew", "bsq", "rsq", "pAp")),
            MAT_FLUID, MAT_AIR)
        st = {"iterations": iters}
    elif solver == "pcg-hip":
        buf = _get_solve_workspace(core, grid.shape, grid.p.device)
        st = core.pcg_solve(grid.p.data_ptr(), ws.div_u.data_ptr(), grid.mat.data_ptr(), nx, ny, nz,
                            dx, dt, rho, float(cfg.solver.pressure_tol), mi, buf.data_ptr(),
                            buf.numel(), warm_start=False)
    elif solver == "mgpcg-hip":
        from .learned import PRECOND_MG_CHEB, learned_pcg_solve

        st = learned_pcg_solve(core, grid, ws, dt, rho, float(cfg.solver.pressure_tol), mi, precond=PRECOND_MG_CHEB)
    else:  # lpcg-hip
        from .learned import learned_pcg_solve

        st = learned_pcg_solve(core, grid, ws, dt, rho, float(cfg.solver.pressure_tol), mi, preconditioner)
    LAST_SOLVE_STATS.clear()
    LAST_SOLVE_STATS.update({"solver": solver, **st})


def _bits_to_floats(t) -> np.ndarray:
    return t.cpu().numpy().view(np.float32)


def pressure_projection(grid: Grid, ws: FluidWorkspace, cfg: ScenarioConfig, dt: float, rho: float,
                        device: str = "cuda", preconditioner=None) -> float:
    """Enforce incompressibility (fluid.py:612-734); returns max|div u| after
    the correction.  ``preconditioner`` is the trained network used by
    ``lpcg-hip``."""
    core = _get_sim_core()
    nx, ny, nz = grid.shape
    dx = grid.dx
    solver = _select_pressure_solver()
    mi = int(cfg.solver.pressure_max_iter)
    if solver in ("jacobi", "rust-jacobi"):
        # One fused call: zero p, divergence, sweeps, gradient, divergence, max.
        core.projection_jacobi(grid.p.data_ptr(), ws.p_tmp.data_ptr(), ws.div_u.data_ptr(),
                               grid.ux.data_ptr(), grid.uy.data_ptr(), grid.uz.data_ptr(),
                               grid.mat.data_ptr(), nx, ny, nz, dx, dt, rho, mi, ws.max_bits.data_ptr(),
                               MAT_FLUID, MAT_AIR)
    else:
        core.compute_divergence(ws.div_u.data_ptr(), grid.ux.data_ptr(), grid.uy.data_ptr(),
                                grid.uz.data_ptr(), grid.mat.data_ptr(), nx, ny, nz, dx, MAT_FLUID)
        _solve(core, solver, grid, ws, cfg, dt, rho, preconditioner)
        core.subtract_pressure_gradient(grid.ux.data_ptr(), grid.uy.data_ptr(), grid.uz.data_ptr(),
                                        grid.p.data_ptr(), grid.mat.data_ptr(), nx, ny, nz, dx, dt, rho,
                                        MAT_FLUID)
        # the diagnostic divergence and its max|div| in one pass (max|u| too)
        core.divergence_max(ws.div_u.data_ptr(), grid.ux.data_ptr(), grid.uy.data_ptr(), grid.uz.data_ptr(),
                            grid.mat.data_ptr(), nx, ny, nz, dx, ws.max_bits.data_ptr(), MAT_FLUID)
    return float(_bits_to_floats(ws.max_bits)[0])


def enforce_no_slip(grid: Grid, device: str = "cuda") -> None:
    """Zero every MAC face that touches a non-fluid cell or the domain edge
    (fluid.py:153-162), in one launch."""
    core = _get_sim_core()
    core.enforce_no_slip(grid.ux.data_ptr(), grid.uy.data_ptr(), grid.uz.data_ptr(), grid.mat.data_ptr(),
                         *grid.shape, MAT_FLUID)


def compute_max_velocity(grid: Grid, ws: FluidWorkspace | None = None) -> float:
    """max|u| over all MAC faces (fluid.py:807-833), reduced on the device;
    one 4-byte readback."""
    import torch

    core = _get_sim_core()
    out = ws.max_bits[1:] if ws is not None else torch.zeros(1, dtype=torch.int32, device=grid.ux.device)
    core.max_velocity(grid.ux.data_ptr(), grid.uy.data_ptr(), grid.uz.data_ptr(), *grid.shape, out.data_ptr())
    return float(_bits_to_floats(out)[0])


def compute_cfl_dt(grid: Grid, cfg: ScenarioConfig, ws: FluidWorkspace | None = None,
                   u_max: float | None = None) -> float:
    """CFL-limited time step from the current max velocity (fluid.py:836-844).
    ``u_max`` may be passed in, e.g. ``ws.u_max`` after ``projection_step``,
    to skip the reduction."""
    if u_max is None:
        u_max = compute_max_velocity(grid, ws=ws)
    if u_max < 1e-8:
        return cfg.solver.max_dt_s
    return cfg.solver.cfl_safety_factor * grid.dx / u_max


def projection_step(grid: Grid, ws: FluidWorkspace, cfg: ScenarioConfig, dt: float, rho: float,
                    device: str = "cuda", preconditioner=None) -> float:
    """enforce_no_slip, pressure_projection, enforce_no_slip (pipeline.py:
    606-620) with identical results in fewer passes: the first no-slip runs
    inside the first divergence pass, the second is omitted (the faces it
    would zero are +0.0 already: the first pass zeroed them and the gradient
    subtraction writes only interior faces between fluid cells), and the
    diagnostic divergence pass also reduces max|u|.  Returns max|div u| as
    pressure_projection does and leaves max|u| in ``ws.u_max``, both from one
    8-byte readback."""
    core = _get_sim_core()
    nx, ny, nz = grid.shape
    dx = grid.dx
    solver = _select_pressure_solver()
    faces = (grid.ux.data_ptr(), grid.uy.data_ptr(), grid.uz.data_ptr())
    if solver in ("jacobi", "rust-jacobi"):
        core.projection_step_jacobi(grid.p.data_ptr(), ws.p_tmp.data_ptr(), ws.div_u.data_ptr(), *faces,
                                    grid.mat.data_ptr(), nx, ny, nz, dx, dt, rho,
                                    int(cfg.solver.pressure_max_iter), ws.max_bits.data_ptr(), MAT_FLUID,
                                    MAT_AIR)
    else:
        core.divergence_no_slip(ws.div_u.data_ptr(), *faces, grid.mat.data_ptr(), nx, ny, nz, dx, MAT_FLUID)
        _solve(core, solver, grid, ws, cfg, dt, rho, preconditioner)
        core.subtract_pressure_gradient(*faces, grid.p.data_ptr(), grid.mat.data_ptr(), nx, ny, nz, dx, dt,
                                        rho, MAT_FLUID)
        core.divergence_max(ws.div_u.data_ptr(), *faces, grid.mat.data_ptr(), nx, ny, nz, dx,
                            ws.max_bits.data_ptr(), MAT_FLUID)
    m = _bits_to_floats(ws.max_bits)
    ws.u_max = float(m[1])
    return float(m[0])


def advect_all(grid: Grid, ws: FluidWorkspace, dt: float, device: str = "cuda") -> None:
    """Semi-Lagrangian advection of u and T (fluid.py:737-775): the extended
    temperature, the four advections in one launch into the workspace, and
    the results copied back so grid.ux/uy/uz/T hold them."""
    import torch

    core = _get_sim_core()
    if grid.T is None:
        raise ValueError("advect_all needs grid.T (the cell temperature)")
    if ws.ux_tmp is None:
        z = lambda a: torch.empty_like(a)
        ws.ux_tmp, ws.uy_tmp, ws.uz_tmp = z(grid.ux), z(grid.uy), z(grid.uz)
        ws.T_tmp, ws.T_ext = z(grid.T), z(grid.T)
    core.advect_all(grid.ux.data_ptr(), grid.uy.data_ptr(), grid.uz.data_ptr(), grid.T.data_ptr(),
                    ws.ux_tmp.data_ptr(), ws.uy_tmp.data_ptr(), ws.uz_tmp.data_ptr(), ws.T_tmp.data_ptr(),
                    ws.T_ext.data_ptr(), grid.mat.data_ptr(), *grid.shape, grid.origin, grid.dx, dt, MAT_FLUID)
