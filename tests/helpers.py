"""Test helpers: seeded inputs restated from the reference's own tests and
host<->device plumbing through torch."""

from __future__ import annotations

import numpy as np

# (seed, shape) of the reference's one-step parity cases
# (python/tests/test_pressure_parity.py:112-121).
PARITY_CASES = [
    (1, (16, 16, 16)),
    (2, (24, 12, 20)),
    (3, (8, 32, 8)),
    (42, (20, 20, 20)),
    (101, (12, 24, 16)),
    (202, (16, 12, 24)),
    (303, (28, 16, 14)),
    (404, (10, 10, 30)),
]


def make_random_grid(rng: np.random.Generator, shape):
    """The reference's ``_make_random_grid`` draw sequence
    (test_pressure_parity.py:51-67): air ring, interior 60/15/20/5 % of
    fluid/wall/air/carrot, p_old ~ N(0,1), div ~ 0.1 N(0,1), all from one rng."""
    nx, ny, nz = shape
    mat = np.full(shape, 2, dtype=np.int32)
    choice = rng.choice([0, 1, 2, 3], size=(nx - 2, ny - 2, nz - 2),
                        p=[0.60, 0.15, 0.20, 0.05]).astype(np.int32)
    mat[1:-1, 1:-1, 1:-1] = choice
    p_old = rng.standard_normal(shape).astype(np.float32)
    div_u = rng.standard_normal(shape).astype(np.float32) * 0.1
    return mat, p_old, div_u


def parity_case(seed, shape):
    """Inputs and physical parameters of one reference parity case
    (test_pressure_parity.py:136-141)."""
    rng = np.random.default_rng(seed)
    mat, p_old, div_u = make_random_grid(rng, shape)
    dx = float(rng.uniform(0.0005, 0.004))
    dt = float(rng.uniform(0.0001, 0.01))
    rho = float(rng.uniform(900.0, 1100.0))
    return mat, p_old, div_u, dx, dt, rho


def fluid_everywhere_edge_case(seed, shape):
    """Fluid touching the domain edge (no air ring): out-of-bounds neighbours
    must act as air."""
    rng = np.random.default_rng(seed)
    mat = rng.choice([0, 1, 2, 3], size=shape, p=[0.7, 0.1, 0.15, 0.05]).astype(np.int32)
    p_old = rng.standard_normal(shape).astype(np.float32)
    div_u = (rng.standard_normal(shape) * 0.1).astype(np.float32)
    return mat, p_old, div_u


def dev(torch, a):
    return torch.from_numpy(np.ascontiguousarray(a)).cuda()


def host(t):
    return t.detach().cpu().numpy()


def same_bits(a, b) -> bool:
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def bit_diff_report(a, b) -> str:
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    bad = a.view(np.uint32) != b.view(np.uint32)
    if not bad.any():
        return "identical"
    idx = np.argwhere(bad)[:5]
    return (f"{int(bad.sum())} cells differ; first {idx.tolist()}: "
            f"{a[tuple(idx[0])]!r} vs {b[tuple(idx[0])]!r}; max abs diff {np.abs(a - b).max():.3e}")


class OracleRun:
    """The oracle's counterpart of boilsim.regression.Run: advection,
    backward-Euler conduction and no-slip + Jacobi projection + no-slip per
    step, on host arrays.  The conduction's ping-pong partner and snapshot
    persist across steps like the device workspace (an odd sweep count copies
    the partner into T, air cells included, as conduct_one_step does)."""

    def __init__(self, state: dict, cfg):
        import oracle
        from boilsim.thermal import MaterialProps

        self.oracle = oracle
        self.cfg = cfg
        self.mat = np.ascontiguousarray(state["mat"], np.int32)
        self.f = {k: np.array(state[k], np.float32) for k in ("ux", "uy", "uz", "T")}
        self.origin, self.dx = state["origin"], state["dx"]
        self.t_alt = np.zeros_like(self.f["T"])
        self.t_old = np.zeros_like(self.f["T"])
        self.props = MaterialProps.from_pot(cfg.pot_material)
        self.p = np.zeros_like(self.f["T"])
        self.div = np.zeros_like(self.f["T"])
        self.max_div = self.u_max = float("nan")

    def step(self):
        o, c, f = self.oracle, self.cfg, self.f
        f["ux"], f["uy"], f["uz"], f["T"], _ = o.advect(f["ux"], f["uy"], f["uz"], f["T"], self.mat, self.origin,
                                                        self.dx, c.dt)
        o.conduction_solve(f["T"], self.t_alt, self.t_old, self.mat, self.props.k, self.props.rho, self.props.c_p,
                           self.dx, c.dt, c.h_conv, c.ambient_c + 273.15, c.n_cond)
        self.p, _, self.div, self.max_div, self.u_max = o.projection_step(
            f["ux"], f["uy"], f["uz"], self.mat, self.dx, c.dt, c.rho, int(c.solver.pressure_max_iter))

    def fields(self) -> dict:
        return {"mat": self.mat, "p": self.p, "ux": self.f["ux"], "uy": self.f["uy"], "uz": self.f["uz"],
                "T": self.f["T"], "div_u": self.div}
