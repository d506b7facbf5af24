"""Writes tests/golden/reduced_steps_20x20x16.bsst: the state after 5 steps
of advection + conduction + no-slip/Jacobi projection/no-slip, computed by the
oracle from boilsim.regression.initial_state(20, 20, 16, 2e-3), in the
boilsim.statefile format.  The GPU regression test must reproduce it bit for
bit; the CPU test checks that the oracle still does.

    python tests/golden/make_state_fixture.py
"""

from __future__ import annotations

import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "boiling-sim_amd"), os.path.dirname(HERE)]

FIXTURE = os.path.join(HERE, "reduced_steps_20x20x16.bsst")
SHAPE, DX, STEPS, PRESSURE_ITERS = (20, 20, 16), 2.0e-3, 5, 60


def config():
    from boilsim.regression import StepConfig

    cfg = StepConfig()
    cfg.solver.pressure_max_iter = PRESSURE_ITERS
    return cfg


def compute():
    from boilsim.regression import initial_state
    from helpers import OracleRun

    run = OracleRun(initial_state(*SHAPE, DX), config())
    for _ in range(STEPS):
        run.step()
    return run


def main():
    from boilsim import statefile

    run = compute()
    cfg = config()
    statefile.save_state(FIXTURE, run.fields(), {
        "nx": SHAPE[0], "ny": SHAPE[1], "nz": SHAPE[2], "dx": DX, "steps": STEPS, "dt": cfg.dt,
        "n_cond": cfg.n_cond, "pressure_iters": PRESSURE_ITERS, "max_div": run.max_div, "u_max": run.u_max,
        "made_by": "oracle (tests/golden/make_state_fixture.py)"})
    print(f"wrote {FIXTURE} ({os.path.getsize(FIXTURE)} bytes): max|div| {run.max_div:.4e}, max|u| {run.u_max:.4f}")


if __name__ == "__main__":
    main()
