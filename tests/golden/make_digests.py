"""Writes tests/golden/config_digests.json: sha256 digests that freeze the
synthetic inputs (the pot masks of the three grids, the seeded face
velocities) and the oracle's config-1 outputs (64x64x48, seed-7 velocities
with the plume, dt 1e-3, rho 958, dx 2 mm): the divergence, 200 Jacobi
sweeps from p = 0 (p and p_tmp) and the projection's corrected faces and
max|div|; and config 2's (110x110x70 and 440x440x280, the same seed and
sweeps): divergence, p and p_tmp.  The CPU suite checks that the fixture
generators and the oracle still produce them; the GPU suite that the device
path does, bit for bit.

    python tests/golden/make_digests.py
"""

from __future__ import annotations

import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [ROOT, os.path.join(ROOT, "boiling-sim_amd")]

OUT = os.path.join(HERE, "config_digests.json")
SEED, SWEEPS = 7, 200


def sha(a) -> str:
    import numpy as np

    a = np.ascontiguousarray(a)
    return hashlib.sha256(a.view(np.uint8).tobytes()).hexdigest()


def config1():
    """Inputs and oracle outputs of config 1."""
    import numpy as np

    import oracle
    from boilsim.fixtures import CONFIGS, DT, RHO, pot_mask, seeded_velocities

    cfg = CONFIGS["small"]
    mat = pot_mask(*cfg.shape)
    ux, uy, uz = seeded_velocities(*cfg.shape, seed=SEED, plume=True)
    div = oracle.divergence(ux, uy, uz, mat, cfg.dx)
    p = np.zeros(cfg.shape, np.float32)
    p_tmp = np.zeros(cfg.shape, np.float32)
    oracle.pressure_solve(p, p_tmp, div, mat, cfg.dx, DT, RHO, SWEEPS)
    fx, fy, fz = ux.copy(), uy.copy(), uz.copy()
    _, _, div_after, max_div = oracle.projection(fx, fy, fz, mat, cfg.dx, DT, RHO, SWEEPS)
    return {"mat": mat, "ux": ux, "uy": uy, "uz": uz, "div": div, "p": p, "p_tmp": p_tmp,
            "ux_after": fx, "uy_after": fy, "uz_after": fz, "div_after": div_after,
            "max_div_after": float(max_div)}


def config2(name: str) -> dict:
    """Config 2 on the 110x110x70 ("dev") or 440x440x280 ("prod") pot:
    seed-7 velocities with the plume, the oracle's divergence and 200 Jacobi
    sweeps from p = 0 (p and p_tmp)."""
    import numpy as np

    import oracle
    from boilsim.fixtures import CONFIGS, DT, RHO, pot_mask, seeded_velocities

    cfg = CONFIGS[name]
    mat = pot_mask(*cfg.shape)
    ux, uy, uz = seeded_velocities(*cfg.shape, seed=SEED, plume=True)
    div = oracle.divergence(ux, uy, uz, mat, cfg.dx)
    del ux, uy
    p = np.zeros(cfg.shape, np.float32)
    p_tmp = np.zeros(cfg.shape, np.float32)
    oracle.pressure_solve(p, p_tmp, div, mat, cfg.dx, DT, RHO, SWEEPS)
    return {"uz": uz, "div": div, "p": p, "p_tmp": p_tmp}


def reference_gate() -> dict:
    """The reference's 200-step gate case (python/tests/test_pressure_parity.py:
    159-224): single_carrot.yaml voxelised at dx = 4 mm (the oracle's
    restatement of build_pot_geometry), N(0, 0.05) seed-7 velocities, 200
    Jacobi iterations, dt 0.01, rho 997; the oracle's projection."""
    from oracle import pot_geometry as pg

    import oracle

    mat, _ = pg.build_pot_mask(pg.GATE_DX)
    ux, uy, uz = pg.gate_velocities(mat.shape)
    d = {"mat": mat, "ux": ux, "uy": uy, "uz": uz}
    fx, fy, fz = ux.copy(), uy.copy(), uz.copy()
    p, p_tmp, div_after, max_div = oracle.projection(fx, fy, fz, mat, pg.GATE_DX, pg.GATE_DT, pg.GATE_RHO,
                                                     pg.GATE_ITERS)
    d.update({"p": p, "p_tmp": p_tmp, "ux_after": fx, "uy_after": fy, "uz_after": fz, "div_after": div_after})
    return {"shape": list(mat.shape), "max_div_after": max_div, "sha256": {k: sha(v) for k, v in d.items()}}


def compute() -> dict:
    from boilsim.fixtures import CONFIGS, pot_mask

    out = {"masks": {name: sha(pot_mask(*c.shape)) for name, c in CONFIGS.items()}}
    c1 = config1()
    out["config1"] = {"seed": SEED, "sweeps": SWEEPS, "max_div_after": c1.pop("max_div_after"),
                      "sha256": {k: sha(v) for k, v in c1.items()}}
    out["config2"] = {}
    for name in ("dev", "prod"):
        out["config2"][name] = {"seed": SEED, "sweeps": SWEEPS,
                                "sha256": {k: sha(v) for k, v in config2(name).items()}}
    out["reference_gate"] = reference_gate()
    return out


def main():
    d = compute()
    with open(OUT, "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)
        f.write("\n")
    print(f"wrote {OUT}")


if __name__ == "__main__":
    main()
