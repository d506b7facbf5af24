"""Held-out evaluation of the learned preconditioner on its training grid.

    python -m boilsim.heldout [--seeds 200] [--out summary.md]

The 200 held-out right-hand sides (Gaussian random fields, seeds
10000..10199, correlation length U[4,16] cells, unit variance over the fluid
cells of the 110x110x70 pot) are each solved to a true relative residual tol
(default 1e-6) from x = 0.  On these smooth fields an FP32 iterate cannot get
there (its rounding alone leaves ~1e-5), so every solver is an FP64 iterative
refinement whose residual is formed in FP64, around one of four inner
solvers with FP32 vectors:

  unet     flexible CG with the network alone as the preconditioner
           (bs_fcg_refine_solve, kind 1); a stalled inner solve hands the
           rest of the solve to plain Jacobi-PCG
  vcycle   flexible CG with the shipped V-cycle whose coarsest level the
           network solves (kind 2)
  mg       the same V-cycle with one Jacobi sweep + Chebyshev polish in place
           of the network (kind 3)
  pcg      plain Jacobi-PCG (bs_refine_solve, FP32 vectors)
"""

from __future__ import annotations

import argparse
import json
import time

import numpy as np

from .learned import PRECOND_MG_JACOBI, PRECOND_MG_UNET, PRECOND_UNET, _UNIT, grf_torch

HELDOUT_SEEDS = range(10000, 10200)
SOLVERS = ("unet", "vcycle", "mg", "pcg")
_KIND = {"unet": PRECOND_UNET, "vcycle": PRECOND_MG_UNET, "mg": PRECOND_MG_JACOBI}


class HeldoutSolver:
    """Workspaces and the four refinement solves on one mask."""

    def __init__(self, core, torch, mat_t, net):
        self.core, self.torch, self.mat, self.net = core, torch, mat_t, net
        self.shape = tuple(mat_t.shape)
        self.fluid = mat_t == 0
        nb = max(core.fcg_refine_workspace_bytes(*self.shape, k) for k in _KIND.values())
        nb = max(nb, core.refine_workspace_bytes(*self.shape))
        self.ws = torch.empty(nb, dtype=torch.uint8, device=mat_t.device)
        self.x64 = torch.zeros(self.shape, dtype=torch.float64, device=mat_t.device)

    def rhs(self, seed: int):
        return grf_torch(self.torch, self.fluid, seed)

    def solve(self, solver: str, b, tol: float = 1e-6, max_inner: int = 40000) -> tuple[dict, float]:
        """Solves A x = b into self.x64 (cold start); returns (stats, ms)."""
        core, torch = self.core, self.torch
        self.x64.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if solver == "pcg":
            st = core.refine_solve(self.x64.data_ptr(), b.data_ptr(), self.mat.data_ptr(), *self.shape, *_UNIT, tol,
                                   self.ws.data_ptr(), self.ws.numel(), storage="fp32", max_outer=60,
                                   max_inner=max_inner)
        else:
            st = core.fcg_refine_solve(self.x64.data_ptr(), b.data_ptr(), self.mat.data_ptr(), *self.shape, *_UNIT,
                                       tol, _KIND[solver], self.net.params.data_ptr(), self.net.packed.data_ptr(),
                                       self.ws.data_ptr(), self.ws.numel(), max_outer=60, max_inner=max_inner)
        torch.cuda.synchronize()
        return st, (time.perf_counter() - t0) * 1e3


def evaluate(core, torch, mat_t, net, seeds=HELDOUT_SEEDS, solvers=SOLVERS, tol: float = 1e-6, log=None) -> dict:
    """Per-solver summary over `seeds` (see the module docstring)."""
    hs = HeldoutSolver(core, torch, mat_t, net)
    rows = {s: [] for s in solvers}
    ells = []
    for n, seed in enumerate(seeds):
        b, ell = hs.rhs(int(seed))
        ells.append(ell)
        for s in solvers:
            st, ms = hs.solve(s, b, tol)
            rows[s].append((ms, st["inner_iterations"], st["outer"], st["converged"], st["relres_true"],
                            st.get("fallback", 0), st.get("pcg_iterations", 0)))
        if log and (n + 1) % 20 == 0:
            log(f"{n + 1}/{len(seeds)} right-hand sides: " + ", ".join(
                f"{s} {np.median([r[0] for r in rows[s]]):.1f} ms" for s in solvers))
    out = {"seeds": [int(seeds[0]), int(seeds[-1])], "n": len(ells), "tol": tol, "shape": list(hs.shape),
           "ell_range": [float(min(ells)), float(max(ells))], "solvers": {}}
    for s in solvers:
        a = np.array(rows[s], dtype=np.float64)
        out["solvers"][s] = {
            "ms_median": float(np.median(a[:, 0])), "ms_mean": float(a[:, 0].mean()), "ms_max": float(a[:, 0].max()),
            "iterations_median": float(np.median(a[:, 1])), "iterations_mean": float(a[:, 1].mean()),
            "iterations_max": int(a[:, 1].max()), "refine_steps_mean": float(a[:, 2].mean()),
            "converged": int(a[:, 3].sum()), "relres_true_max": float(a[:, 4].max()),
            "fallback": int(a[:, 5].sum()), "pcg_iterations_mean": float(a[:, 6].mean()),
        }
    return out


def markdown(summary: dict, cmd: str, device: str) -> str:
    lines = [
        "# Held-out evaluation of the learned preconditioner, 110x110x70 pot",
        "",
        f"Command: `{cmd}`  ",
        f"Device: {device}",
        "",
        f"{summary['n']} held-out right-hand sides (GRF seeds {summary['seeds'][0]}..{summary['seeds'][1]}, "
        f"correlation length {summary['ell_range'][0]:.1f}..{summary['ell_range'][1]:.1f} cells), each solved "
        f"from x = 0 to a true relative residual {summary['tol']:g} by FP64 iterative refinement around the "
        "inner solver named (FP32 vectors; FP64 residual, iterate and dot products).",
        "",
        "| inner solver | converged | median ms | mean ms | median iterations | max iterations | "
        "refinement steps (mean) | stalls handed to PCG | max true rel-res |",
        "|---|---|---|---|---|---|---|---|---|",
    ]
    names = {"unet": "network alone as the flexible-CG preconditioner (kind 1)",
             "vcycle": "V-cycle, network on the coarsest level (kind 2, shipped)",
             "mg": "V-cycle, Jacobi sweep + Chebyshev polish on the coarsest level (kind 3, no network)",
             "pcg": "plain Jacobi-PCG"}
    for s, v in summary["solvers"].items():
        lines.append(f"| {names.get(s, s)} | {v['converged']}/{summary['n']} | {v['ms_median']:.2f} | "
                     f"{v['ms_mean']:.2f} | {v['iterations_median']:.0f} | {v['iterations_max']} | "
                     f"{v['refine_steps_mean']:.2f} | {v['fallback']} | {v['relres_true_max']:.2e} |")
    return "\n".join(lines) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--seeds", type=int, default=200)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--out", default=None, help="write a markdown summary here")
    ap.add_argument("--json", default=None, help="write the summary as JSON here")
    a = ap.parse_args(argv)
    import torch

    import sim_core
    from .fixtures import CONFIGS, pot_mask
    from .learned import default_net

    core = sim_core.cuda.SimCore(0)
    mat = torch.from_numpy(pot_mask(*CONFIGS["dev"].shape)).cuda()
    log = lambda m: print(f"[heldout] {m}", flush=True)
    summary = evaluate(core, torch, mat, default_net(), seeds=range(10000, 10000 + a.seeds), tol=a.tol, log=log)
    log(json.dumps(summary))
    if a.json:
        with open(a.json, "w") as f:
            json.dump(summary, f, indent=1)
    if a.out:
        with open(a.out, "w") as f:
            f.write(markdown(summary, f"python -m boilsim.heldout --seeds {a.seeds} --tol {a.tol:g}",
                             core.device_info()))
    return summary


if __name__ == "__main__":
    main()
