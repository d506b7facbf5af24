"""Time sequence of correlated pressure right-hand sides and a warm-started
solve driver.

Each frame's right-hand side is an AR(1) Gaussian random field
f_t = 0.95 f_{t-1} + sqrt(1 - 0.95^2) xi_t (xi_t a unit-variance GRF with
correlation length 12 cells, masked to the fluid) plus a Gaussian plume
source that drifts around the pot axis.  Consecutive pressures are close, so
every solve starts from the previous FP64 solution, as a time-stepping
simulation's projection would.  Because these fields are smooth, |x| is
some 10^2 times |b| and an FP32 iterate cannot reach a 1e-6 relative
residual; both solvers here are FP64 iterative refinements (``refine_solve``
around Jacobi-PCG, ``fcg_refine_solve`` around the learned flexible CG).

Run the full 100-frame sequence at 440x440x280 with
``python -m boilsim.sequence --frames 100``.
"""

from __future__ import annotations

import argparse
import json
import math
import time

import numpy as np

from .learned import _UNIT, PRECOND_MG_CHEB, PRECOND_MG_UNET, grf_torch

STEP_CORR = 0.95
ELL = 12.0


def plume_source(torch, shape, t, period=100, device="cuda"):
    """Gaussian blob of amplitude 2 and width 8*nx/64 cells whose centre circles
    the pot axis at half the inner radius, mid-depth, once per `period` frames."""
    nx, ny, nz = shape
    sx, sz = nx / 64.0, nz / 48.0
    ang = 2.0 * math.pi * t / period
    cx = nx / 2.0 + 13.0 * sx * math.cos(ang)
    cy = ny / 2.0 + 13.0 * sx * math.sin(ang)
    cz = 22.0 * sz
    sig = 8.0 * sx
    X = torch.arange(nx, device=device, dtype=torch.float32)[:, None, None] + 0.5 - cx
    Y = torch.arange(ny, device=device, dtype=torch.float32)[None, :, None] + 0.5 - cy
    Z = torch.arange(nz, device=device, dtype=torch.float32)[None, None, :] + 0.5 - cz
    return 2.0 * torch.exp(-(X * X + Y * Y + Z * Z) / (2.0 * sig * sig))


def ar1_frames(torch, mat_d, n_frames, seed=0, ell=ELL):
    """Yields (t, b_t) on the device, b_t float32 zero off the fluid."""
    fluid = mat_d == 0
    a = STEP_CORR
    c = math.sqrt(1.0 - a * a)
    f = None
    for t in range(n_frames):
        xi, _ = grf_torch(torch, fluid, seed=seed * 100003 + t, ell=ell)
        f = xi if f is None else a * f + c * xi
        b = f + plume_source(torch, tuple(mat_d.shape), t, device=mat_d.device)
        b = torch.where(fluid, b, torch.zeros_like(b)).contiguous()
        yield t, b


class SequenceSolver:
    """Warm-started FP64 solves of consecutive frames with one method:
    "learned" (refinement around the learned flexible CG), "mg" (the same
    flexible CG and V-cycle with 60 Chebyshev steps instead of the network on
    the coarsest level) or "pcg-fp32" / "pcg-fp16" (refinement around
    Jacobi-PCG with that vector storage)."""

    def __init__(self, core, torch, mat_d, method, net=None, tol=1e-6):
        self.core, self.torch, self.mat_d, self.method, self.net, self.tol = core, torch, mat_d, method, net, tol
        self.shape = tuple(mat_d.shape)
        self.precond = {"learned": PRECOND_MG_UNET, "mg": PRECOND_MG_CHEB}.get(method)
        if self.precond is not None:
            nbytes = core.fcg_refine_workspace_bytes(*self.shape, self.precond)
        else:
            nbytes = core.refine_workspace_bytes(*self.shape)
        self.ws = torch.empty(nbytes, dtype=torch.uint8, device="cuda")
        self.x = torch.zeros(self.shape, dtype=torch.float64, device="cuda")
        self.started = False

    def solve(self, b, tol=None, x=None, warm=None) -> dict:
        tol = self.tol if tol is None else tol
        x = self.x if x is None else x
        dx, dt, rho = _UNIT
        args = (x.data_ptr(), b.data_ptr(), self.mat_d.data_ptr(), *self.shape, dx, dt, rho, tol)
        if warm is None:
            warm = self.started or x is not self.x
        if self.precond is not None:
            net = self.net if self.precond == PRECOND_MG_UNET else None
            st = self.core.fcg_refine_solve(*args, self.precond, net.params.data_ptr() if net else 0,
                                            net.packed.data_ptr() if net else 0, self.ws.data_ptr(),
                                            self.ws.numel(), inner_tol=1e-3, warm_start=warm)
        else:
            st = self.core.refine_solve(*args, self.ws.data_ptr(), self.ws.numel(), storage=self.method[4:],
                                        warm_start=warm)
        if x is self.x:
            self.started = True
        return st


def run_sequence(core, torch, mat_d, n_frames, methods=("learned", "pcg-fp32"), net=None, tol=1e-6,
                 error_every=0, seed=0, log=None):
    """Solves every frame with each method (warm-started from its own previous
    solution).  Every `error_every` frames (and the last) the error of each
    method is stated against an FP64 reference solve of the same frame from
    x = 0 (independent of every method's iterate) to a relative residual of
    1e-11.  Returns {method: [per-frame dict]}."""
    solvers = {m: SequenceSolver(core, torch, mat_d, m, net=net, tol=tol) for m in methods}
    ref = SequenceSolver(core, torch, mat_d, "pcg-fp32", tol=1e-11)
    out = {m: [] for m in methods}
    fluid = mat_d == 0
    for t, b in ar1_frames(torch, mat_d, n_frames, seed=seed):
        check = t == n_frames - 1 or (error_every and t % error_every == 0)
        for m, s in solvers.items():
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            st = s.solve(b)
            torch.cuda.synchronize()
            rec = {"frame": t, "ms": (time.perf_counter() - t0) * 1e3, **st}
            out[m].append(rec)
        if check:
            xr = torch.zeros_like(solvers[methods[0]].x)
            ref.solve(b, x=xr, warm=False)
            nr = torch.linalg.vector_norm(xr[fluid]).item()
            for m, s in solvers.items():
                out[m][-1]["err_vs_fp64"] = torch.linalg.vector_norm((s.x - xr)[fluid]).item() / max(nr, 1e-300)
        if log:
            log(t, {m: out[m][-1] for m in methods})
    return out


def summarize(res: dict) -> dict:
    s = {}
    for m, recs in res.items():
        ms = np.array([r["ms"] for r in recs])
        its = np.array([r["inner_iterations"] for r in recs])
        errs = [r["err_vs_fp64"] for r in recs if "err_vs_fp64" in r]
        s[m] = {
            "frames": len(recs),
            "ms_per_solve_median": float(np.median(ms)),
            "ms_per_solve_mean": float(ms.mean()),
            "ms_first_frame": float(ms[0]),
            "iterations_mean": float(its.mean()),
            "iterations_first_frame": int(its[0]),
            "all_converged": bool(all(r["converged"] for r in recs)),
            "max_relres_true": float(max(r["relres_true"] for r in recs)),
            "max_err_vs_fp64": float(max(errs)) if errs else None,
        }
    return s


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--frames", type=int, default=100)
    ap.add_argument("--grid", default="prod", choices=["small", "dev", "prod"])
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--methods", default="learned,mg,pcg-fp32,pcg-fp16")
    ap.add_argument("--error-every", type=int, default=10)
    ap.add_argument("--out", default=None, help="JSON file for the per-frame records and the summary")
    a = ap.parse_args(argv)
    import torch

    import sim_core

    from .fixtures import CONFIGS, pot_mask
    from .unet import UNet

    if sim_core.cuda is None:
        raise RuntimeError("libboilsim_hip.so is not loadable")
    core = sim_core.cuda.SimCore(0)
    cfg = CONFIGS[a.grid]
    mat_d = torch.from_numpy(pot_mask(*cfg.shape)).cuda()
    methods = tuple(a.methods.split(","))

    def log(t, recs):
        parts = [f"{m}: {r['ms']:7.1f} ms {r['inner_iterations']:5d} its" +
                 (f" err {r['err_vs_fp64']:.2e}" if "err_vs_fp64" in r else "") for m, r in recs.items()]
        print(f"frame {t:3d}  " + "  |  ".join(parts), flush=True)

    res = run_sequence(core, torch, mat_d, a.frames, methods=methods, net=UNet.load(), tol=a.tol,
                       error_every=a.error_every, log=log)
    summ = summarize(res)
    print(json.dumps(summ, indent=1))
    if a.out:
        with open(a.out, "w") as fh:
            json.dump({"grid": cfg.shape, "tol": a.tol, "summary": summ, "frames": res}, fh)


if __name__ == "__main__":
    main()
