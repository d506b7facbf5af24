"""Pressure-projection benchmark on one MI355X.

Default run: the 440x440x280 pot (dx = 0.5 mm, ~54.2 M cells, FP32), inputs
already resident in HBM, and
  * the Jacobi projection leg: divergence -> 200 masked Jacobi sweeps ->
    gradient subtraction -> divergence -> on-device max|div|;
  * roofline of the dominant kernel (the mask-free Jacobi sweep), timed with
    HIP events on the stream it is launched on;
  * the CPU baseline: the oracle's OpenMP restatement on a bounded sample.

Prints ONE JSON line (see the repository's bench contract).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
for _p in (ROOT, os.path.join(ROOT, "boiling-sim_amd")):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0  # MI355X published HBM3E bandwidth, GB/s


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="prod", choices=["small", "dev", "prod"])
    ap.add_argument("--sweeps", type=int, default=200)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--cpu-threads", type=int, default=0, help="0 = all cores in the affinity mask")
    return ap.parse_args()


def cpu_baseline(cfg, mat, div, n_sweeps_sample: int, threads: int):
    """Oracle Jacobi sweeps on the host cores: a bounded sample of the same
    workload, scaled to ms per 200-sweep solve."""
    import oracle

    oracle.set_threads(threads)
    p = np.zeros(cfg.shape, np.float32)
    q = np.zeros(cfg.shape, np.float32)
    t0 = time.perf_counter()
    oracle.pressure_solve(p, q, div, mat, cfg.dx, 1e-3, 958.0, n_sweeps_sample)
    dt = time.perf_counter() - t0
    return dt / n_sweeps_sample * 1e3


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; multi-GPU runs are not part of it",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    import sim_core
    from boilsim.fixtures import CONFIGS, DT, RHO, pot_mask, seeded_velocities

    if sim_core.cuda is None:
        raise RuntimeError("libboilsim_hip.so not loadable: build first")
    core = sim_core.cuda.SimCore(0)
    log(core.device_info())
    cfg = CONFIGS[args.config]
    nx, ny, nz = cfg.shape
    n = cfg.n
    log(f"generating {cfg.shape} pot + seeded velocities")
    mat = pot_mask(nx, ny, nz)
    ux, uy, uz = seeded_velocities(nx, ny, nz, seed=7)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    gmat, gux, guy, guz = d(mat), d(ux), d(uy), d(uz)
    p = torch.zeros(cfg.shape, dtype=torch.float32, device="cuda")
    p_tmp = torch.zeros_like(p)
    div = torch.zeros_like(p)
    bits = torch.zeros(1, dtype=torch.int32, device="cuda")
    torch.cuda.synchronize()

    def projection():
        core.projection_jacobi(p.data_ptr(), p_tmp.data_ptr(), div.data_ptr(), gux.data_ptr(),
                               guy.data_ptr(), guz.data_ptr(), gmat.data_ptr(), nx, ny, nz,
                               cfg.dx, DT, RHO, args.sweeps, bits.data_ptr())

    for _ in range(args.warmup):
        projection()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    t0 = time.perf_counter()
    e0.record()
    for _ in range(args.steps):
        projection()
    e1.record()
    torch.cuda.synchronize()
    wall = (time.perf_counter() - t0) * 1e3 / args.steps
    ms_proj = e0.elapsed_time(e1) / args.steps
    maxdiv = float(bits.cpu().numpy().view(np.float32)[0])
    log(f"jacobi projection: {ms_proj:.3f} ms (host wall {wall:.3f} ms), max|div| {maxdiv:.4e}")

    # Dominant kernel: the mask-free sweep.  Time a 200-sweep solve alone.
    gdiv = div.clone()
    reps = 5
    core_solve = lambda: sim_core._lib.call(
        "bs_jacobi_solve_launch", p.data_ptr(), p_tmp.data_ptr(), gdiv.data_ptr(), gmat.data_ptr(),
        nx, ny, nz, cfg.dx, DT, RHO, args.sweeps, 0, 2, 1, 0)
    core_solve()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        core_solve()
    e1.record()
    torch.cuda.synchronize()
    ms_sweep = e0.elapsed_time(e1) / reps / args.sweeps
    bytes_sweep = 12.0 * n
    achieved = bytes_sweep / (ms_sweep * 1e-3) / 1e9
    log(f"sweep: {ms_sweep * 1e3:.1f} us, {achieved:.0f} GB/s algorithmic (12 B/cell)")

    cpu = None
    if not args.no_cpu:
        threads = args.cpu_threads or len(os.sched_getaffinity(0))
        threads = min(threads, 16)
        sample = 10 if cfg.name == "prod" else 50
        mat_np, div_np = mat, div.cpu().numpy()
        ms_cpu_sweep = cpu_baseline(cfg, mat_np, div_np, sample, threads)
        cpu = {"value": ms_cpu_sweep * args.sweeps, "unit": "ms/projection (200 Jacobi sweeps)",
               "cores": threads, "kind": "port",
               "sample": f"{sample} oracle Jacobi sweeps at {nx}x{ny}x{nz}, scaled to {args.sweeps}"}
        log(f"cpu baseline: {ms_cpu_sweep:.2f} ms/sweep on {threads} threads")

    out = {
        "metric": "ms/projection to rel-res 1e-6 at 440x440x280: learned-PCG vs plain PCG",
        "value": ms_proj,
        "unit": "ms/projection",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_proj,
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded pot mask and N(0,0.01) face velocities + plume)",
        "config": {"workload": f"jacobi projection, {args.sweeps} sweeps, {nx}x{ny}x{nz} pot",
                   "grid": [nx, ny, nz], "dx_m": cfg.dx, "sweeps": args.sweeps},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "jacobi_sweep_kernel<4,encoded,encoded>", "bytes_per_cell": 12},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
