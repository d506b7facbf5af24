"""Pressure-projection benchmark on one MI355X.

Workload (the headline metric): one pressure projection on the 440x440x280
pot (dx = 0.5 mm, ~54.2 M cells, FP32 fields), RHS the divergence of seeded
N(0, 0.01) face velocities (seed 11, no plume), solved to a TRUE relative
residual ||b - A x||_2/||b||_2 <= 1e-6 (FP64 check), inputs resident in HBM:
  * learned-PCG: flexible CG preconditioned by a V-cycle whose 4x-coarsened
    grid is solved by the trained U-Net, its correction polished by
    Chebyshev steps  -> `value`
  * plain PCG: Jacobi-preconditioned CG, FP32 vectors / FP64 scalars
  * (ablations) the same V-cycle with non-learned coarse solves
ms/projection is the whole projection: divergence -> solve -> gradient
subtraction -> divergence -> on-device max|div|.  Each step restores the
input velocities (untimed) so every projection solves the same system.

Also measured live, with HIP events on the stream each kernel runs on:
`roofline` for the learned solve's dominant kernel (fcg_spmv_dir_kernel) and
its two other fine-level passes (timed inside learned solves run with pass
timing on), the Jacobi sweep's two-sweep kernel, the PCG iteration, the
200-sweep Jacobi projection of configs 1-2 at all three grids, conduction,
the face passes and advection; the warm-started sequence of correlated
smooth right-hand sides (AR(1) Gaussian random fields plus a drifting plume,
FP64 refinement around the learned flexible CG vs around Jacobi-PCG); and
the CPU baseline: the oracle's plain PCG projection at 440x440x280, solved in
full on the host cores, and its 200-sweep Jacobi projections.

Prints ONE JSON line on stdout; progress goes to stderr.
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
METRIC = "ms/projection to rel-res 1e-6 at 440x440x280: learned-PCG vs plain PCG"

# Algorithmic bytes per cell: FP32 fields 4 B; m is the class bytes per cell
# a kernel reads (the reference's int32 mask: m = 4; this build's one-byte
# class code: m = 1).
# Jacobi sweep, 12 + m: p_in, div read, p_out written.  Inside the fused
# solve the class rides in p itself (NaN sentinels), so the middle sweeps
# read no mask at all (m = 0, 12 B/cell); B_JACOBI_REF is the reference's
# layout (m = 4).
B_JACOBI = 12.0
B_JACOBI_REF = 16.0
# PCG iteration, fused-minimum convention 40 + 2m with m = 1.
B_PCG_ITER = 42.0
# The learned flexible-CG iteration's three fine-level passes:
#   fcg_spmv_dir_kernel: code, z, p_old, x read; p, Ap, x written = 25
#   pre2_restrict_kernel<true, 64>: code, r_old, Ap read; r, z written, the
#       restricted residual (1/8 of a field) written = 17.5
#   prolong_post_kernel<48>: code, z_in, r, the coarse correction (1/8) read;
#       z written = 13.5
PASS_BYTES = {"fcg_spmv_dir_kernel": 25.0, "pre2_restrict_kernel<true, 64>": 17.5, "prolong_post_kernel<48>": 13.5}
B_LEARNED_FINE = sum(PASS_BYTES.values())  # 56
# HBM bytes per cell one launch actually moves at 440x440x280, measured with
# rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (one counter per run) and
# calibrated by kernels of known traffic in the same runs (a one-field
# streaming read counts 0.5 of its bytes on FETCH_SIZE; a one-field fill
# writes exactly its bytes on WRITE_SIZE).  jacobi_pair_kernel: 266,018.7 KiB
# x 1.99976 read + 220,094.4 KiB written per launch.
TRAFFIC_JACOBI_PAIR = 14.207
# The learned iteration's fine-level passes, measured the same way over
# `bench.py --legs passes` (reads = FETCH_SIZE x 1.99965 from max_abs_kernel,
# writes = WRITE_SIZE; means over the 45 launches that ran): read + written
# KiB per launch 709,727 + 647,096, 517,636 + 450,094 and 602,806 + 216,228.
TRAFFIC_PASSES = {"fcg_spmv_dir_kernel": 25.63, "pre2_restrict_kernel<true, 64>": 18.28,
                  "prolong_post_kernel<48>": 15.47}


# The same flexible CG and V-cycle with non-learned solves of the 4x-coarsened
# level (preconditioner kinds of bs_fcg_solve).
ABLATIONS = {"mg-jacobi": 3, "mg-chebyshev60": 5, "mg-deep": 6}
ABLATION_DOC = {
    "mg-jacobi": "one Jacobi sweep + the 20 Chebyshev polish steps (less work than the network)",
    "mg-chebyshev60": "one Jacobi sweep + 60 Chebyshev steps over [0.001, 2] (the best non-learned coarse solve "
                      "measured; more work than the network + its 20-step polish)",
    "mg-deep": "two more geometric levels (55x55x35, 28x28x18) with Chebyshev smoothing and a 24-iteration "
               "Jacobi-PCG bottom solve, + the 20 polish steps",
}


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", default="prod", choices=["small", "dev", "prod"])
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-ablation", action="store_true")
    ap.add_argument("--legs", default="learned,passes,pcg,ablation,jacobi,jacobi_configs,mixed,thermal,faces,"
                                      "advection,sequence,cpu",
                    help="comma list of the legs to run (the JSON fields of skipped legs are null)")
    ap.add_argument("--seq-frames", type=int, default=100,
                    help="frames of the warm-started AR(1) sequence leg (0: skip)")
    ap.add_argument("--seq-methods", default="learned,mg,pcg-fp32",
                    help="methods of the sequence leg (learned, mg, pcg-fp32, pcg-fp16)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="host threads of the CPU baseline (0: every core in the process's affinity)")
    return ap.parse_args()


def jacobi_roofline(core, sim_core, torch, p, gdiv, ux0, uy0, uz0, gmat, cfg, nx, ny, nz, n):
    """HIP-event timings on the null stream (the launcher's stream) of the
    fused Jacobi driver: 200 sweeps from p = 0 run the first sweep (mask read),
    98 two-sweep passes of `jacobi_pair_kernel`, three single mask-free sweeps
    and a decode; 4 sweeps run the same single sweeps without the pairs, so
    (T200 - T4) / 98 is the average pair launch.  Returns (us per pair launch,
    us per sweep over the whole 200-sweep driver)."""
    from boilsim.fixtures import DT, RHO

    ptr = lambda t: t.data_ptr()
    core.compute_divergence(ptr(gdiv), ptr(ux0), ptr(uy0), ptr(uz0), ptr(gmat), nx, ny, nz, cfg.dx)
    p_tmp = torch.zeros_like(p)
    solve = lambda k: sim_core._lib.call("bs_jacobi_solve_launch", ptr(p), ptr(p_tmp), ptr(gdiv), ptr(gmat), nx, ny,
                                         nz, cfg.dx, DT, RHO, k, 0, 2, 1, 0)

    def timed(k, reps):
        solve(k)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            solve(k)
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e3  # us per call

    t200 = timed(200, 3)
    t4 = timed(4, 10)
    us_pair = (t200 - t4) / 98
    us_sweep = t200 / 200
    log(f"jacobi: {us_pair:.1f} us per two-sweep launch ({2 * B_JACOBI * n / (us_pair * 1e-6) / 1e9:.0f} GB/s "
        f"algorithmic), {us_sweep:.1f} us per sweep over the 200-sweep driver")
    return us_pair, us_sweep


def conduction_leg(core, torch, gmat, mat, cfg, nx, ny, nz, n):
    """Backward-Euler Jacobi conduction (thermal.py:259-388, 631-665) on the
    pot: one step of the reference's default 15 sweeps (dt 0.05 s, hot pot
    base), and the sweep kernel alone, 20 launches timed with HIP events on
    the null stream.  Algorithmic bytes per sweep: T_iter, T_old, mat read,
    T_dst written = 16 B/cell."""
    from boilsim.thermal import MaterialProps, allocate_thermal_workspace, conduct_implicit

    props = MaterialProps.from_pot("steel_304")
    k, rho, cp = props.device(torch)
    t0 = np.full(cfg.shape, 295.15, np.float32)
    base = (mat == 1) & (np.arange(nz)[None, None, :] < nz // 6)
    t0[base] = 500.0
    T = torch.from_numpy(t0).cuda()
    ws = allocate_thermal_workspace(torch, cfg.shape)
    dt = 0.05
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    conduct_implicit(core, T, gmat, (k, rho, cp), ws, cfg.dx, dt, 15)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(3):
        conduct_implicit(core, T, gmat, (k, rho, cp), ws, cfg.dx, dt, 15)
    e1.record()
    torch.cuda.synchronize()
    ms_step = e0.elapsed_time(e1) / 3
    src, dst = T, ws.T_iter_alt
    ptr = lambda t: t.data_ptr()
    sweep = lambda: core.conduction_sweep(ptr(dst), ptr(src), ptr(ws.T_old_snapshot), ptr(gmat), ptr(k), ptr(rho),
                                          ptr(cp), k.numel(), nx, ny, nz, cfg.dx, dt, 10.0, 295.15)
    sweep()
    torch.cuda.synchronize()
    e0.record()
    for _ in range(20):
        sweep()
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) / 20 * 1e3
    gbs = 16.0 * n / (us * 1e-6) / 1e9
    log(f"conduction: {us:.1f} us/sweep ({gbs:.0f} GB/s algorithmic, 16 B/cell), {ms_step:.2f} ms per 15-sweep step")
    return {"us_per_sweep": us, "achieved_gbs": gbs, "frac_of_peak": gbs / HBM_PEAK_GBS, "bytes_per_cell": 16.0,
            "ms_per_step_15_sweeps": ms_step, "kernel": "conduction_sweep_kernel<4>"}


def faces_leg(core, torch, gmat, ux0, uy0, uz0, gux, guy, guz, div, cfg, nx, ny, nz, n, reps=10):
    """The face passes around the projection (pipeline.py:606-620,
    fluid.py:111-162, 726-734, 796-833), unfused against fused, each variant
    timed with HIP events on the null stream over `reps` runs.
      pre:  enforce_no_slip + compute_divergence  vs  divergence_no_slip,
            on the seed-11 faces (before any no-slip)
      post: compute_divergence + max|div| + enforce_no_slip +
            compute_max_velocity  vs  divergence_max, on no-slipped faces
    The gradient subtraction between them is the same in both and not timed.
    Algorithmic bytes of each fused pass: mat and the three face arrays read,
    div written = 20 B/cell; the pre pass also stores every face it zeroes
    (on these faces, all the faces that are not between two fluid cells:
    4 B x that count per cell)."""
    ptr = lambda t: t.data_ptr()
    F = (ptr(gux), ptr(guy), ptr(guz))
    mb = torch.zeros(2, dtype=torch.int32, device="cuda")
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def reset(no_slip):
        gux.copy_(ux0)
        guy.copy_(uy0)
        guz.copy_(uz0)
        if no_slip:
            core.enforce_no_slip(*F, ptr(gmat), nx, ny, nz)
        torch.cuda.synchronize()

    def timed(fn, no_slip):
        ts = []
        for _ in range(reps + 1):
            reset(no_slip)
            e0.record()
            fn()
            e1.record()
            torch.cuda.synchronize()
            ts.append(e0.elapsed_time(e1) * 1e3)
        return float(np.mean(ts[1:]))

    div_ = lambda: core.compute_divergence(ptr(div), *F, ptr(gmat), nx, ny, nz, cfg.dx)
    ns = lambda: core.enforce_no_slip(*F, ptr(gmat), nx, ny, nz)
    res = {
        "pre_unfused_us": timed(lambda: (ns(), div_()), False),
        "pre_fused_us": timed(lambda: core.divergence_no_slip(ptr(div), *F, ptr(gmat), nx, ny, nz, cfg.dx), False),
        "post_unfused_us": timed(lambda: (div_(), core.max_abs(ptr(div), n, ptr(mb)), ns(),
                                          core.max_velocity(*F, nx, ny, nz, ptr(mb) + 4)), True),
        "post_fused_us": timed(lambda: core.divergence_max(ptr(div), *F, ptr(gmat), nx, ny, nz, cfg.dx, ptr(mb)),
                               True),
        "enforce_no_slip_us": timed(ns, False),
        "max_velocity_us": timed(lambda: core.max_velocity(*F, nx, ny, nz, ptr(mb)), True),
        "divergence_us": timed(div_, True),
    }
    fl = torch.nn.functional.pad((gmat == 0).to(torch.uint8), (1, 1, 1, 1, 1, 1)).bool()
    kept = int((fl[:-1, 1:-1, 1:-1] & fl[1:, 1:-1, 1:-1]).sum() + (fl[1:-1, :-1, 1:-1] & fl[1:-1, 1:, 1:-1]).sum() +
               (fl[1:-1, 1:-1, :-1] & fl[1:-1, 1:-1, 1:]).sum())
    del fl
    zeroed = ((nx + 1) * ny * nz + nx * (ny + 1) * nz + nx * ny * (nz + 1) - kept) / n
    res["zeroed_faces_per_cell"] = zeroed
    res["bytes_per_cell_pre_fused"] = 20.0 + 4.0 * zeroed
    res["pre_fused_gbs"] = res["bytes_per_cell_pre_fused"] * n / (res["pre_fused_us"] * 1e-6) / 1e9
    res["post_fused_gbs"] = 20.0 * n / (res["post_fused_us"] * 1e-6) / 1e9
    res["saved_us_per_step"] = (res["pre_unfused_us"] + res["post_unfused_us"] - res["pre_fused_us"] -
                                res["post_fused_us"])
    res["bytes_per_cell_post_fused"] = 20.0
    log(f"faces: pre {res['pre_unfused_us']:.1f} -> {res['pre_fused_us']:.1f} us "
        f"({res['pre_fused_gbs']:.0f} GB/s), post {res['post_unfused_us']:.1f} -> {res['post_fused_us']:.1f} us "
        f"({res['post_fused_gbs']:.0f} GB/s); enforce_no_slip {res['enforce_no_slip_us']:.1f} us, "
        f"max_velocity {res['max_velocity_us']:.1f} us, divergence {res['divergence_us']:.1f} us")
    return res


def smooth_faces(torch, faces, umax=0.05, passes=3, width=5):
    """A smooth velocity field for the advection timing: each face array
    box-filtered `passes` times (width `width` cells, zero-padded) and scaled
    to max|u| = umax m/s, which is how a projected flow looks to the sampler
    (neighbouring cells' departure points in the same cells)."""
    out = []
    for a in faces:
        x = a[None, None]
        for _ in range(passes):
            x = torch.nn.functional.avg_pool3d(x, width, stride=1, padding=width // 2)
        out.append(x[0, 0].contiguous())
    m = max(float(a.abs().max()) for a in out)
    return [a * (umax / m) for a in out]


def advection_leg(core, torch, gmat, mat, ux0, uy0, uz0, cfg, nx, ny, nz, n, reps=10):
    """Semi-Lagrangian advection (fluid.py:737-775) of a temperature field with
    a hot pot base and of the velocity itself, with the geometry's origin and
    the CFL-limited dt (0.4 dx / max|u|), each pass timed with HIP events on
    the null stream over `reps` launches, for two velocity fields:
      smooth: the seed-11 faces box-filtered (3 x 5 cells) to max|u| 5 cm/s,
              what the sampler sees in a projected flow;
      random: the seed-11 faces themselves (independent per face), the worst
              case: neighbouring cells' departure points scatter over
              different rows, so each gather touches ~12 L1 lines.
    Algorithmic bytes of the advection pass: ux, uy, uz, T_ext, T and mat
    read once, four outputs written = 40 B/cell (the 32 gathers per output
    hit L1/L2)."""
    ptr = lambda t: t.data_ptr()
    t0 = np.full(cfg.shape, 295.15, np.float32)
    t0[(mat == 1) & (np.arange(nz)[None, None, :] < nz // 6)] = 480.0
    T = torch.from_numpy(t0).cuda()
    outs = [torch.empty_like(a) for a in (ux0, uy0, uz0, T)]
    t_ext = torch.empty_like(T)
    origin = (-nx * cfg.dx / 2, -ny * cfg.dx / 2, -4 * cfg.dx)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e3

    res = {}
    us_ext = timed(lambda: core.extend_temperature(ptr(t_ext), ptr(T), ptr(gmat), nx, ny, nz))
    for name, (u, v, w) in (("smooth", smooth_faces(torch, (ux0, uy0, uz0))), ("random", (ux0, uy0, uz0))):
        umax = max(float(a.abs().max()) for a in (u, v, w))
        dt = 0.4 * cfg.dx / umax
        us = timed(lambda: core.advect(*(ptr(o) for o in outs), ptr(u), ptr(v), ptr(w), ptr(T), ptr(t_ext),
                                       ptr(gmat), nx, ny, nz, origin, cfg.dx, dt))
        gbs = 40.0 * n / (us * 1e-6) / 1e9
        res[name] = {"us_advect": us, "achieved_gbs": gbs, "frac_of_peak": gbs / HBM_PEAK_GBS, "dt_s": dt,
                     "max_u": umax}
        log(f"advection ({name} velocity): {us:.1f} us per advect pass ({gbs:.0f} GB/s algorithmic, 40 B/cell), "
            f"dt {dt * 1e3:.3f} ms (CFL 0.4)")
    log(f"temperature extension: {us_ext:.1f} us ({12.0 * n / (us_ext * 1e-6) / 1e9:.0f} GB/s, 12 B/cell)")
    res.update({"us_extend": us_ext, "bytes_per_cell": 40.0,
                "kernel": "advect_kernel (u_x, u_y, u_z and T in one pass)"})
    return res


def pass_leg(core, torch, projection, cal_src, n, reps=3):
    """The learned solve's fine-level passes timed inside `reps` learned
    projections run with pass timing on (bs_fcg_pass_timing: the solves
    launch eagerly and bracket fcg_spmv_dir_kernel, pre2_restrict_kernel<true, 64>
    and prolong_post_kernel with HIP events on the solve's stream).  A plain
    device copy of one field runs first: its FETCH_SIZE / WRITE_SIZE counters
    calibrate a counter run over this leg (bench.py --legs passes)."""
    cal = torch.empty_like(cal_src)
    cal.copy_(cal_src)  # n floats read, n floats written
    torch.cuda.synchronize()
    core.fcg_pass_timing_start()
    try:
        for _ in range(reps):
            projection("learned")
        torch.cuda.synchronize()
    finally:
        res = core.fcg_pass_timing_stop()
    out = {}
    for k, v in res.items():
        b = PASS_BYTES[k]
        gbs = b * n / (v["us_mean"] * 1e-6) / 1e9 if v["us_mean"] > 0 else float("nan")
        tr = TRAFFIC_PASSES.get(k)
        out[k] = {"us_median": v["us_median"], "us_mean": v["us_mean"], "launches": v["launches"],
                  "bytes_per_cell": b, "achieved_gbs": gbs, "frac": gbs / HBM_PEAK_GBS,
                  "traffic_bytes_per_cell": tr, "traffic_over_algorithmic": (tr / b) if tr else None}
        log(f"pass {k}: {v['us_mean']:.1f} us mean ({v['us_median']:.1f} median, {v['launches']} launches), "
            f"{gbs:.0f} GB/s algorithmic at {b} B/cell ({gbs / HBM_PEAK_GBS:.2f} of peak)")
    return out


def jacobi_configs_leg(core, torch, reps=10, warm=3):
    """Configs 1-2: ms per 200-sweep Jacobi projection (bs_projection_jacobi:
    zero p, divergence, 200 sweeps, gradient, divergence + max|div|) of the
    seed-7 velocities with the plume, at 64x64x48, 110x110x70 and
    440x440x280, and at 110x110x68 (the grid of the reference's published
    3.86 / 5.86 ms on an RTX 6000 Ada, analytic pot mask), HIP events on the
    null stream (the launcher's), median of `reps` after `warm`."""
    from boilsim.fixtures import CONFIGS, DT, RHO, pot_mask, seeded_velocities

    out = {}
    grids = [(CONFIGS[k].shape, CONFIGS[k].dx) for k in ("small", "dev", "prod")] + [((110, 110, 68), 2e-3)]
    for shape, dx in grids:
        mat = pot_mask(*shape)
        u0 = [torch.from_numpy(a).cuda() for a in seeded_velocities(*shape, seed=7, plume=True)]
        u = [a.clone() for a in u0]
        gm = torch.from_numpy(mat).cuda()
        p, pt, dv = (torch.empty(shape, dtype=torch.float32, device="cuda") for _ in range(3))
        mb = torch.zeros(2, dtype=torch.int32, device="cuda")
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        ts = []
        for r in range(warm + reps):
            for a, b in zip(u, u0):
                a.copy_(b)
            torch.cuda.synchronize()
            e0.record()
            core.projection_jacobi(p.data_ptr(), pt.data_ptr(), dv.data_ptr(), *(a.data_ptr() for a in u),
                                   gm.data_ptr(), *shape, dx, DT, RHO, 200, mb.data_ptr())
            e1.record()
            torch.cuda.synchronize()
            if r >= warm:
                ts.append(e0.elapsed_time(e1))
        key = "x".join(map(str, shape))
        out[key] = float(np.median(ts))
        log(f"jacobi projection {key}, 200 sweeps: {out[key]:.3f} ms (median of {reps})")
        del u0, u, gm, p, pt, dv
    out["reference_published_ms_110x110x68"] = {"hand_written_cuda": 3.86, "warp": 5.86,
                                                "hardware": "RTX 6000 Ada (960 GB/s)"}
    return out


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_leg(core, torch, mat, vel11, cfg, max_threads, d):
    """The CPU baseline: the oracle (this repository's C restatement of the
    reference arithmetic, -O3 -ffp-contract=off, OpenMP) on the host cores,
    reported beside the GPU, never a target.
      * one full plain-PCG projection at 440x440x280 (divergence, Jacobi-PCG
        to a true relative residual 1e-6 with residual replacement -- the
        semantics of the GPU's plain PCG, bs_pcg_solve --, gradient,
        divergence + max|div|), timed whole;
      * the 200-sweep Jacobi projection of configs 1-2 at the three grids;
      * one thread: 3 PCG iterations at 440 and the 64x64x48 projection;
      * the shipped Jacobi path against the oracle on config 1 (bit parity)."""
    import oracle
    from boilsim.fixtures import CONFIGS, DT, RHO, pot_mask, seeded_velocities

    aff = len(os.sched_getaffinity(0))
    threads = aff if max_threads <= 0 else max(1, min(aff, max_threads))
    oracle.set_threads(threads)
    model = cpu_model()
    ux, uy, uz = (a.copy() for a in vel11)
    t0 = time.perf_counter()
    div = oracle.divergence(ux, uy, uz, mat, cfg.dx)
    x, it, restarts, rel = oracle.pcg_true(div, mat, cfg.dx, DT, RHO, 1e-6, 20000)
    oracle.gradient(ux, uy, uz, x, mat, cfg.dx, DT, RHO)
    div2 = oracle.divergence(ux, uy, uz, mat, cfg.dx)
    maxdiv = float(np.abs(div2).max())
    t_full = time.perf_counter() - t0
    log(f"cpu: full plain-PCG projection at {cfg.shape}: {t_full:.1f} s, {it} iterations, true relres {rel:.3e}, "
        f"{threads} threads ({model})")
    del ux, uy, uz, div, div2, x
    jac = {}
    for name in ("small", "dev", "prod"):
        c = CONFIGS[name]
        m = pot_mask(*c.shape)
        v = seeded_velocities(*c.shape, seed=7, plume=True)
        t0 = time.perf_counter()
        oracle.projection(*v, m, c.dx, DT, RHO, 200)
        jac["x".join(map(str, c.shape))] = (time.perf_counter() - t0) * 1e3
        log(f"cpu: 200-sweep Jacobi projection {c.shape}: {jac['x'.join(map(str, c.shape))]:.1f} ms")
        del v, m
    # one thread
    oracle.set_threads(1)
    div_np = oracle.divergence(*vel11, mat, cfg.dx)
    t0 = time.perf_counter()
    _, it1, _ = oracle.pcg_ref(div_np, mat, cfg.dx, DT, RHO, 1e-30, 3)
    ms_it1 = (time.perf_counter() - t0) * 1e3 / max(it1, 1)
    c1 = CONFIGS["small"]
    m1 = pot_mask(*c1.shape)
    v1 = seeded_velocities(*c1.shape, seed=7, plume=True)
    t0 = time.perf_counter()
    oracle.projection(*[a.copy() for a in v1], m1, c1.dx, DT, RHO, 200)
    ms_j1 = (time.perf_counter() - t0) * 1e3
    oracle.set_threads(threads)
    log(f"cpu: one thread: {ms_it1:.0f} ms per PCG iteration at 440 (3 sampled, set-up included), "
        f"64x64x48 Jacobi projection {ms_j1:.1f} ms")
    # parity of the shipped Jacobi path with the oracle on config 1
    d1 = oracle.divergence(*v1, m1, c1.dx)
    pr, tr = np.zeros(c1.shape, np.float32), np.zeros(c1.shape, np.float32)
    oracle.pressure_solve(pr, tr, d1, m1, c1.dx, DT, RHO, 200)
    pg, tg = torch.zeros(c1.shape, device="cuda"), torch.zeros(c1.shape, device="cuda")
    dg, mg = d(d1), d(m1)
    core.pressure_solve(pg.data_ptr(), tg.data_ptr(), dg.data_ptr(), mg.data_ptr(), *c1.shape, c1.dx, DT, RHO, 200)
    torch.cuda.synchronize()
    got = pg.cpu().numpy()
    fl = m1 == 0
    relp = np.abs(got - pr)[fl] / np.maximum(np.abs(pr[fl]), 1e-6)
    parity = {"case": "64x64x48 pot, seed-7 velocities + plume, 200 Jacobi sweeps from 0",
              "bits_equal": bool(np.array_equal(got.view(np.uint32), pr.view(np.uint32))),
              "max_rel_fluid": float(relp.max()), "gate": 1e-4}
    log(f"parity (Jacobi, 200 sweeps, 64x64x48): bits equal {parity['bits_equal']}, max rel {parity['max_rel_fluid']:.1e}")
    return {"value": t_full * 1e3, "unit": "ms/projection", "cores": threads, "kind": "port",
            "sample": f"one full plain-PCG projection at {cfg.shape[0]}x{cfg.shape[1]}x{cfg.shape[2]} (full solve, "
                      f"not a sample): divergence, the oracle's Jacobi-PCG to a true relative residual 1e-6 ({it} "
                      f"iterations, {restarts} residual replacement(s), true relres {rel:.2e}), gradient, "
                      f"divergence + max|div|, on {threads} threads",
            "cpu_model": model, "host_cores_in_affinity": aff, "iterations": it, "relres_true": rel,
            "max_div_after": maxdiv, "jacobi_ms_per_projection_200_sweeps": jac,
            "single_thread": {"ms_per_pcg_iteration_440": ms_it1, "jacobi_projection_64x64x48_ms": ms_j1},
            "parity": parity}


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1", "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    import sim_core
    from boilsim.fixtures import CONFIGS, DT, RHO, pot_mask, seeded_velocities
    from boilsim.learned import PRECOND_MG_UNET, default_net

    if sim_core.cuda is None:
        raise RuntimeError("libboilsim_hip.so not loadable: run __graft_entry__.build() first")
    core = sim_core.cuda.SimCore(0)
    log(core.device_info())
    cfg = CONFIGS[args.config]
    nx, ny, nz = cfg.shape
    n = cfg.n
    log(f"generating the {nx}x{ny}x{nz} pot and seeded velocities (seed 11, no plume)")
    mat = pot_mask(nx, ny, nz)
    ux, uy, uz = seeded_velocities(nx, ny, nz, seed=11, plume=False)
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).cuda()
    gmat = d(mat)
    ux0, uy0, uz0 = d(ux), d(uy), d(uz)
    gux, guy, guz = ux0.clone(), uy0.clone(), uz0.clone()
    p = torch.zeros(cfg.shape, dtype=torch.float32, device="cuda")
    div = torch.zeros_like(p)
    bits = torch.zeros(2, dtype=torch.int32, device="cuda")  # max|div|, max|u| as float bits
    net = default_net()
    ws_pcg = torch.empty(core.pcg_workspace_bytes(nx, ny, nz), dtype=torch.uint8, device="cuda")
    ws_f = {k: torch.empty(core.fcg_workspace_bytes(nx, ny, nz, k), dtype=torch.uint8, device="cuda")
            for k in set(ABLATIONS.values()) | {PRECOND_MG_UNET}}
    torch.cuda.synchronize()
    ptr = lambda t: t.data_ptr()

    es0, es1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)

    def projection(solver):
        core.compute_divergence(ptr(div), ptr(gux), ptr(guy), ptr(guz), ptr(gmat), nx, ny, nz, cfg.dx)
        es0.record()
        if solver == "pcg":
            st = core.pcg_solve(ptr(p), ptr(div), ptr(gmat), nx, ny, nz, cfg.dx, DT, RHO, args.tol, 20000,
                                ptr(ws_pcg), ws_pcg.numel())
        else:
            k = PRECOND_MG_UNET if solver == "learned" else ABLATIONS[solver]
            st = core.fcg_solve(ptr(p), ptr(div), ptr(gmat), nx, ny, nz, cfg.dx, DT, RHO, args.tol, 5000, k,
                                ptr(net.params), ptr(net.packed), ptr(ws_f[k]), ws_f[k].numel())
        es1.record()
        core.subtract_pressure_gradient(ptr(gux), ptr(guy), ptr(guz), ptr(p), ptr(gmat), nx, ny, nz, cfg.dx, DT, RHO)
        # diagnostic divergence with max|div| (and max|u|) in one pass, as pressure_projection
        core.divergence_max(ptr(div), ptr(gux), ptr(guy), ptr(guz), ptr(gmat), nx, ny, nz, cfg.dx, ptr(bits))
        return st

    def leg(solver):
        """`warmup` untimed then `steps` timed projections (HIP events on the
        null stream around the whole projection and around the solve)."""
        times, solve, stats = [], [], None
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for s in range(args.warmup + args.steps):
            gux.copy_(ux0)
            guy.copy_(uy0)
            guz.copy_(uz0)
            torch.cuda.synchronize()
            e0.record()
            stats = projection(solver)
            e1.record()
            torch.cuda.synchronize()
            if s >= args.warmup:
                times.append(e0.elapsed_time(e1))
                solve.append(es0.elapsed_time(es1))
        maxdiv = float(bits.cpu().numpy().view(np.float32)[0])
        ms = float(np.mean(times))
        stats = dict(stats, median_ms=float(np.median(times)), solve_ms_median=float(np.median(solve)))
        log(f"{solver}: {ms:.2f} ms/projection (median {stats['median_ms']:.2f}, solve {stats['solve_ms_median']:.2f}),"
            f" iters {stats['iterations']}, true relres {stats['relres_true']:.3e}, converged {stats['converged']}, "
            f"max|div| after {maxdiv:.3e}")
        return ms, times, stats, maxdiv

    def fresh_projection(solver):
        gux.copy_(ux0)
        guy.copy_(uy0)
        guz.copy_(uz0)
        return projection(solver)

    # Initial divergence (for the report).
    core.compute_divergence(ptr(div), ptr(gux), ptr(guy), ptr(guz), ptr(gmat), nx, ny, nz, cfg.dx)
    core.max_abs(ptr(div), n, ptr(bits))
    torch.cuda.synchronize()
    maxdiv0 = float(bits.cpu().numpy().view(np.float32)[0])

    legs = set(args.legs.split(","))
    if args.no_cpu:
        legs.discard("cpu")
    if args.no_ablation:
        legs.discard("ablation")
    if args.seq_frames <= 0:
        legs.discard("sequence")
    nan = float("nan")
    ms_l, t_l, st_l, md_l = leg("learned") if "learned" in legs else (nan, [], None, None)
    ms_p, t_p, st_p, md_p = leg("pcg") if "pcg" in legs else (nan, [], None, None)
    abl = {}
    if "ablation" in legs:
        for name in ABLATIONS:
            ms_a, _, st_a, _ = leg(name)
            abl[name] = {"ms_per_projection": ms_a, "solve_ms_median": st_a["solve_ms_median"],
                         "iterations": st_a["iterations"], "relres_true": st_a["relres_true"],
                         "converged": st_a["converged"], "fallback": st_a["fallback"],
                         "coarse_solve": ABLATION_DOC[name]}

    gdiv = torch.empty_like(p)  # the initial divergence (the solves' RHS)
    core.compute_divergence(ptr(gdiv), ptr(ux0), ptr(uy0), ptr(uz0), ptr(gmat), nx, ny, nz, cfg.dx)
    passes = pass_leg(core, torch, fresh_projection, gdiv, n) if "passes" in legs else None

    # Roofline of the Jacobi sweep: 200 mask-free sweeps timed with HIP events
    # on the null stream, the stream the launcher uses.  A plain device copy of
    # one field (n floats) runs first: its FETCH_SIZE / WRITE_SIZE counters
    # calibrate the traffic measured for the sweep kernel.
    us_sweep = us_pair = jac_gbs = nan
    if "jacobi" in legs:
        cal = torch.empty_like(p)
        cal.copy_(gdiv)  # n floats read, n floats written
        torch.cuda.synchronize()
        us_pair, us_sweep = jacobi_roofline(core, sim_core, torch, p, gdiv, ux0, uy0, uz0, gmat, cfg, nx, ny, nz, n)
        jac_gbs = 2 * B_JACOBI * n / (us_pair * 1e-6) / 1e9
    jcfg = jacobi_configs_leg(core, torch) if "jacobi_configs" in legs else None
    us_iter = st_p["solve_ms_median"] * 1e3 / max(st_p["iterations"], 1) if st_p else nan
    pcg_gbs = B_PCG_ITER * n / (us_iter * 1e-6) / 1e9 if st_p else nan
    # learned iteration: the solve's median time over its iterations (coarse
    # levels, the network and set-up included, so a lower bound on the fine
    # level's rate)
    us_iter_l = st_l["solve_ms_median"] * 1e3 / max(st_l["iterations"], 1) if st_l else nan
    learned_gbs = B_LEARNED_FINE * n / (us_iter_l * 1e-6) / 1e9 if st_l else nan
    thermal = None
    if "thermal" in legs:
        thermal = conduction_leg(core, torch, gmat, mat, cfg, nx, ny, nz, n)
    adv = None
    if "advection" in legs:
        adv = advection_leg(core, torch, gmat, mat, ux0, uy0, uz0, cfg, nx, ny, nz, n)
    faces = None
    if "faces" in legs:
        faces = faces_leg(core, torch, gmat, ux0, uy0, uz0, gux, guy, guz, div, cfg, nx, ny, nz, n)

    mixed = None
    if "mixed" in legs:
        # Config 3's mixed-precision variant: FP64 refinement around Jacobi-PCG
        # with FP32 / FP16 / BF16 inner vectors, same RHS, true rel-res tol.
        ws_r = torch.empty(core.refine_workspace_bytes(nx, ny, nz), dtype=torch.uint8, device="cuda")
        x64 = torch.zeros(cfg.shape, dtype=torch.float64, device="cuda")
        mixed = {}
        for storage in ("fp32", "fp16", "bf16"):
            for rep in range(2):  # the first call also builds the partial-sum layout caches
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                st = core.refine_solve(ptr(x64), ptr(gdiv), ptr(gmat), nx, ny, nz, cfg.dx, DT, RHO, args.tol,
                                       ptr(ws_r), ws_r.numel(), storage=storage)
                torch.cuda.synchronize()
                ms = (time.perf_counter() - t0) * 1e3
            mixed[storage] = {"ms_per_solve": ms, "iterations": st["inner_iterations"], "steps": st["outer"],
                              "relres_true": st["relres_true"], "converged": st["converged"]}
            log(f"refinement {storage}: {ms:.1f} ms, {st['inner_iterations']} its, {st['outer']} steps, "
                f"true relres {st['relres_true']:.3e}")
        del ws_r, x64

    seq = None
    if "sequence" in legs:
        from boilsim.sequence import run_sequence, summarize

        methods = tuple(args.seq_methods.split(","))
        log(f"AR(1) sequence: {args.seq_frames} frames, {', '.join(methods)} (FP64 refinement, warm starts)")

        def seq_log(t, recs):
            if t % 10 == 9 or t == args.seq_frames - 1:
                log(f"  frame {t}: " + ", ".join(f"{m} {r['ms']:.1f} ms {r['inner_iterations']} its"
                                               for m, r in recs.items()))

        res = run_sequence(core, torch, gmat, args.seq_frames, methods=methods, net=net, tol=args.tol, log=seq_log)
        seq = summarize(res)
        for m, v in seq.items():
            log(f"  {m}: median {v['ms_per_solve_median']:.1f} ms/solve, {v['iterations_mean']:.1f} its, "
                f"max relres {v['max_relres_true']:.2e}, err vs FP64 {v['max_err_vs_fp64']:.2e}")
        if "learned" in seq and "pcg-fp32" in seq:
            seq["speedup_learned_vs_pcg_fp32"] = (seq["pcg-fp32"]["ms_per_solve_median"] /
                                                  seq["learned"]["ms_per_solve_median"])
        if "learned" in seq and "mg" in seq:
            seq["ms_ratio_learned_over_mg"] = (seq["learned"]["ms_per_solve_median"] /
                                               seq["mg"]["ms_per_solve_median"])
        seq["rhs"] = "AR(1) GRF (step corr 0.95, ell 12 cells, unit variance) + drifting plume; warm starts"

    cpu = None
    if "cpu" in legs:
        cpu = cpu_leg(core, torch, mat, (ux, uy, uz), cfg, args.cpu_threads, d)

    fin = lambda v: v if v == v else None  # NaN (leg skipped) -> null
    K0 = "fcg_spmv_dir_kernel"
    out = {
        "metric": METRIC,
        "value": fin(ms_l),
        "unit": "ms/projection",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": fin(ms_l),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic: seeded pot mask, N(0,0.01) face velocities (seed 11), trained U-Net weights",
        "config": {"workload": f"pressure projection, {nx}x{ny}x{nz} pot, true rel-res {args.tol:g}",
                   "grid": [nx, ny, nz], "dx_m": cfg.dx, "rhs": "divergence of N(0,0.01) faces, seed 11",
                   "learned_preconditioner": "FCG + Jacobi V-cycle, U-Net (8/16/32) coarse solve + 20 Chebyshev steps"},
        "plain_pcg": None if not st_p else {
            "ms_per_projection": ms_p, "median_ms": st_p["median_ms"], "solve_ms_median": st_p["solve_ms_median"],
            "iterations": st_p["iterations"], "relres_true": st_p["relres_true"], "max_div_after": md_p},
        "learned_pcg": None if not st_l else {
            "ms_per_projection": ms_l, "median_ms": st_l["median_ms"], "solve_ms_median": st_l["solve_ms_median"],
            "iterations": st_l["iterations"], "relres_true": st_l["relres_true"],
            "fallback": st_l["fallback"], "max_div_after": md_l, "times_ms": t_l},
        "speedup_vs_plain_pcg": fin(ms_p / ms_l),
        "ablations_no_network": abl or None,
        "max_div_before": maxdiv0,
        "roofline": None if not passes else {
            "bound": "hbm", "achieved": passes[K0]["achieved_gbs"], "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": passes[K0]["frac"],
            "traffic": (TRAFFIC_PASSES[K0] * n if (cfg.name == "prod" and TRAFFIC_PASSES.get(K0)) else None),
            "traffic_unit": "bytes/launch (PMC FETCH_SIZE + WRITE_SIZE, calibrated)",
            "kernel": K0 + " (the learned solve's dominant kernel: direction p = z + beta p, SpMV Ap and p.Ap, "
                           "the previous iteration's x += alpha p, one i-marching pass)",
            "bytes_per_cell": PASS_BYTES[K0], "us_per_launch": passes[K0]["us_mean"],
            "us_median": passes[K0]["us_median"], "launches_timed": passes[K0]["launches"]},
        "roofline_fine_passes": passes,
        "roofline_jacobi": {"bound": "hbm", "achieved": fin(jac_gbs), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": fin(jac_gbs / HBM_PEAK_GBS),
                            "frac_at_reference_layout": fin(jac_gbs / HBM_PEAK_GBS * B_JACOBI_REF / B_JACOBI),
                            "traffic": TRAFFIC_JACOBI_PAIR * n if (cfg.name == "prod" and TRAFFIC_JACOBI_PAIR) else None,
                            "traffic_unit": "bytes/launch (PMC, calibrated)",
                            "kernel": "jacobi_pair_kernel (two mask-free sweeps per launch)",
                            "bytes_per_cell": 2 * B_JACOBI,
                            "convention": "12 B/cell per sweep (p_in, div, p_out; the class rides in p as NaN "
                                          "sentinels, m = 0); frac_at_reference_layout counts 12 + m with the "
                                          "reference's int32 mask (m = 4, 16 B/cell)",
                            "us_per_launch": fin(us_pair), "us_per_sweep_200_sweep_driver": fin(us_sweep)},
        "jacobi_ms_per_projection": jcfg,
        "roofline_pcg_iteration": {"bound": "hbm", "achieved": fin(pcg_gbs), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                   "frac": fin(pcg_gbs / HBM_PEAK_GBS), "bytes_per_cell": B_PCG_ITER,
                                   "convention": "fused-minimum 40 + 2m B/cell, m = 1 (one-byte class code)",
                                   "us_per_iteration": fin(us_iter)},
        "roofline_learned_iteration": {"bound": "hbm", "achieved": fin(learned_gbs), "peak": HBM_PEAK_GBS,
                                       "unit": "GB/s", "frac": fin(learned_gbs / HBM_PEAK_GBS),
                                       "bytes_per_cell": B_LEARNED_FINE, "us_per_iteration": fin(us_iter_l),
                                       "note": "fine-level bytes over the whole solve time per iteration "
                                               "(coarse levels, network and set-up included)"},
        "mixed_precision_refinement": mixed,
        "thermal_conduction": thermal,
        "no_slip_faces": faces,
        "advection": adv,
        "sequence_ar1": seq,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
