"""Pressure-projection benchmark on one MI355X.

Workload (the headline metric): one pressure projection on the 440x440x280
pot (dx = 0.5 mm, ~54.2 M cells, FP32 fields), RHS the divergence of seeded
N(0, 0.01) face velocities (seed 11, no plume), solved to a TRUE relative
residual ||b - A x||_2/||b||_2 <= 1e-6 (FP64 check), inputs resident in HBM:
  * learned-PCG: flexible CG preconditioned by a damped-Jacobi V-cycle whose
    4x-coarsened grid is solved by the trained U-Net, its correction polished
    by Chebyshev steps  -> `value`
  * plain PCG: Jacobi-preconditioned CG, FP32 vectors / FP64 scalars
  * (ablation) the same V-cycle with one Jacobi sweep from zero instead of
    the network, followed by the same Chebyshev polish
ms/projection is the whole projection: divergence -> solve -> gradient
subtraction -> divergence -> on-device max|div|.  Each step restores the
input velocities (untimed) so every projection solves the same system.

Also measured live: the roofline of the mask-free Jacobi sweep and of the
PCG iteration kernels (HIP events on the launch stream), the first frames of
the warm-started sequence of correlated smooth right-hand sides (AR(1)
Gaussian random fields plus a drifting plume, FP64 refinement around the
learned flexible CG vs around Jacobi-PCG), and the CPU baseline (the
oracle's PCG on a bounded sample on the host cores).

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

# Algorithmic bytes per cell (FP32 fields, one-byte class code where used).
B_JACOBI = 12.0   # p_in, div, p_out (class carried in p, no mask read)
B_PCG_ITER = 47.0  # spmv 4+1+4, update 4+4+(4+4)+(4+4)+1, direction 4+1+(4+4)
# Fine-level algorithmic bytes per cell of one learned flexible-CG iteration:
# direction + spmv + deferred x step 25 (code, z, p_old, x; p, Ap, x),
# residual update + pre-smoothing + residual + restriction 17.5 (code, r_old,
# Ap; r, z, rc/8 x 4), prolongation + post-sweep + r.z 13.5 (code, z_in,
# coarse e/8 x 4, r; z).
B_LEARNED_FINE = 56.0
# HBM bytes per cell one two-sweep launch of jacobi_pair_kernel actually
# moves at 440x440x280, measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE (one counter per run) and calibrated by kernels of known traffic
# in the same runs (a one-field streaming read counts 0.5 of its bytes on
# FETCH_SIZE; a one-field fill writes exactly its bytes on WRITE_SIZE).
# 266,018.7 KiB x 1.99976 read + 220,094.4 KiB written per launch.
TRAFFIC_JACOBI_PAIR = 14.207


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
    ap.add_argument("--cpu-iters", type=int, default=12, help="oracle PCG iterations in the CPU sample")
    ap.add_argument("--legs", default="learned,pcg,ablation,jacobi,mixed,thermal,faces,advection,sequence,cpu",
                    help="comma list of the legs to run (the JSON fields of skipped legs are null)")
    ap.add_argument("--seq-frames", type=int, default=6,
                    help="frames of the warm-started AR(1) sequence leg (0: skip; the full run is "
                         "python -m boilsim.sequence --frames 100)")
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


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU only; run with --gpus 1", "n_gpus": args.gpus}))
        sys.exit(2)
    import torch

    import sim_core
    from boilsim.fixtures import CONFIGS, DT, RHO, pot_mask, seeded_velocities
    from boilsim.learned import PRECOND_MG_JACOBI, PRECOND_MG_UNET
    from boilsim.unet import UNet

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
    net = UNet.load()
    ws_pcg = torch.empty(core.pcg_workspace_bytes(nx, ny, nz), dtype=torch.uint8, device="cuda")
    ws_f = {k: torch.empty(core.fcg_workspace_bytes(nx, ny, nz, k), dtype=torch.uint8, device="cuda")
            for k in (PRECOND_MG_UNET, PRECOND_MG_JACOBI)}
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
            k = PRECOND_MG_UNET if solver == "learned" else PRECOND_MG_JACOBI
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
    abl = None
    if "ablation" in legs:
        ms_a, _, st_a, _ = leg("mg-jacobi")
        abl = {"ms_per_projection": ms_a, "iterations": st_a["iterations"], "relres_true": st_a["relres_true"]}

    # Roofline of the Jacobi sweep: 200 mask-free sweeps timed with HIP events
    # on the null stream, the stream the launcher uses.  A plain device copy of
    # one field (n floats) runs first: its FETCH_SIZE / WRITE_SIZE counters
    # calibrate the traffic measured for the sweep kernel.
    us_sweep = us_pair = jac_gbs = nan
    gdiv = torch.empty_like(p)  # the initial divergence (the solves' RHS)
    core.compute_divergence(ptr(gdiv), ptr(ux0), ptr(uy0), ptr(uz0), ptr(gmat), nx, ny, nz, cfg.dx)
    if "jacobi" in legs:
        cal = torch.empty_like(p)
        cal.copy_(gdiv)  # n floats read, n floats written
        torch.cuda.synchronize()
        us_pair, us_sweep = jacobi_roofline(core, sim_core, torch, p, gdiv, ux0, uy0, uz0, gmat, cfg, nx, ny, nz, n)
        jac_gbs = 2 * B_JACOBI * n / (us_pair * 1e-6) / 1e9
    us_iter = ms_p * 1e3 / max(st_p["iterations"], 1) if st_p else nan
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

        log(f"AR(1) sequence: {args.seq_frames} frames, learned vs Jacobi-PCG (FP64 refinement, warm starts)")
        res = run_sequence(core, torch, gmat, args.seq_frames, methods=("learned", "pcg-fp32", "pcg-fp16"),
                           net=net, tol=args.tol)
        seq = summarize(res)
        for m, v in seq.items():
            log(f"  {m}: median {v['ms_per_solve_median']:.1f} ms/solve, {v['iterations_mean']:.1f} its, "
                f"max relres {v['max_relres_true']:.2e}, err vs FP64 {v['max_err_vs_fp64']:.2e}")
        seq["speedup_learned_vs_pcg_fp32"] = (seq["pcg-fp32"]["ms_per_solve_median"] /
                                              seq["learned"]["ms_per_solve_median"])
        seq["rhs"] = "AR(1) GRF (step corr 0.95, ell 12 cells, unit variance) + drifting plume; warm starts"

    cpu = None
    if "cpu" in legs and st_p:
        import oracle

        threads = min(len(os.sched_getaffinity(0)), 16)
        oracle.set_threads(threads)
        div_np = gdiv.cpu().numpy()
        t0 = time.perf_counter()
        _, it_c, _ = oracle.pcg_ref(div_np, mat, cfg.dx, DT, RHO, 1e-30, args.cpu_iters)
        dt_c = time.perf_counter() - t0
        ms_per_it = dt_c * 1e3 / max(it_c, 1)
        cpu = {"value": ms_per_it * st_p["iterations"], "unit": "ms/projection (plain PCG solve, same iterations)",
               "cores": threads, "kind": "port",
               "sample": f"{it_c} iterations of the oracle's FP32/FP64 Jacobi-PCG at {nx}x{ny}x{nz} on {threads} "
                         f"threads ({dt_c:.1f} s incl. set-up), scaled to the GPU plain-PCG count "
                         f"{st_p['iterations']}"}
        log(f"cpu baseline: {ms_per_it:.1f} ms/iteration on {threads} threads")
        # the same on one thread (a serial loop over cells), a shorter sample
        oracle.set_threads(1)
        t0 = time.perf_counter()
        _, it_1, _ = oracle.pcg_ref(div_np, mat, cfg.dx, DT, RHO, 1e-30, 3)
        ms_1 = (time.perf_counter() - t0) * 1e3 / max(it_1, 1)
        oracle.set_threads(threads)
        cpu["single_thread"] = {"ms_per_iteration": ms_1, "iterations_sampled": it_1,
                                "ms_per_projection": ms_1 * st_p["iterations"]}
        log(f"cpu baseline: {ms_1:.1f} ms/iteration on 1 thread")
        # parity of the shipped Jacobi path with the oracle on config 1
        # (64x64x48 pot, seed-7 velocities with the plume, 200 sweeps from 0)
        from boilsim.fixtures import CONFIGS as _C

        c1 = _C["small"]
        m1 = pot_mask(*c1.shape)
        v1 = seeded_velocities(*c1.shape, seed=7, plume=True)
        d1 = oracle.divergence(*v1, m1, c1.dx)
        pr, tr = np.zeros(c1.shape, np.float32), np.zeros(c1.shape, np.float32)
        oracle.pressure_solve(pr, tr, d1, m1, c1.dx, DT, RHO, 200)
        pg, tg = torch.zeros(c1.shape, device="cuda"), torch.zeros(c1.shape, device="cuda")
        dg, mg = d(d1), d(m1)
        core.pressure_solve(ptr(pg), ptr(tg), ptr(dg), ptr(mg), *c1.shape, c1.dx, DT, RHO, 200)
        torch.cuda.synchronize()
        got = pg.cpu().numpy()
        fl = m1 == 0
        rel = np.abs(got - pr)[fl] / np.maximum(np.abs(pr[fl]), 1e-6)
        cpu["parity"] = {"case": "64x64x48 pot, seed-7 velocities + plume, 200 Jacobi sweeps from 0",
                         "bits_equal": bool(np.array_equal(got.view(np.uint32), pr.view(np.uint32))),
                         "max_rel_fluid": float(rel.max()), "gate": 1e-4}
        log(f"parity (Jacobi, 200 sweeps, 64x64x48): bits equal {cpu['parity']['bits_equal']}, "
            f"max rel {cpu['parity']['max_rel_fluid']:.1e}")

    fin = lambda v: v if v == v else None  # NaN (leg skipped) -> null
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
        "ablation_mg_jacobi_coarse": abl,
        "max_div_before": maxdiv0,
        "roofline": {"bound": "hbm", "achieved": fin(jac_gbs), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": fin(jac_gbs / HBM_PEAK_GBS),
                     "traffic": TRAFFIC_JACOBI_PAIR * n if (cfg.name == "prod" and TRAFFIC_JACOBI_PAIR) else None,
                     "traffic_unit": "bytes/launch (PMC, calibrated)",
                     "kernel": "jacobi_pair_kernel (two mask-free sweeps per launch)",
                     "bytes_per_cell": 2 * B_JACOBI, "us_per_launch": fin(us_pair),
                     "us_per_sweep_200_sweep_driver": fin(us_sweep)},
        "roofline_pcg_iteration": {"bound": "hbm", "achieved": fin(pcg_gbs), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                   "frac": fin(pcg_gbs / HBM_PEAK_GBS), "bytes_per_cell": B_PCG_ITER,
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
