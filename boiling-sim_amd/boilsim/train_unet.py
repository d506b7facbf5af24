"""Train the learned preconditioner on the GPU with the library's kernels.

    python -m boilsim.train_unet --out weights.npz [--pairs 2000 --epochs 3]

Pairs come from the plain solver on the 110x110x70 pot (dx = 2 mm) with RHS
seeds 0..pairs-1; 200 held-out seeds (10000..10199) measure the loss and the
flexible-CG iteration counts against plain PCG."""

from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np


def log(msg):
    print(f"[train] {msg}", flush=True)


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--pairs", type=int, default=2000)
    ap.add_argument("--heldout", type=int, default=200)
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--eval-solves", type=int, default=20)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args(argv)
    import torch

    import sim_core
    from boilsim.fixtures import CONFIGS, pot_mask
    from boilsim.learned import (PRECOND_UNET, _UNIT, evaluate_loss, fluid_count, generate_pairs, grf_torch,
                                 train)
    from boilsim.unet import UNet, build_codes

    core = sim_core.cuda.SimCore(0)
    cfg = CONFIGS["dev"]
    mat = torch.from_numpy(pot_mask(*cfg.shape)).cuda()
    code = build_codes(mat)
    nfl = fluid_count(mat)
    t0 = time.time()
    rs, es = generate_pairs(core, mat, range(args.pairs), log=log)
    rh, eh = generate_pairs(core, mat, range(10000, 10000 + args.heldout), log=log)
    torch.cuda.synchronize()
    t_pairs = time.time() - t0
    log(f"{len(rs)} training + {len(rh)} held-out pairs in {t_pairs:.1f}s")
    net = UNet(seed=args.seed)
    l0 = evaluate_loss(net, code, nfl, rh, eh)
    log(f"held-out loss before training {l0:.4f}")
    t0 = time.time()
    hist = train(net, code, nfl, rs, es, epochs=args.epochs, batch=args.batch, lr=args.lr, seed=args.seed, log=log)
    torch.cuda.synchronize()
    t_train = time.time() - t0
    l1 = evaluate_loss(net, code, nfl, rh, eh)
    log(f"held-out loss after training {l1:.4f} (train {t_train:.1f}s)")
    net.save(args.out, loss_before=l0, loss_after=l1, train_hist=np.array(hist), pairs=args.pairs,
             epochs=args.epochs)
    # Held-out solves: flexible CG with the network vs plain PCG, tol 1e-6.
    shape = cfg.shape
    nb = core.fcg_workspace_bytes(*shape, PRECOND_UNET)
    wsf = torch.empty(nb, dtype=torch.uint8, device="cuda")
    wsp = torch.empty(core.pcg_workspace_bytes(*shape), dtype=torch.uint8, device="cuda")
    x = torch.zeros(shape, dtype=torch.float32, device="cuda")
    res = []
    for s in range(10000, 10000 + args.eval_solves):
        b, _ = grf_torch(torch, mat == 0, s)
        x.zero_()
        torch.cuda.synchronize()
        t = time.perf_counter()
        sp = core.pcg_solve(x.data_ptr(), b.data_ptr(), mat.data_ptr(), *shape, *_UNIT, 1e-6, 20000,
                            wsp.data_ptr(), wsp.numel())
        torch.cuda.synchronize()
        tp = time.perf_counter() - t
        x.zero_()
        t = time.perf_counter()
        sf = core.fcg_solve(x.data_ptr(), b.data_ptr(), mat.data_ptr(), *shape, *_UNIT, 1e-6, 20000, PRECOND_UNET,
                            net.params.data_ptr(), net.packed.data_ptr(), wsf.data_ptr(), wsf.numel())
        torch.cuda.synchronize()
        tf = time.perf_counter() - t
        res.append((sp["iterations"], sf["iterations"], sf["pcg_iterations"], tp * 1e3, tf * 1e3,
                    sf["relres_true"], sf["converged"]))
    a = np.array(res, dtype=np.float64)
    summary = {"pcg_iters": a[:, 0].mean(), "fcg_iters": a[:, 1].mean(), "fcg_fallback_pcg_iters": a[:, 2].mean(),
               "pcg_ms": a[:, 3].mean(), "fcg_ms": a[:, 4].mean(), "fcg_relres_max": a[:, 5].max(),
               "all_converged": bool(a[:, 6].all()), "loss_before": l0, "loss_after": l1,
               "t_pairs_s": t_pairs, "t_train_s": t_train, "train_hist": hist}
    log("summary " + json.dumps(summary))
    return summary


if __name__ == "__main__":
    main()
