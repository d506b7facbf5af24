"""Solves issued from two host threads at once (ctypes releases the GIL for
the call): the solver entry points serialise on the library's lock, so each
result is bit for bit the one of the same solve run alone — a learned
flexible CG (captured graphs, pinned polling) in one thread against a plain
true-residual PCG and another flexible CG in the other."""

from __future__ import annotations

import threading

import numpy as np
import pytest

import oracle
from helpers import dev, host

pytestmark = pytest.mark.gpu


def _case(torch, seed):
    from boilsim.fixtures import CONFIGS, DT, RHO, pot_mask, seeded_velocities

    cfg = CONFIGS["dev"]
    mat = pot_mask(*cfg.shape)
    div = oracle.divergence(*seeded_velocities(*cfg.shape, seed=seed, plume=False), mat, cfg.dx)
    return cfg, dev(torch, mat), dev(torch, div), DT, RHO


def test_concurrent_solves_match_sequential(simcore, torch_gpu):
    torch = torch_gpu
    from boilsim.learned import default_net

    net = default_net()
    jobs = []
    for seed, kind in ((11, "fcg"), (21, "pcg"), (31, "fcg")):
        cfg, mat_d, div_d, dt, rho = _case(torch, seed)
        nb = simcore.fcg_workspace_bytes(*cfg.shape, 2) if kind == "fcg" else simcore.pcg_workspace_bytes(*cfg.shape)
        jobs.append(dict(kind=kind, cfg=cfg, mat=mat_d, div=div_d, dt=dt, rho=rho,
                         ws=torch.empty(nb, dtype=torch.uint8, device="cuda"),
                         x=torch.zeros(cfg.shape, dtype=torch.float32, device="cuda")))

    def run(j):
        j["x"].zero_()
        torch.cuda.synchronize()
        c = j["cfg"]
        if j["kind"] == "fcg":
            st = simcore.fcg_solve(j["x"].data_ptr(), j["div"].data_ptr(), j["mat"].data_ptr(), *c.shape, c.dx,
                                   j["dt"], j["rho"], 1e-6, 5000, 2, net.params.data_ptr(), net.packed.data_ptr(),
                                   j["ws"].data_ptr(), j["ws"].numel())
        else:
            st = simcore.pcg_solve(j["x"].data_ptr(), j["div"].data_ptr(), j["mat"].data_ptr(), *c.shape, c.dx,
                                   j["dt"], j["rho"], 1e-6, 20000, j["ws"].data_ptr(), j["ws"].numel())
        torch.cuda.synchronize()
        return st

    alone = []
    for j in jobs:
        st = run(j)
        alone.append((st["iterations"], host(j["x"]).copy()))
    print(f"[threads] alone: {[a[0] for a in alone]}", flush=True)

    out = [None] * len(jobs)
    err = []

    def worker(idxs):
        try:
            for i in idxs:
                st = run(jobs[i])
                out[i] = (st["iterations"], host(jobs[i]["x"]).copy())
        except Exception as e:  # reported below
            err.append(e)

    ts = [threading.Thread(target=worker, args=([0],)), threading.Thread(target=worker, args=([1, 2],))]
    for t in ts:
        t.start()
    for t in ts:
        t.join(timeout=120)
    assert not any(t.is_alive() for t in ts), "a solver thread did not finish"
    assert not err, err
    print(f"[threads] concurrent: {[o[0] for o in out]}", flush=True)
    for (ia, xa), (ib, xb) in zip(alone, out):
        assert ia == ib
        assert np.array_equal(xa.view(np.uint32), xb.view(np.uint32))
