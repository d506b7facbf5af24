"""GPU tests of the warm-started AR(1) sequence (config 5's right-hand sides):
every frame of every method reaches the tolerance by the oracle's FP64
residual, and its error is stated against the oracle's own FP64 PCG solution
of the same frame to 1e-12 (independent of the device), on the 110x110x70
pot (the learned flexible CG, the same V-cycle without the network, and
Jacobi-PCG) and, for the learned method, at the production grid 440x440x280.  The
reference cold-starts every projection (python/boilingsim/fluid.py:627-629);
warm starts are new here, so only the oracle can judge them."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from helpers import host

pytestmark = pytest.mark.gpu

TOL = 1e-6


def test_sequence_dev_against_oracle(simcore, torch_gpu):
    torch = torch_gpu
    from boilsim.fixtures import pot_mask
    from boilsim.learned import default_net
    from boilsim.sequence import SequenceSolver, ar1_frames

    mat = pot_mask(110, 110, 70)
    mat_d = torch.from_numpy(mat).cuda()
    fl = mat == 0
    methods = ("learned", "mg", "pcg-fp32")
    solvers = {m: SequenceSolver(simcore, torch, mat_d, m, net=default_net(), tol=TOL) for m in methods}
    its = {m: [] for m in methods}
    x_ref = None
    for t, b in ar1_frames(torch, mat_d, 4):
        b64 = host(b).astype(np.float64)
        # The oracle's FP64 PCG, warm-started from its own previous frame.
        x_ref, _ = oracle.pcg_f64(b64, mat, tol=1e-12, x0=x_ref)
        for m, s in solvers.items():
            print(f"[sequence 110] frame {t} {m}", flush=True)
            st = s.solve(b)
            x = host(s.x)
            rel2, _ = oracle.residual_f64(x, b64, mat)
            err = np.linalg.norm((x - x_ref)[fl]) / np.linalg.norm(x_ref[fl])
            print(f"  {st}; oracle relres {rel2:.3e}; error vs oracle FP64 {err:.3e}")
            assert st["converged"] == 1 and rel2 <= 1.05 * TOL
            assert err < 1e-4
            its[m].append(st["inner_iterations"])
    # Warm starts cut the plain method's iterations after the first frame;
    # both V-cycle methods (with and without the network) need far fewer.
    assert max(its["pcg-fp32"][1:]) < its["pcg-fp32"][0]
    assert 4 * np.mean(its["learned"]) < np.mean(its["pcg-fp32"])
    assert 4 * np.mean(its["mg"]) < np.mean(its["pcg-fp32"])


def test_sequence_production_grid_against_oracle(simcore, torch_gpu):
    """Four warm-started frames at 440x440x280 with the learned method; the
    oracle's FP64 residual of every frame's FP64 iterate <= 1.05 tol, and the
    warm-started frames need fewer iterations than the cold first one."""
    torch = torch_gpu
    from boilsim.fixtures import CONFIGS, pot_mask
    from boilsim.learned import default_net
    from boilsim.sequence import SequenceSolver, ar1_frames

    mat = pot_mask(*CONFIGS["prod"].shape)
    mat_d = torch.from_numpy(mat).cuda()
    s = SequenceSolver(simcore, torch, mat_d, "learned", net=default_net(), tol=TOL)
    its = []
    for t, b in ar1_frames(torch, mat_d, 4):
        print(f"[sequence 440] frame {t}", flush=True)
        st = s.solve(b)
        rel2, _ = oracle.residual_f64(host(s.x), host(b).astype(np.float64), mat)
        print(f"  {st}; oracle relres {rel2:.3e}", flush=True)
        assert st["converged"] == 1 and rel2 <= 1.05 * TOL
        assert abs(rel2 - st["relres_true"]) <= 1e-3 * rel2
        its.append(st["inner_iterations"])
    assert max(its[1:]) < its[0]
