"""GPU test of the warm-started sequence driver on the 110x110x70 pot: every
frame of both methods reaches the tolerance on the FP64 residual, warm starts
cut the plain method's iterations after the first frame, the learned method
needs far fewer iterations, and both agree with an FP64 reference solve."""

from __future__ import annotations

import pytest

pytestmark = pytest.mark.gpu


def test_sequence_dev(simcore, torch_gpu):
    torch = torch_gpu
    from boilsim.fixtures import pot_mask
    from boilsim.sequence import run_sequence, summarize
    from boilsim.unet import UNet

    mat_d = torch.from_numpy(pot_mask(110, 110, 70)).cuda()
    res = run_sequence(simcore, torch, mat_d, 4, methods=("learned", "pcg-fp32"), net=UNet.load(), tol=1e-6,
                       error_every=1)
    s = summarize(res)
    print(s)
    for m in ("learned", "pcg-fp32"):
        assert s[m]["all_converged"] and s[m]["max_relres_true"] <= 1e-6
        assert s[m]["max_err_vs_fp64"] < 1e-3
    plain = [r["inner_iterations"] for r in res["pcg-fp32"]]
    assert max(plain[1:]) < plain[0]
    assert s["learned"]["iterations_mean"] * 4 < s["pcg-fp32"]["iterations_mean"]
