"""Learned preconditioner: training-data generation from the plain solver,
training with the library's own forward/backward/Adam kernels, and the
learned flexible-CG solve used by the ``lpcg-hip`` projection path.

Training pairs (residual, correction) come from plain Jacobi-PCG runs on
seeded right-hand sides: for RHS b (a masked, unit-variance Gaussian random
field, correlation length U[4,16] cells) the solver is stopped after k
iterations at x_k, then run on to x* (relative residual 1e-9); the pair is
r_k = b - A x_k and e_k = x* - x_k, so A e_k = r_k.  The network is trained
on the A-norm loss (z-e)^T A (z-e) / (e.r) of its correction z = N(r)."""

from __future__ import annotations

import math
import time

import numpy as np

from .geometry import MAT_AIR, MAT_FLUID
from .unet import NUM_PARAMS, UNet, build_codes

# b is passed to the solvers as "div" with dx = 1, dt = -1, rho = 1, which makes
# their RHS scale -rho*dx*dx/dt exactly 1.
_UNIT = (1.0, -1.0, 1.0)

K_CHOICES = (0, 1, 2, 4, 8, 16, 32, 64, 128)

PRECOND_UNET, PRECOND_MG_UNET, PRECOND_MG_JACOBI = 1, 2, 3
# The same V-cycle without the network: 60 Chebyshev steps on the coarsest
# level (5), or two more geometric levels and a bottom PCG (6).
PRECOND_MG_CHEB, PRECOND_MG_DEEP = 5, 6
NETWORK_KINDS = (PRECOND_UNET, PRECOND_MG_UNET, 4)


def grf_torch(torch, fluid, seed: int, ell: float | None = None, ell_range=(4.0, 16.0)):
    """Unit-variance (over fluid) Gaussian random field on the device, zero off
    fluid; white noise from torch's generator seeded with `seed`, smoothed by a
    Gaussian of std `ell` cells in Fourier space.  Returns (field, ell)."""
    g = torch.Generator(device=fluid.device)
    g.manual_seed(int(seed))
    if ell is None:
        u = torch.rand(1, generator=g, device=fluid.device).item()
        ell = ell_range[0] + (ell_range[1] - ell_range[0]) * u
    shape = tuple(fluid.shape)
    noise = torch.randn(shape, generator=g, device=fluid.device, dtype=torch.float32)
    f = torch.fft.rfftn(noise)
    kx = torch.fft.fftfreq(shape[0], device=fluid.device)[:, None, None]
    ky = torch.fft.fftfreq(shape[1], device=fluid.device)[None, :, None]
    kz = torch.fft.rfftfreq(shape[2], device=fluid.device)[None, None, :]
    k2 = (kx * kx + ky * ky + kz * kz) * (2.0 * math.pi) ** 2
    f = f * torch.exp(-0.5 * k2 * ell * ell)
    x = torch.fft.irfftn(f, s=shape)
    vals = x[fluid]
    vals = (vals - vals.mean()) / vals.std().clamp_min(1e-30)
    out = torch.zeros(shape, dtype=torch.float32, device=fluid.device)
    out[fluid] = vals
    return out, ell


class PairGenerator:
    """(r_k, e_k) pairs from the library's plain PCG on one mask."""

    def __init__(self, core, mat_t):
        import torch

        self.torch = torch
        self.core = core
        self.mat = mat_t
        self.shape = tuple(mat_t.shape)
        self.fluid = mat_t == MAT_FLUID
        nb = core.pcg_workspace_bytes(*self.shape)
        self.ws = torch.empty(nb, dtype=torch.uint8, device=mat_t.device)
        self.ap = torch.empty(self.shape, dtype=torch.float32, device=mat_t.device)

    def solve(self, x, b, tol, max_iter, warm, restarts):
        return self.core.pcg_solve(x.data_ptr(), b.data_ptr(), self.mat.data_ptr(), *self.shape, *_UNIT, tol,
                                   max_iter, self.ws.data_ptr(), self.ws.numel(), warm_start=warm,
                                   max_restarts=restarts)

    def residual(self, x, b):
        from sim_core import _lib

        _lib.call("bs_laplacian_spmv_launch", self.ap.data_ptr(), x.data_ptr(), self.mat.data_ptr(), *self.shape,
                  MAT_FLUID, MAT_AIR)
        return b - self.ap

    def pair(self, seed: int, k: int):
        torch = self.torch
        b, _ = grf_torch(torch, self.fluid, seed)
        x = torch.zeros(self.shape, dtype=torch.float32, device=b.device)
        if k > 0:
            self.solve(x, b, 0.0, k, False, 0)  # exactly k iterations
        xk = x.clone()
        # On to the FP32 attainable accuracy (true relative residual ~1e-7).
        self.solve(x, b, 1e-7, 20000, k > 0, 2)
        return self.residual(xk, b), x - xk


def generate_pairs(core, mat_t, seeds, k_choices=K_CHOICES, log=None):
    gen = PairGenerator(core, mat_t)
    rs, es = [], []
    t0 = time.time()
    for n, s in enumerate(seeds):
        k = k_choices[s % len(k_choices)]
        r, e = gen.pair(int(s), k)
        rs.append(r)
        es.append(e)
        if log and (n + 1) % 200 == 0:
            log(f"pairs {n + 1}/{len(seeds)} ({time.time() - t0:.1f}s)")
    return rs, es


def train(net: UNet, code, n_fluid: int, rs, es, epochs: int = 3, batch: int = 8, lr: float = 2e-3,
          seed: int = 0, log=None):
    """Adam on the A-norm loss with the library's kernels; returns the mean
    loss of each epoch."""
    import torch

    from sim_core import _lib

    shape = tuple(code.shape)
    ws = torch.zeros(UNet.workspace_bytes(shape, training=True), dtype=torch.uint8, device=code.device)
    z = torch.zeros(shape, dtype=torch.float32, device=code.device)
    grads = torch.zeros(NUM_PARAMS, dtype=torch.float32, device=code.device)
    m = torch.zeros_like(grads)
    v = torch.zeros_like(grads)
    loss2 = torch.zeros((batch, 2), dtype=torch.float64, device=code.device)
    rng = np.random.default_rng(seed)
    step = 0
    hist = []
    n = len(rs)
    steps_total = epochs * (n // batch)
    for ep in range(epochs):
        order = rng.permutation(n)
        losses = []
        for b0 in range(0, n - batch + 1, batch):
            grads.zero_()
            for q, idx in enumerate(order[b0:b0 + batch]):
                net.forward(rs[idx], code, n_fluid, z, ws, training=True)
                net.backward(rs[idx], es[idx], code, grads, loss2[q], ws)
            step += 1
            cur_lr = lr * 0.5 * (1.0 + math.cos(math.pi * step / steps_total))
            _lib.call("bs_unet_adam", net.params.data_ptr(), grads.data_ptr(), m.data_ptr(), v.data_ptr(), step,
                      cur_lr, 1.0 / batch, 0)
            net.pack()
            l = loss2.cpu().numpy()
            losses.extend((l[:, 0] / l[:, 1]).tolist())
        hist.append(float(np.mean(losses)))
        if log:
            log(f"epoch {ep + 1}/{epochs}: mean A-norm loss {hist[-1]:.4f}")
    return hist


def evaluate_loss(net: UNet, code, n_fluid: int, rs, es) -> float:
    import torch

    shape = tuple(code.shape)
    ws = torch.zeros(UNet.workspace_bytes(shape, training=True), dtype=torch.uint8, device=code.device)
    z = torch.zeros(shape, dtype=torch.float32, device=code.device)
    grads = torch.zeros(NUM_PARAMS, dtype=torch.float32, device=code.device)
    loss2 = torch.zeros(2, dtype=torch.float64, device=code.device)
    out = []
    for r, e in zip(rs, es):
        net.forward(r, code, n_fluid, z, ws, training=True)
        net.backward(r, e, code, grads, loss2, ws)
        l = loss2.cpu().numpy()
        out.append(l[0] / l[1])
    return float(np.mean(out))


_FCG_WS = {"key": None, "buf": None}
_REFINE_WS = {"key": None, "buf": None, "x64": None}
_DEFAULT_NET = {"net": None}

# Refinement steps of one lpcg-hip projection at most (each cold-starts the
# inner solver on the new residual).
MAX_REFINE_STEPS = 20


def default_net() -> UNet:
    """The shipped network, loaded (read, copied to the device and packed)
    once per process, so every projection reuses the same parameter buffers
    (and the flexible CG's captured graphs, which are keyed on them)."""
    if _DEFAULT_NET["net"] is None:
        _DEFAULT_NET["net"] = UNet.load()
    return _DEFAULT_NET["net"]


def learned_pcg_solve(core, grid, ws, dt, rho, tol, max_iter, net: UNet | None = None,
                      precond: int = PRECOND_MG_UNET, warm_start: bool = False, refine: bool = True) -> dict:
    """The lpcg-hip solve of the projection: flexible CG with the learned
    preconditioner on grid.p from ws.div_u (b = -rho dx^2 div / dt).

    FP32 vectors cannot reach every tolerance: on right-hand sides with a
    smooth component (a buoyant plume's divergence, say) the true relative
    residual of an FP32 iterate stalls near 1e-6.  With ``refine`` an
    unconverged solve continues from its iterate in FP64 iterative refinement
    around the same learned solver (bs_fcg_refine_solve) and grid.p receives
    the refined solution rounded to FP32; the stats then carry
    ``refined: True`` and the refinement's true residual.  ``max_iter``
    bounds the whole solve: the refinement's inner iterations get what the
    FP32 solve left of it (and at most MAX_REFINE_STEPS steps); ``capped``
    says a cap ended an unconverged solve."""
    import torch

    if net is None and precond in NETWORK_KINDS:
        net = default_net()
    params = net.params.data_ptr() if net is not None else 0
    packed = net.packed.data_ptr() if net is not None else 0
    shape = grid.shape
    key = (shape, precond)
    if _FCG_WS["key"] != key:
        nb = core.fcg_workspace_bytes(*shape, precond)
        _FCG_WS["buf"] = torch.empty(nb, dtype=torch.uint8, device=grid.p.device)
        _FCG_WS["key"] = key
    buf = _FCG_WS["buf"]
    st = core.fcg_solve(grid.p.data_ptr(), ws.div_u.data_ptr(), grid.mat.data_ptr(), *shape, grid.dx, dt, rho,
                        tol, max_iter, precond, params, packed, buf.data_ptr(), buf.numel(), warm_start=warm_start)
    used = st["iterations"] + st["pcg_iterations"]
    if st["converged"] or not refine or used >= max_iter:
        if not st["converged"]:
            st["capped"] = used >= max_iter
        return st
    if _REFINE_WS["key"] != key:
        nb = core.fcg_refine_workspace_bytes(*shape, precond)
        _REFINE_WS["buf"] = torch.empty(nb, dtype=torch.uint8, device=grid.p.device)
        _REFINE_WS["x64"] = torch.empty(shape, dtype=torch.float64, device=grid.p.device)
        _REFINE_WS["key"] = key
    x64, rbuf = _REFINE_WS["x64"], _REFINE_WS["buf"]
    x64.copy_(grid.p)
    rs = core.fcg_refine_solve(x64.data_ptr(), ws.div_u.data_ptr(), grid.mat.data_ptr(), *shape, grid.dx, dt, rho,
                               tol, precond, params, packed, rbuf.data_ptr(),
                               rbuf.numel(), max_outer=MAX_REFINE_STEPS, max_inner=max_iter - used,
                               warm_start=True, x32_ptr=grid.p.data_ptr())
    out = dict(st)
    out.update({"refined": True, "fp32_iterations": st["iterations"], "fp32_relres_true": st["relres_true"],
                "iterations": st["iterations"] + rs["inner_iterations"], "refine_steps": rs["outer"],
                "refine_fallback": rs["fallback"], "converged": rs["converged"], "relres_true": rs["relres_true"]})
    if not rs["converged"]:
        out["capped"] = ("max_iter" if used + rs["inner_iterations"] >= max_iter else
                         "refine_steps" if rs["outer"] >= MAX_REFINE_STEPS else False)
    return out


def fluid_count(mat_t) -> int:
    return int((mat_t == MAT_FLUID).sum().item())


def codes_for(mat_t):
    return build_codes(mat_t)
