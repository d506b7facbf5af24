"""Empty grids through every grid entry point of the C ABI.  The reference's
launchers return success without touching anything when a dimension is
<= 0 (crates/cuda-kernels/src/jacobi_pressure.cu:193, 238;
pressure_solve_pcg.cu:120-123, which also reports 0 iterations): every entry
point here does the same, leaving sentinel-filled buffers bit for bit as they
were, and reports 0 for the statistics and device maxima it returns."""

from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SHAPES = [(0, 5, 5), (5, 0, 5), (5, 5, 0), (0, 0, 0)]
SENTINEL = 777.25
N_BUF = 8 * 8 * 8


class Bufs:
    """Named float32 buffers of N_BUF sentinels (the largest face array of any
    shape above fits), int32 material codes, uint8 class codes, float64
    buffers and two 32-bit max slots preset to all ones."""

    def __init__(self, torch):
        self.torch = torch
        self.f = {}
        self.mat = torch.zeros(N_BUF, dtype=torch.int32, device="cuda")
        self.code = torch.full((N_BUF,), 7, dtype=torch.uint8, device="cuda")
        self.bits = torch.full((2,), -1, dtype=torch.int32, device="cuda")
        self.f64 = torch.full((N_BUF,), SENTINEL, dtype=torch.float64, device="cuda")
        self.part = torch.full((N_BUF,), SENTINEL, dtype=torch.float64, device="cuda")

    def __getitem__(self, name):
        if name not in self.f:
            self.f[name] = self.torch.full((N_BUF,), SENTINEL, dtype=self.torch.float32, device="cuda")
        return self.f[name].data_ptr()

    def untouched(self):
        self.torch.cuda.synchronize()
        for name, t in self.f.items():
            assert bool((t == SENTINEL).all()), name
        assert bool((self.code == 7).all())
        assert bool((self.f64 == SENTINEL).all()) and bool((self.part == SENTINEL).all())
        assert bool((self.mat == 0).all())


def _bits(b):
    b.torch.cuda.synchronize()
    return b.bits.cpu().numpy().view(np.uint32).tolist()


@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_empty_grid_is_a_no_op(simcore, torch_gpu, shape):
    torch = torch_gpu
    from boilsim.learned import default_net

    print(f"[empty] shape {shape}", flush=True)
    net = default_net()
    b = Bufs(torch)
    nx, ny, nz = shape
    g = (nx, ny, nz)
    mat = b.mat.data_ptr()
    dx, dt, rho = 0.002, 0.001, 1000.0
    origin = (-0.1, -0.1, 0.0)

    simcore.scale_array(b["s"], 0, 3.0)
    simcore.jacobi_pressure_step(b["pn"], b["p"], b["div"], mat, *g, dx, dt, rho)
    simcore.pressure_solve(b["p"], b["pt"], b["div"], mat, *g, dx, dt, rho, 50)
    simcore.compute_divergence(b["div"], b["ux"], b["uy"], b["uz"], mat, *g, dx)
    simcore.subtract_pressure_gradient(b["ux"], b["uy"], b["uz"], b["p"], mat, *g, dx, dt, rho)
    simcore.enforce_no_slip(b["ux"], b["uy"], b["uz"], mat, *g)
    simcore.divergence_no_slip(b["div"], b["ux"], b["uy"], b["uz"], mat, *g, dx)
    simcore.extend_temperature(b["te"], b["t"], mat, *g)
    simcore.advect(b["uxn"], b["uyn"], b["uzn"], b["tn"], b["ux"], b["uy"], b["uz"], b["t"], b["te"], mat, *g,
                   origin, dx, dt)
    simcore.advect_all(b["ux"], b["uy"], b["uz"], b["t"], b["uxn"], b["uyn"], b["uzn"], b["tn"], b["te"], mat, *g,
                       origin, dx, dt)
    simcore.conduction_sweep(b["tn"], b["t"], b["to"], mat, b["k"], b["rho"], b["cp"], 4, *g, dx, dt, 10.0, 293.15)
    simcore.conduction_solve(b["t"], b["tn"], b["to"], mat, b["k"], b["rho"], b["cp"], 4, *g, dx, dt, 10.0,
                             293.15, 20)
    simcore.mg_smooth_restrict(b.code.data_ptr(), b["r"], b["z"], b["rc"], *g, fused=False)
    simcore.mg_chebyshev(b.code.data_ptr(), b["r"], b["z"], b["tmp"], *g, 10, 0.02)
    assert simcore.mg_prolong_post(b.code.data_ptr(), b["z"], b["ec"], b["r"], b["zo"], b["zn"], b["tmp"],
                                   b.part.data_ptr(), N_BUF, *g, fused=False) == 0

    assert simcore.pcg_workspace_bytes(*g) == 0
    assert simcore.fcg_workspace_bytes(*g, 2) == 0
    assert simcore.refine_workspace_bytes(*g) == 0
    assert simcore.fcg_refine_workspace_bytes(*g, 2) == 0
    ws = b["ws"]
    its = simcore.pressure_solve_pcg(b["p"], b["div"], mat, *g, dx, dt, rho, 1e-6, 100, b["wb"], b["wr"], b["wz"],
                                     b["wp"], b["wap"], b["wdot"], *(b[f"d{i}"] for i in range(7)))
    assert its == 0
    stats = [
        simcore.pcg_solve(b["p"], b["div"], mat, *g, dx, dt, rho, 1e-6, 100, ws, 0),
        simcore.pcg_solve_noalloc(b["p"], b["div"], mat, *g, dx, dt, rho, 1e-6, 100),
        simcore.fcg_solve(b["p"], b["div"], mat, *g, dx, dt, rho, 1e-6, 100, 2, net.params.data_ptr(),
                          net.packed.data_ptr(), ws, 0),
        simcore.fcg_solve(b["p"], b["div"], mat, *g, dx, dt, rho, 1e-6, 100, 5, 0, 0, ws, 0),
        simcore.refine_solve(b.f64.data_ptr(), b["div"], mat, *g, dx, dt, rho, 1e-6, ws, 0, storage="fp16",
                             x32_ptr=b["x32"]),
        simcore.fcg_refine_solve(b.f64.data_ptr(), b["div"], mat, *g, dx, dt, rho, 1e-6, 2, net.params.data_ptr(),
                                 net.packed.data_ptr(), ws, 0, x32_ptr=b["x32"]),
    ]
    for st in stats:
        assert all(v == 0 for v in st.values()), st
    b.untouched()

    # Entry points that report device maxima write 0 (the maximum over no cells).
    for call in (
        lambda p: simcore.max_abs(b["div"], 0, p),
        lambda p: simcore.max_velocity(b["ux"], b["uy"], b["uz"], *g, p),
        lambda p: simcore.divergence_max(b["div"], b["ux"], b["uy"], b["uz"], mat, *g, dx, p),
        lambda p: simcore.projection_jacobi(b["p"], b["pt"], b["div"], b["ux"], b["uy"], b["uz"], mat, *g, dx, dt,
                                            rho, 50, p),
        lambda p: simcore.projection_step_jacobi(b["p"], b["pt"], b["div"], b["ux"], b["uy"], b["uz"], mat, *g, dx,
                                                 dt, rho, 50, p),
    ):
        b.bits.fill_(-1)
        call(b.bits.data_ptr())
        got = _bits(b)
        assert got[0] == 0
        b.untouched()
    for call in (
        lambda p: simcore.divergence_max(b["div"], b["ux"], b["uy"], b["uz"], mat, *g, dx, p),
        lambda p: simcore.projection_step_jacobi(b["p"], b["pt"], b["div"], b["ux"], b["uy"], b["uz"], mat, *g, dx,
                                                 dt, rho, 50, p),
    ):
        b.bits.fill_(-1)
        call(b.bits.data_ptr())
        assert _bits(b) == [0, 0]


def test_negative_dims_through_the_abi(simcore, torch_gpu):
    """SimCore refuses a negative size as the reference's usize extraction
    does; the C ABI itself takes int dimensions, and a negative one is the
    same silent no-op as 0."""
    import ctypes

    from sim_core import _lib

    torch = torch_gpu
    b = Bufs(torch)
    mat = b.mat.data_ptr()
    with pytest.raises(OverflowError):
        simcore.pressure_solve(b["p"], b["pt"], b["div"], mat, -3, 5, 5, 0.002, 0.001, 1000.0, 10)
    _lib.call("bs_jacobi_pressure_launch", b["pn"], b["p"], b["div"], mat, 5, -3, 5, 0.002, 0.001, 1000.0, 0, 2)
    _lib.call("bs_pressure_solve_launch", b["p"], b["pt"], b["div"], mat, 5, 5, -3, 0.002, 0.001, 1000.0, 10, 0, 2)
    st = _lib.SolveStats()
    st.iterations = 99
    _lib.call("bs_pcg_solve", b["p"], b["div"], mat, -3, 5, 5, 0.002, 0.001, 1000.0, 0, 2, 1e-6, 100, 0, 4,
              b["ws"], 0, None, ctypes.byref(st))
    assert st.iterations == 0
    fs = _lib.FcgStats()
    fs.iterations = 99
    _lib.call("bs_fcg_solve", b["p"], b["div"], mat, 5, -3, 5, 0.002, 0.001, 1000.0, 0, 2, 1e-6, 100, 0, 5,
              None, None, b["ws"], 0, None, ctypes.byref(fs))
    assert fs.iterations == 0
    b.untouched()
