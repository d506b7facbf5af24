"""GPU tests of the backward-Euler Jacobi conduction (SimCore.conduction_sweep /
conduction_solve against the oracle's restatement of thermal.py:259-388 and
631-665): bit-identical sweeps on random materials with fluid on the domain
edges, the ping-pong driver's three buffers, and the pot geometry."""

from __future__ import annotations

import numpy as np
import pytest

import oracle
from helpers import bit_diff_report, dev, host, same_bits

pytestmark = pytest.mark.gpu

H_CONV, T_AMB = 10.0, 295.15


def _props(torch, pot="steel_304"):
    from boilsim.thermal import MaterialProps

    P = MaterialProps.from_pot(pot)
    return P, P.device(torch)


def _case(seed, shape):
    rng = np.random.default_rng(seed)
    mat = rng.choice([0, 1, 2, 3], size=shape, p=[0.45, 0.2, 0.25, 0.1]).astype(np.int32)
    t_iter = (290.0 + 110.0 * rng.random(shape)).astype(np.float32)
    t_old = (290.0 + 110.0 * rng.random(shape)).astype(np.float32)
    dx = float(rng.uniform(5e-4, 4e-3))
    dt = float(rng.uniform(0.01, 2.0))
    return mat, t_iter, t_old, dx, dt


@pytest.mark.parametrize("seed,shape", [(1, (16, 16, 16)), (2, (24, 12, 20)), (3, (9, 7, 13)), (4, (33, 17, 44)),
                                        (5, (5, 40, 2)), (6, (64, 64, 48))])
@pytest.mark.parametrize("pot", ["steel_304", "copper"])
def test_sweep_bit_exact(simcore, torch_gpu, seed, shape, pot):
    torch = torch_gpu
    P, (k, rho, cp) = _props(torch, pot)
    mat, t_iter, t_old, dx, dt = _case(seed, shape)
    marker = np.full(shape, -7.25, np.float32)
    ref = marker.copy()
    oracle.conduction_sweep(ref, t_iter, t_old, mat, P.k, P.rho, P.c_p, dx, dt, H_CONV, T_AMB)
    # keep every device tensor referenced until the launch has completed
    out, ti, to, md = dev(torch, marker), dev(torch, t_iter), dev(torch, t_old), dev(torch, mat)
    simcore.conduction_sweep(out.data_ptr(), ti.data_ptr(), to.data_ptr(), md.data_ptr(), k.data_ptr(),
                             rho.data_ptr(), cp.data_ptr(), k.numel(), *shape, dx, dt, H_CONV, T_AMB)
    torch.cuda.synchronize()
    got = host(out)
    assert same_bits(got, ref), bit_diff_report(got, ref)
    assert (got[mat == 2] == -7.25).all()  # air cells untouched


def test_sweep_misaligned_views(simcore, torch_gpu):
    """Views offset by one element (no 16-byte alignment) take the scalar path."""
    torch = torch_gpu
    P, (k, rho, cp) = _props(torch)
    shape = (12, 10, 16)
    n = 12 * 10 * 16
    mat, t_iter, t_old, dx, dt = _case(9, shape)
    ref = np.zeros(shape, np.float32)
    oracle.conduction_sweep(ref, t_iter, t_old, mat, P.k, P.rho, P.c_p, dx, dt, H_CONV, T_AMB)
    buf = torch.zeros(4 * (n + 1), dtype=torch.float32, device="cuda")
    ti = buf[1:n + 1]
    to = buf[n + 2:2 * n + 2]
    out = buf[2 * n + 3:3 * n + 3]
    ti.copy_(torch.from_numpy(t_iter.ravel()))
    to.copy_(torch.from_numpy(t_old.ravel()))
    mb = torch.zeros(n + 1, dtype=torch.int32, device="cuda")
    mv = mb[1:]
    mv.copy_(torch.from_numpy(mat.ravel()))
    simcore.conduction_sweep(out.data_ptr(), ti.data_ptr(), to.data_ptr(), mv.data_ptr(), k.data_ptr(),
                             rho.data_ptr(), cp.data_ptr(), k.numel(), *shape, dx, dt, H_CONV, T_AMB)
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(shape)
    assert same_bits(got, ref), bit_diff_report(got, ref)


@pytest.mark.parametrize("n_iter", [1, 2, 3, 15, 16])
def test_solve_driver_buffers(simcore, torch_gpu, n_iter):
    """conduct_one_step's implicit branch: T, the ping-pong partner and the
    T_old snapshot all end bit-identical to the reference's driver."""
    torch = torch_gpu
    P, (k, rho, cp) = _props(torch)
    shape = (20, 18, 24)
    mat, t0, _, dx, dt = _case(21, shape)
    alt0 = np.full(shape, 1.5, np.float32)
    t_r, alt_r, old_r = t0.copy(), alt0.copy(), np.zeros(shape, np.float32)
    oracle.conduction_solve(t_r, alt_r, old_r, mat, P.k, P.rho, P.c_p, dx, dt, H_CONV, T_AMB, n_iter)
    t, alt, old = dev(torch, t0), dev(torch, alt0), torch.zeros(shape, dtype=torch.float32, device="cuda")
    md = dev(torch, mat)
    simcore.conduction_solve(t.data_ptr(), alt.data_ptr(), old.data_ptr(), md.data_ptr(), k.data_ptr(),
                             rho.data_ptr(), cp.data_ptr(), k.numel(), *shape, dx, dt, H_CONV, T_AMB, n_iter)
    torch.cuda.synchronize()
    for g, r in ((host(t), t_r), (host(alt), alt_r), (host(old), old_r)):
        assert same_bits(g, r), bit_diff_report(g, r)


def test_pot_conduction_step(simcore, torch_gpu):
    """The 110x110x70 pot with a hot base, one step of the reference's default
    15 sweeps through boilsim.thermal.conduct_implicit: bit-identical to the
    oracle; heat flows from the base into the water."""
    torch = torch_gpu
    from boilsim.fixtures import pot_mask
    from boilsim.thermal import allocate_thermal_workspace, conduct_implicit

    P, props_dev = _props(torch)
    shape = (110, 110, 70)
    mat = pot_mask(*shape)
    t0 = np.full(shape, 295.15, np.float32)
    base = (mat == 1) & (np.arange(70)[None, None, :] < 8)
    t0[base] = 500.0
    dx, dt = 2e-3, 0.05
    t_r, alt_r, old_r = t0.copy(), np.zeros(shape, np.float32), np.zeros(shape, np.float32)
    oracle.conduction_solve(t_r, alt_r, old_r, mat, P.k, P.rho, P.c_p, dx, dt, H_CONV, 22.0 + 273.15, 15)
    T, md = dev(torch, t0), dev(torch, mat)
    ws = allocate_thermal_workspace(torch, shape)
    conduct_implicit(simcore, T, md, props_dev, ws, dx, dt, 15)
    torch.cuda.synchronize()
    got = host(T)
    assert same_bits(got, t_r), bit_diff_report(got, t_r)
    water = mat == 0
    assert got[water].max() > 295.2


def test_too_many_materials_is_an_error(simcore, torch_gpu):
    torch = torch_gpu
    k = torch.ones(17, device="cuda")
    t = torch.zeros((4, 4, 4), device="cuda")
    m = torch.zeros((4, 4, 4), dtype=torch.int32, device="cuda")
    with pytest.raises(RuntimeError):
        simcore.conduction_sweep(t.data_ptr(), t.data_ptr() + 0, t.data_ptr(), m.data_ptr(), k.data_ptr(),
                                 k.data_ptr(), k.data_ptr(), 17, 4, 4, 4, 1e-3, 0.1, H_CONV, T_AMB)


@pytest.mark.parametrize("shape,n_iter", [((96, 96, 64), 15), ((64, 440, 280), 9), ((37, 29, 44), 16),
                                          ((9, 7, 8), 15)])
@pytest.mark.parametrize("geometry", ["pot", "random"])
def test_driver_bit_exact_larger_grids(simcore, torch_gpu, shape, n_iter, geometry):
    """All three buffers stay bit-identical to the reference's ping-pong on
    larger grids: the pot (uniform-material chunks take the short form) and
    random materials with fluid on the domain edges."""
    torch = torch_gpu
    from boilsim.fixtures import pot_mask

    P, (k, rho, cp) = _props(torch, "aluminum")
    if geometry == "pot":
        mat = pot_mask(*shape)
        rng = np.random.default_rng(2)
        t0 = (295.0 + 200.0 * rng.random(shape)).astype(np.float32)
        dx, dt = 1e-3, 0.2
    else:
        mat, t0, _, dx, dt = _case(31, shape)
    alt0 = np.full(shape, -3.0, np.float32)
    t_r, alt_r, old_r = t0.copy(), alt0.copy(), np.zeros(shape, np.float32)
    oracle.conduction_solve(t_r, alt_r, old_r, mat, P.k, P.rho, P.c_p, dx, dt, H_CONV, T_AMB, n_iter)
    t, alt, old = dev(torch, t0), dev(torch, alt0), torch.zeros(shape, dtype=torch.float32, device="cuda")
    md = dev(torch, mat)
    simcore.conduction_solve(t.data_ptr(), alt.data_ptr(), old.data_ptr(), md.data_ptr(), k.data_ptr(),
                             rho.data_ptr(), cp.data_ptr(), k.numel(), *shape, dx, dt, H_CONV, T_AMB, n_iter)
    torch.cuda.synchronize()
    for g, r in ((host(t), t_r), (host(alt), alt_r), (host(old), old_r)):
        assert same_bits(g, r), bit_diff_report(g, r)
