"""ORACLE — test infrastructure only.

ctypes front end of ``liborc.so`` (oracle/cpu_ref.c), the CPU restatement of
the reference's pressure-projection arithmetic, plus :mod:`oracle.np_ref`, an
independent numpy restatement used to cross-check it.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg import this
package, and only as the checker or the timed CPU baseline — never as part of
the product path.

Parity pinning (what fixes each function, since the reference ships no golden
vectors and Warp cannot be imported here):

* ``jacobi_step`` / ``pressure_solve`` — bit-exact agreement with the
  independent numpy restatement; known answers p=i^2 (Ap = -2) and p linear
  (Ap = 0) from test_pressure_cg.py:62-70; the fixed point of a converged sweep
  solves A p = b (checked against scipy's sparse solve in FP64); odd/even
  canonicalisation (N and N+1 sweeps differ by exactly one sweep).
* ``divergence`` / ``gradient`` — closed forms (linear velocity -> constant
  divergence, discrete gradient of a linear p), the fluid-face skip rules,
  and brute-force Python loops on tiny grids.
* ``laplacian_f64`` / ``pcg_f64`` / ``pcg_ref`` / ``residual_f64`` — scipy
  sparse matrix assembled independently from the class rules, scipy spsolve,
  and the reference's tests' gates (test_pressure_cg.py:49-434).
* ``enforce_no_slip`` / ``max_velocity`` / ``projection_step`` (the
  no-slip faces of fluid.py:111-162 and compute_max_velocity, fluid.py:796-833)
  — bit-exact agreement with the independent numpy restatement
  (``np_ref.enforce_no_slip``) and with per-face brute-force loops on tiny
  grids, the reference's own check (test_fluid.py:94-117: every face touching a
  wall cell is zero), the untouched interior fluid-fluid faces, idempotence,
  -0.0 becoming +0.0 as the reference's store of 0.0 makes it, and the
  hydrostatic closed box of test_fluid.py:199-217 (zero velocity stays zero).
* ``extend_temperature`` / ``advect`` (oracle/advect_ref.c, the
  semi-Lagrangian advection of fluid.py:33-103, 338-492, 737-775) —
  bit-exact agreement with the independent vectorised numpy restatement
  (``np_ref.advect``); per-cell brute force of the extension's neighbour
  order; zero velocity on a power-of-two grid returns the fields unchanged;
  a uniform velocity moving a whole number of cells shifts them exactly; a
  linear field under a uniform sub-cell velocity is reproduced to rounding
  (trilinear interpolation is exact on linear fields); the reference's
  20-step temperature-drift gate (test_fluid.py:147-184).
* ``conduction_sweep`` / ``conduction_solve`` (oracle/thermal_ref.c, the
  backward-Euler Jacobi conduction of thermal.py:259-388 and 631-665) —
  closed forms (one solid cell cooled through one air face; a water/steel
  cell pair with the harmonic face conductivity; the uniform field as a fixed
  point), the converged sweep against an independently assembled FP64
  backward-Euler system solved by scipy, exact energy conservation of that
  solve in a closed box, and the discrete amplification factor of a cosine
  mode (test_thermal.py:135-296 restated for the implicit scheme).
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from ctypes import POINTER, c_double, c_float, c_int, c_int64, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")
_lib = None

MAT_FLUID, MAT_AIR = 0, 2


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = c_void_p
        I = c_int
        F = c_float
        L.orc_set_threads.argtypes = [I]
        L.orc_get_threads.restype = I
        L.orc_divergence.argtypes = [P, P, P, P, P, I, I, I, F, I]
        L.orc_jacobi_step.argtypes = [P, P, P, P, I, I, I, F, F, F, I, I]
        L.orc_pressure_solve.argtypes = [P, P, P, P, I, I, I, F, F, F, I, I, I]
        L.orc_conduction_sweep.argtypes = [P, P, P, P, P, P, P, I, I, I, F, F, F, F, I]
        L.orc_conduction_solve.argtypes = [P, P, P, P, P, P, P, I, I, I, F, F, F, F, I, I]
        L.orc_gradient.argtypes = [P, P, P, P, P, I, I, I, F, F, F, I]
        L.orc_max_abs.argtypes = [P, c_int64]
        L.orc_max_abs.restype = c_double
        L.orc_projection.argtypes = [P, P, P, P, P, P, P, I, I, I, F, F, F, I, I, I]
        L.orc_projection.restype = c_double
        L.orc_enforce_no_slip.argtypes = [P, P, P, P, I, I, I, I]
        L.orc_extend_temperature.argtypes = [P, P, P, I, I, I, I]
        L.orc_advect.argtypes = [P, P, P, P, P, P, P, P, P, P, I, I, I, F, F, F, F, F, I]
        L.orc_max_velocity.argtypes = [P, P, P, I, I, I]
        L.orc_max_velocity.restype = c_double
        L.orc_laplacian_f64.argtypes = [P, P, P, I, I, I, I, I]
        L.orc_diag.argtypes = [P, P, I, I, I, I, I]
        L.orc_rhs_f64.argtypes = [P, P, P, I, I, I, F, F, F, I]
        L.orc_residual_f64.argtypes = [P, P, P, I, I, I, I, I, P, P]
        L.orc_pcg_f64.argtypes = [P, P, P, I, I, I, c_double, I, I, I]
        L.orc_pcg_f64.restype = I
        L.orc_pcg_ref.argtypes = [P, P, P, I, I, I, F, F, F, F, I, I, I, P]
        L.orc_pcg_ref.restype = I
        L.orc_pcg_true.argtypes = [P, P, P, I, I, I, F, F, F, c_double, I, I, I, I, P, P]
        L.orc_pcg_true.restype = I
        _lib = L
    return _lib


def _p(a: np.ndarray):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


def get_threads() -> int:
    return int(lib().orc_get_threads())


def divergence(ux, uy, uz, mat, dx, mat_fluid=MAT_FLUID):
    mat = _i32(mat)
    nx, ny, nz = mat.shape
    ux, uy, uz = _f32(ux), _f32(uy), _f32(uz)
    assert ux.shape == (nx + 1, ny, nz) and uy.shape == (nx, ny + 1, nz) and uz.shape == (nx, ny, nz + 1)
    div = np.empty(mat.shape, np.float32)
    lib().orc_divergence(_p(div), _p(ux), _p(uy), _p(uz), _p(mat), nx, ny, nz, float(dx), mat_fluid)
    return div


def jacobi_step(p_old, div, mat, dx, dt, rho, mat_fluid=MAT_FLUID, mat_air=MAT_AIR):
    mat = _i32(mat)
    nx, ny, nz = mat.shape
    p_old, div = _f32(p_old), _f32(div)
    out = np.empty(mat.shape, np.float32)
    lib().orc_jacobi_step(_p(out), _p(p_old), _p(div), _p(mat), nx, ny, nz, float(dx), float(dt),
                          float(rho), mat_fluid, mat_air)
    return out


def pressure_solve(p, p_tmp, div, mat, dx, dt, rho, n_iter, mat_fluid=MAT_FLUID, mat_air=MAT_AIR):
    """In place on p and p_tmp (float32 C-contiguous arrays), like the reference."""
    mat = _i32(mat)
    nx, ny, nz = mat.shape
    div = _f32(div)
    assert p.dtype == np.float32 and p_tmp.dtype == np.float32
    lib().orc_pressure_solve(_p(p), _p(p_tmp), _p(div), _p(mat), nx, ny, nz, float(dx), float(dt),
                             float(rho), int(n_iter), mat_fluid, mat_air)


def gradient(ux, uy, uz, p, mat, dx, dt, rho, mat_fluid=MAT_FLUID):
    """In place on ux, uy, uz."""
    mat = _i32(mat)
    nx, ny, nz = mat.shape
    p = _f32(p)
    lib().orc_gradient(_p(ux), _p(uy), _p(uz), _p(p), _p(mat), nx, ny, nz, float(dx), float(dt),
                       float(rho), mat_fluid)


def projection(ux, uy, uz, mat, dx, dt, rho, n_iter, mat_fluid=MAT_FLUID, mat_air=MAT_AIR):
    """Full Jacobi projection, in place on ux/uy/uz.  Returns (p, p_tmp, div, max_div)."""
    mat = _i32(mat)
    nx, ny, nz = mat.shape
    p = np.empty(mat.shape, np.float32)
    p_tmp = np.empty(mat.shape, np.float32)
    div = np.empty(mat.shape, np.float32)
    m = lib().orc_projection(_p(p), _p(p_tmp), _p(div), _p(ux), _p(uy), _p(uz), _p(mat), nx, ny, nz,
                             float(dx), float(dt), float(rho), int(n_iter), mat_fluid, mat_air)
    return p, p_tmp, div, float(m)


def enforce_no_slip(ux, uy, uz, mat, mat_fluid=MAT_FLUID):
    """In place on ux, uy, uz (float32 C-contiguous face arrays)."""
    mat = _i32(mat)
    nx, ny, nz = mat.shape
    for a, shp in ((ux, (nx + 1, ny, nz)), (uy, (nx, ny + 1, nz)), (uz, (nx, ny, nz + 1))):
        assert a.dtype == np.float32 and a.flags.c_contiguous and a.shape == shp
    lib().orc_enforce_no_slip(_p(ux), _p(uy), _p(uz), _p(mat), nx, ny, nz, mat_fluid)


def max_velocity(ux, uy, uz):
    ux, uy, uz = _f32(ux), _f32(uy), _f32(uz)
    nx, ny, nz = uy.shape[0], ux.shape[1], ux.shape[2]
    return float(lib().orc_max_velocity(_p(ux), _p(uy), _p(uz), nx, ny, nz))


def projection_step(ux, uy, uz, mat, dx, dt, rho, n_iter, mat_fluid=MAT_FLUID, mat_air=MAT_AIR):
    """pipeline.py:606-620 with the Jacobi solver, in place on the faces:
    enforce_no_slip, pressure_projection, enforce_no_slip.  Returns
    (p, p_tmp, div, max_div, max_u) with max_u = compute_max_velocity of the
    final faces."""
    enforce_no_slip(ux, uy, uz, mat, mat_fluid)
    p, p_tmp, div, md = projection(ux, uy, uz, mat, dx, dt, rho, n_iter, mat_fluid, mat_air)
    enforce_no_slip(ux, uy, uz, mat, mat_fluid)
    return p, p_tmp, div, md, max_velocity(ux, uy, uz)


def extend_temperature(t, mat, mat_fluid=MAT_FLUID):
    mat = _i32(mat)
    nx, ny, nz = mat.shape
    t = _f32(t)
    out = np.empty(mat.shape, np.float32)
    lib().orc_extend_temperature(_p(out), _p(t), _p(mat), nx, ny, nz, mat_fluid)
    return out


def advect(ux, uy, uz, t, mat, origin, dx, dt, mat_fluid=MAT_FLUID):
    """advect_all (fluid.py:737-775): returns (ux_new, uy_new, uz_new, t_new, t_ext)."""
    mat = _i32(mat)
    nx, ny, nz = mat.shape
    ux, uy, uz, t = _f32(ux), _f32(uy), _f32(uz), _f32(t)
    assert ux.shape == (nx + 1, ny, nz) and uy.shape == (nx, ny + 1, nz) and uz.shape == (nx, ny, nz + 1)
    t_ext = extend_temperature(t, mat, mat_fluid)
    out = [np.empty_like(a) for a in (ux, uy, uz, t)]
    lib().orc_advect(*(_p(a) for a in out), _p(ux), _p(uy), _p(uz), _p(t), _p(t_ext), _p(mat), nx, ny, nz,
                     *(float(c) for c in origin), float(dx), float(dt), mat_fluid)
    return (*out, t_ext)


def laplacian_f64(x, mat, mat_fluid=MAT_FLUID, mat_air=MAT_AIR):
    mat = _i32(mat)
    nx, ny, nz = mat.shape
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.empty(mat.shape, np.float64)
    lib().orc_laplacian_f64(_p(y), _p(x), _p(mat), nx, ny, nz, mat_fluid, mat_air)
    return y


def diag(mat, mat_fluid=MAT_FLUID, mat_air=MAT_AIR):
    mat = _i32(mat)
    nx, ny, nz = mat.shape
    d = np.empty(mat.shape, np.float64)
    lib().orc_diag(_p(d), _p(mat), nx, ny, nz, mat_fluid, mat_air)
    return d


def rhs_f64(div, mat, dx, dt, rho, mat_fluid=MAT_FLUID):
    mat = _i32(mat)
    nx, ny, nz = mat.shape
    div = _f32(div)
    b = np.empty(mat.shape, np.float64)
    lib().orc_rhs_f64(_p(b), _p(div), _p(mat), nx, ny, nz, float(dx), float(dt), float(rho), mat_fluid)
    return b


def residual_f64(x, b, mat, mat_fluid=MAT_FLUID, mat_air=MAT_AIR):
    """(||b - A x||_2/||b||_2, ||b - A x||_inf/||b||_inf) in FP64."""
    mat = _i32(mat)
    nx, ny, nz = mat.shape
    x = np.ascontiguousarray(x, dtype=np.float64)
    b = np.ascontiguousarray(b, dtype=np.float64)
    r2 = c_double()
    ri = c_double()
    lib().orc_residual_f64(_p(x), _p(b), _p(mat), nx, ny, nz, mat_fluid, mat_air,
                           ctypes.addressof(r2), ctypes.addressof(ri))
    return r2.value, ri.value


def pcg_f64(b, mat, tol=1e-12, max_iter=100000, x0=None, mat_fluid=MAT_FLUID, mat_air=MAT_AIR):
    """FP64 Jacobi-PCG oracle. Returns (x, iterations)."""
    mat = _i32(mat)
    nx, ny, nz = mat.shape
    b = np.ascontiguousarray(b, dtype=np.float64)
    x = np.zeros(mat.shape, np.float64) if x0 is None else np.array(x0, dtype=np.float64, order="C")
    it = lib().orc_pcg_f64(_p(x), _p(b), _p(mat), nx, ny, nz, float(tol), int(max_iter), mat_fluid, mat_air)
    return x, int(it)


def pcg_true(div, mat, dx, dt, rho, tol, max_iter, max_restarts=4, mat_fluid=MAT_FLUID, mat_air=MAT_AIR):
    """Jacobi-PCG (FP32 vectors, FP64 scalars) to a true relative residual
    tol with residual replacement, the semantics of bs_pcg_solve.  Returns
    (x, iterations, restarts, true relative residual)."""
    mat = _i32(mat)
    nx, ny, nz = mat.shape
    div = _f32(div)
    x = np.empty(mat.shape, np.float32)
    rel = c_double()
    rst = ctypes.c_int()
    it = lib().orc_pcg_true(_p(x), _p(div), _p(mat), nx, ny, nz, float(dx), float(dt), float(rho), float(tol),
                            int(max_iter), int(max_restarts), mat_fluid, mat_air, ctypes.addressof(rel),
                            ctypes.addressof(rst))
    if it < 0:
        raise MemoryError("orc_pcg_true: allocation failed")
    return x, int(it), int(rst.value), rel.value


def pcg_ref(div, mat, dx, dt, rho, tol, max_iter, mat_fluid=MAT_FLUID, mat_air=MAT_AIR):
    """Reference-semantics Jacobi-PCG (FP32 vectors, FP64 scalars).
    Returns (p, iterations, bsq)."""
    mat = _i32(mat)
    nx, ny, nz = mat.shape
    div = _f32(div)
    p = np.empty(mat.shape, np.float32)
    bsq = c_double()
    it = lib().orc_pcg_ref(_p(p), _p(div), _p(mat), nx, ny, nz, float(dx), float(dt), float(rho),
                           float(tol), int(max_iter), mat_fluid, mat_air, ctypes.addressof(bsq))
    return p, int(it), bsq.value


def conduction_sweep(t_dst, t_iter, t_old, mat, k, rho, cp, dx, dt, h_conv_air, t_amb, mat_air=MAT_AIR):
    """One backward-Euler Jacobi sweep into t_dst (float32, in place; air
    cells of t_dst are left as they are)."""
    mat = _i32(mat)
    nx, ny, nz = mat.shape
    t_iter, t_old = _f32(t_iter), _f32(t_old)
    k, rho, cp = _f32(k), _f32(rho), _f32(cp)
    assert t_dst.dtype == np.float32 and t_dst.flags.c_contiguous
    lib().orc_conduction_sweep(_p(t_dst), _p(t_iter), _p(t_old), _p(mat), _p(k), _p(rho), _p(cp), nx, ny, nz,
                               float(dx), float(dt), float(h_conv_air), float(t_amb), mat_air)


def conduction_solve(t, t_alt, t_old, mat, k, rho, cp, dx, dt, h_conv_air, t_amb, n_iter, mat_air=MAT_AIR):
    """conduct_one_step's backward-Euler branch, in place on t, t_alt, t_old."""
    mat = _i32(mat)
    nx, ny, nz = mat.shape
    k, rho, cp = _f32(k), _f32(rho), _f32(cp)
    for a in (t, t_alt, t_old):
        assert a.dtype == np.float32 and a.flags.c_contiguous
    lib().orc_conduction_solve(_p(t), _p(t_alt), _p(t_old), _p(mat), _p(k), _p(rho), _p(cp), nx, ny, nz,
                               float(dx), float(dt), float(h_conv_air), float(t_amb), int(n_iter), mat_air)
