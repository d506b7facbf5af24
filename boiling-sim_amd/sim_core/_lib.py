"""ctypes binding of libboilsim_hip.so, the gfx950 pressure-projection library.

Every entry point takes raw device addresses (Python ints) and returns a
hipError_t; :func:`check` turns a nonzero code into ``RuntimeError`` the way the
reference maps ``cudaError_t`` to ``anyhow`` to ``PyRuntimeError``
(crates/cuda-kernels/src/lib.rs:267-269, crates/sim-core/src/cuda.rs:123).
"""

from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, Structure, c_char_p, c_double, c_float, c_int, c_longlong, c_size_t, c_void_p

LIB_NAME = "libboilsim_hip.so"
LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), LIB_NAME)

_lib = None


class SolveStats(Structure):
    """bs_solve_stats (include/boilsim_hip.h)."""

    _fields_ = [
        ("iterations", c_int),
        ("converged", c_int),
        ("restarts", c_int),
        ("reserved", c_int),
        ("relres_recursive", c_double),
        ("relres_true", c_double),
        ("bnorm", c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}



class FcgStats(Structure):
    """bs_fcg_stats (include/boilsim_hip.h)."""

    _fields_ = [
        ("iterations", c_int),
        ("converged", c_int),
        ("fallback", c_int),
        ("pcg_iterations", c_int),
        ("restarts", c_int),
        ("reserved", c_int),
        ("relres_recursive", c_double),
        ("relres_true", c_double),
        ("bnorm", c_double),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class RefineStats(Structure):
    """bs_refine_stats (include/boilsim_hip.h)."""

    _fields_ = [
        ("outer", c_int),
        ("inner_iterations", c_int),
        ("converged", c_int),
        ("storage", c_int),
        ("relres_true", c_double),
        ("bnorm", c_double),
        ("relres_inner_last", c_double),
        ("fallback", c_int),
        ("pcg_iterations", c_int),
    ]

    def as_dict(self) -> dict:
        return {k: getattr(self, k) for k, _ in self._fields_}


VP = c_void_p  # device pointer
I = c_int
F = c_float
S = c_void_p  # hipStream_t

# name -> argtypes (restype is always c_int = hipError_t unless listed in _RESTYPES)
SIGNATURES = {
    "bs_abi_version": [],
    "bs_error_string": [I],
    "bs_device_open": [I, c_char_p, I],
    "bs_device_synchronize": [],
    "bs_scale_inplace_launch": [VP, I, F],
    "bs_jacobi_pressure_launch": [VP, VP, VP, VP, I, I, I, F, F, F, I, I],
    "bs_jacobi_pressure_launch_stream": [VP, VP, VP, VP, I, I, I, F, F, F, I, I, S],
    "bs_pressure_solve_launch": [VP, VP, VP, VP, I, I, I, F, F, F, I, I, I],
    "bs_jacobi_solve_launch": [VP, VP, VP, VP, I, I, I, F, F, F, I, I, I, I, S],
    "bs_div6_check": [VP, S],
    "bs_divergence_launch": [VP, VP, VP, VP, VP, I, I, I, F, I, S],
    "bs_subtract_gradient_launch": [VP, VP, VP, VP, VP, I, I, I, F, F, F, I, S],
    "bs_max_abs_launch": [VP, c_longlong, VP, S],
    "bs_projection_jacobi_launch": [VP, VP, VP, VP, VP, VP, VP, I, I, I, F, F, F, I, I, I, VP, S],
    "bs_enforce_no_slip_launch": [VP, VP, VP, VP, I, I, I, I, S],
    "bs_extend_temperature_launch": [VP, VP, VP, I, I, I, I, S],
    "bs_advect_launch": [VP] * 10 + [I, I, I, F, F, F, F, F, I, S],
    "bs_advect_all_launch": [VP] * 10 + [I, I, I, F, F, F, F, F, I, S],
    "bs_divergence_no_slip_launch": [VP, VP, VP, VP, VP, I, I, I, F, I, S],
    "bs_divergence_max_launch": [VP, VP, VP, VP, VP, I, I, I, F, I, VP, S],
    "bs_max_velocity_launch": [VP, VP, VP, I, I, I, VP, S],
    "bs_projection_step_jacobi_launch": [VP, VP, VP, VP, VP, VP, VP, I, I, I, F, F, F, I, I, I, VP, S],
    "bs_pressure_solve_pcg_launch": [VP, VP, VP, I, I, I, F, F, F, I, I, F, I] + [VP] * 13 + [POINTER(c_int)],
    "bs_pcg_workspace_bytes": [I, I, I],
    "bs_pcg_solve": [VP, VP, VP, I, I, I, F, F, F, I, I, c_double, I, I, I, VP, c_size_t, S, c_void_p],
    "bs_pcg_noalloc_scratch_bytes": [],
    "bs_pcg_solve_noalloc": [VP, VP, VP, I, I, I, F, F, F, I, I, c_double, I, I, I, S, c_void_p],
    "bs_laplacian_spmv_launch": [VP, VP, VP, I, I, I, I, I],
    "bs_diag_inverse_apply_launch": [VP, VP, VP, I, I, I, I, I],
    "bs_dot_launch": [VP, VP, I, VP, VP],
    "bs_dot_launch_f64": [VP, VP, I, VP, VP, VP],
    "bs_axpy_device_launch": [VP, VP, VP, I, I],
    "bs_scaled_axpy_launch": [VP, VP, VP, VP, I],
    "bs_divide_scalars_launch": [VP, VP, VP],
    "bs_compute_b_launch": [VP, VP, VP, I, I, I, F, I],
    "bs_zero_buffer_launch": [VP, I],
    "bs_unet_num_params": [],
    "bs_unet_packed_bytes": [],
    "bs_unet_workspace_bytes": [I, I, I, I],
    "bs_unet_pack": [VP, VP, S],
    "bs_build_codes": [VP, I, I, I, I, I, VP, S],
    "bs_unet_forward": [VP, VP, VP, VP, I, I, I, c_double, VP, VP, c_size_t, I, S],
    "bs_unet_backward": [VP, VP, VP, VP, VP, I, I, I, VP, VP, VP, c_size_t, S],
    "bs_unet_adam": [VP, VP, VP, VP, I, F, F, S],
    "bs_fcg_workspace_bytes": [I, I, I, I],
    "bs_mg_smooth_restrict": [VP, VP, VP, VP, I, I, I, I, c_double, S],
    "bs_mg_chebyshev": [VP, VP, VP, VP, I, I, I, I, c_double, S],
    "bs_mg_prolong_post": [VP, VP, VP, VP, VP, VP, VP, VP, c_size_t, POINTER(c_int), I, I, I, I, c_double, S],
    "bs_fcg_solve": [VP, VP, VP, I, I, I, F, F, F, I, I, c_double, I, I, I, VP, VP, VP, c_size_t, S, c_void_p],
    "bs_fcg_pass_timing": [I, c_void_p, c_void_p, c_void_p],
    "bs_conduction_sweep_launch": [VP, VP, VP, VP, VP, VP, VP, I, I, I, I, F, F, F, F, I, I, S],
    "bs_conduction_solve_launch": [VP, VP, VP, VP, VP, VP, VP, I, I, I, I, F, F, F, F, I, I, I, S],
    "bs_refine_workspace_bytes": [I, I, I],
    "bs_fcg_refine_workspace_bytes": [I, I, I, I],
    "bs_fcg_refine_solve": [VP, VP, VP, VP, I, I, I, F, F, F, I, I, c_double, I, c_double, I, I, I, VP, VP, VP,
                            c_size_t, S, c_void_p],
    "bs_refine_solve": [VP, VP, VP, VP, I, I, I, F, F, F, I, I, c_double, I, I, c_double, I, I, VP, c_size_t, S,
                        c_void_p],
}

_RESTYPES = {
    "bs_error_string": c_char_p,
    "bs_pcg_workspace_bytes": c_size_t,
    "bs_pcg_noalloc_scratch_bytes": c_size_t,
    "bs_unet_num_params": c_size_t,
    "bs_unet_packed_bytes": c_size_t,
    "bs_unet_workspace_bytes": c_size_t,
    "bs_fcg_workspace_bytes": c_size_t,
    "bs_refine_workspace_bytes": c_size_t,
    "bs_fcg_refine_workspace_bytes": c_size_t,
}


def load():
    """Load the library once.  Raises RuntimeError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `python -c \"import __graft_entry__ as g; g.build()\"` "
            "or `make -C boiling-sim_amd/csrc`."
        )
    lib = ctypes.CDLL(LIB_PATH, mode=ctypes.RTLD_GLOBAL)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = _RESTYPES.get(name, c_int)
    _lib = lib
    return lib


def exported_symbols() -> list[str]:
    return sorted(SIGNATURES)


def error_string(code: int) -> str:
    try:
        s = load().bs_error_string(int(code))
        return s.decode() if s else "unknown"
    except Exception:  # pragma: no cover - only on a broken install
        return "unknown"


def check(name: str, rc: int) -> None:
    if rc != 0:
        raise RuntimeError(f"{name} returned hipError_t {rc}: {error_string(rc)}")


def call(name: str, *args):
    """Call entry point `name` and raise RuntimeError on a nonzero hipError_t."""
    rc = getattr(load(), name)(*args)
    check(name, rc)
    return rc
