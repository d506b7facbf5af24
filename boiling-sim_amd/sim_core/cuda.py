"""``sim_core.cuda.SimCore`` on MI355X: same Python signatures as the reference's
PyO3 class (crates/sim-core/src/cuda.rs:17-412), backed by libboilsim_hip.so.

Pointer contract (crates/sim-core/src/cuda.rs:52-60, python/boilingsim/fluid.py:
646-651): every ``*_ptr`` is a raw device address as a Python int — for example
``tensor.data_ptr()`` of a ``torch`` tensor on the same device — that points to
at least ``nx*ny*nz`` elements of the right type.  The caller owns every
buffer; no method allocates device memory.  Work is queued on the null stream,
asynchronously, except ``scale_array`` (synchronous, scale.cu:22-34) and
``pressure_solve_pcg`` (returns the iteration count, so it completes before
returning).
"""

from __future__ import annotations

import ctypes

from . import _lib

_I32_MAX = 2**31 - 1


def _usize(name: str, v) -> int:
    """PyO3 ``usize`` extraction followed by the reference's ``try_into::<i32>``
    check (crates/cuda-kernels/src/lib.rs:250-252)."""
    v = int(v)
    if v < 0:
        raise OverflowError(f"can't convert negative int to unsigned ({name}={v})")
    if v > _I32_MAX:
        raise RuntimeError(f"dim {v} overflows i32")
    return v


def _ptr(name: str, v) -> int:
    v = int(v)
    if v < 0 or v >= 2**64:
        raise OverflowError(f"{name}={v} is not a valid u64 device pointer")
    return v


def _i32(name: str, v) -> int:
    v = int(v)
    if not (-(2**31) <= v <= _I32_MAX):
        raise OverflowError(f"{name}={v} does not fit i32")
    return v


def _fused(v) -> int:
    """Kernel-shape selector of the V-cycle diagnostic entries: False/0 the
    split kernels, True/1 the single pass on 256-thread tiles, 2 the single
    pass on the fine level's taller tiles."""
    if v is True or v is False:
        return int(v)
    v = int(v)
    if v not in (0, 1, 2):
        raise ValueError(f"fused={v} is not 0, 1 or 2")
    return v


class SimCore:
    """Handle on one HIP device (reference: ``SimCore(device_id)``)."""

    def __init__(self, device_id: int = 0):
        device_id = _usize("device_id", device_id)
        buf = ctypes.create_string_buffer(256)
        rc = _lib.load().bs_device_open(device_id, buf, len(buf))
        if rc != 0:
            raise RuntimeError(
                f"HIP device {device_id} unavailable: hipError_t {rc} ({_lib.error_string(rc)}). "
                "If you are on a CPU-only machine, sim_core.cuda cannot be used."
            )
        self.device_id = device_id
        self._info = buf.value.decode()

    def device_info(self) -> str:
        return self._info

    def __repr__(self) -> str:
        return f"SimCore(device_id={self.device_id})"

    # -- FFI smoke (cuda.rs:51-68) -------------------------------------------------
    def scale_array(self, device_ptr: int, n: int, scale: float) -> None:
        n = _usize("n", n)
        if n == 0:
            return
        _lib.call("bs_scale_inplace_launch", _ptr("device_ptr", device_ptr), n, float(scale))

    # -- one Jacobi sweep (cuda.rs:83-126) ----------------------------------------
    def jacobi_pressure_step(self, p_new_ptr, p_old_ptr, div_u_ptr, mat_ptr, nx, ny, nz,
                             dx, dt, rho, mat_fluid=0, mat_air=2) -> None:
        _lib.call(
            "bs_jacobi_pressure_launch",
            _ptr("p_new_ptr", p_new_ptr), _ptr("p_old_ptr", p_old_ptr),
            _ptr("div_u_ptr", div_u_ptr), _ptr("mat_ptr", mat_ptr),
            _usize("nx", nx), _usize("ny", ny), _usize("nz", nz),
            float(dx), float(dt), float(rho), _i32("mat_fluid", mat_fluid), _i32("mat_air", mat_air),
        )

    # -- fused n_iter Jacobi solve (cuda.rs:136-171) -------------------------------
    def pressure_solve(self, p_ptr, p_tmp_ptr, div_u_ptr, mat_ptr, nx, ny, nz, dx, dt, rho,
                       n_iter, mat_fluid=0, mat_air=2) -> None:
        _lib.call(
            "bs_pressure_solve_launch",
            _ptr("p_ptr", p_ptr), _ptr("p_tmp_ptr", p_tmp_ptr),
            _ptr("div_u_ptr", div_u_ptr), _ptr("mat_ptr", mat_ptr),
            _usize("nx", nx), _usize("ny", ny), _usize("nz", nz),
            float(dx), float(dt), float(rho), _usize("n_iter", n_iter),
            _i32("mat_fluid", mat_fluid), _i32("mat_air", mat_air),
        )

    # -- extensions: the Warp-side kernels of pressure_projection ------------------
    def compute_divergence(self, div_u_ptr, ux_ptr, uy_ptr, uz_ptr, mat_ptr, nx, ny, nz, dx,
                           mat_fluid=0, stream=0) -> None:
        """compute_divergence (python/boilingsim/fluid.py:201-219)."""
        _lib.call(
            "bs_divergence_launch",
            _ptr("div_u_ptr", div_u_ptr), _ptr("ux_ptr", ux_ptr), _ptr("uy_ptr", uy_ptr),
            _ptr("uz_ptr", uz_ptr), _ptr("mat_ptr", mat_ptr),
            _usize("nx", nx), _usize("ny", ny), _usize("nz", nz), float(dx),
            _i32("mat_fluid", mat_fluid), _ptr("stream", stream),
        )

    def subtract_pressure_gradient(self, ux_ptr, uy_ptr, uz_ptr, p_ptr, mat_ptr, nx, ny, nz,
                                   dx, dt, rho, mat_fluid=0, stream=0) -> None:
        """subtract_pressure_gradient_{x,y,z} (fluid.py:279-330) in one pass."""
        _lib.call(
            "bs_subtract_gradient_launch",
            _ptr("ux_ptr", ux_ptr), _ptr("uy_ptr", uy_ptr), _ptr("uz_ptr", uz_ptr),
            _ptr("p_ptr", p_ptr), _ptr("mat_ptr", mat_ptr),
            _usize("nx", nx), _usize("ny", ny), _usize("nz", nz),
            float(dx), float(dt), float(rho), _i32("mat_fluid", mat_fluid), _ptr("stream", stream),
        )

    def max_abs(self, x_ptr, n, out_bits_ptr, stream=0) -> None:
        """max|x| left on the device as float bits in a uint32 at out_bits_ptr."""
        _lib.call("bs_max_abs_launch", _ptr("x_ptr", x_ptr), int(n), _ptr("out_bits_ptr", out_bits_ptr),
                  _ptr("stream", stream))

    def projection_jacobi(self, p_ptr, p_tmp_ptr, div_u_ptr, ux_ptr, uy_ptr, uz_ptr, mat_ptr,
                          nx, ny, nz, dx, dt, rho, n_iter, max_div_bits_ptr, mat_fluid=0,
                          mat_air=2, stream=0) -> None:
        """The whole Jacobi pressure_projection (fluid.py:612-734) in one call;
        max|div| is left on the device (uint32 float bits at max_div_bits_ptr)."""
        _lib.call(
            "bs_projection_jacobi_launch",
            _ptr("p_ptr", p_ptr), _ptr("p_tmp_ptr", p_tmp_ptr), _ptr("div_u_ptr", div_u_ptr),
            _ptr("ux_ptr", ux_ptr), _ptr("uy_ptr", uy_ptr), _ptr("uz_ptr", uz_ptr),
            _ptr("mat_ptr", mat_ptr), _usize("nx", nx), _usize("ny", ny), _usize("nz", nz),
            float(dx), float(dt), float(rho), _usize("n_iter", n_iter),
            _i32("mat_fluid", mat_fluid), _i32("mat_air", mat_air),
            _ptr("max_div_bits_ptr", max_div_bits_ptr), _ptr("stream", stream),
        )

    # -- no-slip faces and the CFL velocity maximum (fluid.py:111-162, 796-844) ----
    def enforce_no_slip(self, ux_ptr, uy_ptr, uz_ptr, mat_ptr, nx, ny, nz, mat_fluid=0, stream=0) -> None:
        """enforce_no_slip (fluid.py:111-162): the three face kernels in one launch."""
        _lib.call("bs_enforce_no_slip_launch", _ptr("ux_ptr", ux_ptr), _ptr("uy_ptr", uy_ptr),
                  _ptr("uz_ptr", uz_ptr), _ptr("mat_ptr", mat_ptr), _usize("nx", nx), _usize("ny", ny),
                  _usize("nz", nz), _i32("mat_fluid", mat_fluid), _ptr("stream", stream))

    def divergence_no_slip(self, div_u_ptr, ux_ptr, uy_ptr, uz_ptr, mat_ptr, nx, ny, nz, dx, mat_fluid=0,
                           stream=0) -> None:
        """enforce_no_slip, then compute_divergence of the result, in one pass."""
        _lib.call("bs_divergence_no_slip_launch", _ptr("div_u_ptr", div_u_ptr), _ptr("ux_ptr", ux_ptr),
                  _ptr("uy_ptr", uy_ptr), _ptr("uz_ptr", uz_ptr), _ptr("mat_ptr", mat_ptr), _usize("nx", nx),
                  _usize("ny", ny), _usize("nz", nz), float(dx), _i32("mat_fluid", mat_fluid),
                  _ptr("stream", stream))

    def divergence_max(self, div_u_ptr, ux_ptr, uy_ptr, uz_ptr, mat_ptr, nx, ny, nz, dx, max_bits_ptr,
                       mat_fluid=0, stream=0) -> None:
        """compute_divergence; max|div| and max|u| as float bits in the two
        uint32s at max_bits_ptr."""
        _lib.call("bs_divergence_max_launch", _ptr("div_u_ptr", div_u_ptr), _ptr("ux_ptr", ux_ptr),
                  _ptr("uy_ptr", uy_ptr), _ptr("uz_ptr", uz_ptr), _ptr("mat_ptr", mat_ptr), _usize("nx", nx),
                  _usize("ny", ny), _usize("nz", nz), float(dx), _i32("mat_fluid", mat_fluid),
                  _ptr("max_bits_ptr", max_bits_ptr), _ptr("stream", stream))

    def max_velocity(self, ux_ptr, uy_ptr, uz_ptr, nx, ny, nz, out_bits_ptr, stream=0) -> None:
        """compute_max_velocity's device reduction (fluid.py:807-833); max|u|
        as float bits in the uint32 at out_bits_ptr."""
        _lib.call("bs_max_velocity_launch", _ptr("ux_ptr", ux_ptr), _ptr("uy_ptr", uy_ptr),
                  _ptr("uz_ptr", uz_ptr), _usize("nx", nx), _usize("ny", ny), _usize("nz", nz),
                  _ptr("out_bits_ptr", out_bits_ptr), _ptr("stream", stream))

    def projection_step_jacobi(self, p_ptr, p_tmp_ptr, div_u_ptr, ux_ptr, uy_ptr, uz_ptr, mat_ptr, nx, ny, nz,
                               dx, dt, rho, n_iter, max_bits_ptr, mat_fluid=0, mat_air=2, stream=0) -> None:
        """no-slip, Jacobi pressure_projection, no-slip (pipeline.py:606-620)
        in four passes plus the sweeps; max|div| and max|u| as float bits in
        the two uint32s at max_bits_ptr."""
        _lib.call(
            "bs_projection_step_jacobi_launch",
            _ptr("p_ptr", p_ptr), _ptr("p_tmp_ptr", p_tmp_ptr), _ptr("div_u_ptr", div_u_ptr),
            _ptr("ux_ptr", ux_ptr), _ptr("uy_ptr", uy_ptr), _ptr("uz_ptr", uz_ptr),
            _ptr("mat_ptr", mat_ptr), _usize("nx", nx), _usize("ny", ny), _usize("nz", nz),
            float(dx), float(dt), float(rho), _usize("n_iter", n_iter),
            _i32("mat_fluid", mat_fluid), _i32("mat_air", mat_air),
            _ptr("max_bits_ptr", max_bits_ptr), _ptr("stream", stream),
        )

    # -- semi-Lagrangian advection (fluid.py:33-103, 338-492, 737-775) -------------
    def extend_temperature(self, t_ext_ptr, t_ptr, mat_ptr, nx, ny, nz, mat_fluid=0, stream=0) -> None:
        """extend_temperature_into_solids (fluid.py:338-378)."""
        _lib.call("bs_extend_temperature_launch", _ptr("t_ext_ptr", t_ext_ptr), _ptr("t_ptr", t_ptr),
                  _ptr("mat_ptr", mat_ptr), _usize("nx", nx), _usize("ny", ny), _usize("nz", nz),
                  _i32("mat_fluid", mat_fluid), _ptr("stream", stream))

    def advect(self, ux_new_ptr, uy_new_ptr, uz_new_ptr, t_new_ptr, ux_ptr, uy_ptr, uz_ptr, t_ptr, t_ext_ptr,
               mat_ptr, nx, ny, nz, origin, dx, dt, mat_fluid=0, stream=0) -> None:
        """advect_u{x,y,z} and advect_temperature (fluid.py:381-492) in one launch."""
        ox, oy, oz = (float(c) for c in origin)
        _lib.call("bs_advect_launch", *(_ptr(n, v) for n, v in (
            ("ux_new_ptr", ux_new_ptr), ("uy_new_ptr", uy_new_ptr), ("uz_new_ptr", uz_new_ptr),
            ("t_new_ptr", t_new_ptr), ("ux_ptr", ux_ptr), ("uy_ptr", uy_ptr), ("uz_ptr", uz_ptr), ("t_ptr", t_ptr),
            ("t_ext_ptr", t_ext_ptr), ("mat_ptr", mat_ptr))),
            _usize("nx", nx), _usize("ny", ny), _usize("nz", nz), ox, oy, oz, float(dx), float(dt),
            _i32("mat_fluid", mat_fluid), _ptr("stream", stream))

    def advect_all(self, ux_ptr, uy_ptr, uz_ptr, t_ptr, ux_tmp_ptr, uy_tmp_ptr, uz_tmp_ptr, t_tmp_ptr, t_ext_ptr,
                   mat_ptr, nx, ny, nz, origin, dx, dt, mat_fluid=0, stream=0) -> None:
        """advect_all (fluid.py:737-775): the advected fields end in ux, uy, uz, t."""
        ox, oy, oz = (float(c) for c in origin)
        _lib.call("bs_advect_all_launch", *(_ptr(n, v) for n, v in (
            ("ux_ptr", ux_ptr), ("uy_ptr", uy_ptr), ("uz_ptr", uz_ptr), ("t_ptr", t_ptr),
            ("ux_tmp_ptr", ux_tmp_ptr), ("uy_tmp_ptr", uy_tmp_ptr), ("uz_tmp_ptr", uz_tmp_ptr),
            ("t_tmp_ptr", t_tmp_ptr), ("t_ext_ptr", t_ext_ptr), ("mat_ptr", mat_ptr))),
            _usize("nx", nx), _usize("ny", ny), _usize("nz", nz), ox, oy, oz, float(dx), float(dt),
            _i32("mat_fluid", mat_fluid), _ptr("stream", stream))

    # -- PCG (cuda.rs:368-412) ------------------------------------------------------
    def pressure_solve_pcg(self, p_out_ptr, div_u_ptr, mat_ptr, nx, ny, nz, dx, dt, rho,
                           pressure_tol, max_iter, ws_b_ptr, ws_r_ptr, ws_z_ptr,
                           ws_p_search_ptr, ws_ap_ptr, dot_workspace_ptr, dev_alpha_ptr,
                           dev_beta_ptr, dev_rzold_ptr, dev_rznew_ptr, dev_bsq_ptr, dev_rsq_ptr,
                           dev_pap_ptr, mat_fluid=0, mat_air=2) -> int:
        """Jacobi-PCG with the reference's workspace layout and control flow;
        returns the iteration count.  FP64 dots replace the reference's FP32
        ones; ws_z carries the class codes and reduction partials."""
        iters = ctypes.c_int(0)
        _lib.call(
            "bs_pressure_solve_pcg_launch",
            _ptr("p_out_ptr", p_out_ptr), _ptr("div_u_ptr", div_u_ptr), _ptr("mat_ptr", mat_ptr),
            _usize("nx", nx), _usize("ny", ny), _usize("nz", nz),
            float(dx), float(dt), float(rho), _i32("mat_fluid", mat_fluid), _i32("mat_air", mat_air),
            float(pressure_tol), _usize("max_iter", max_iter),
            _ptr("ws_b_ptr", ws_b_ptr), _ptr("ws_r_ptr", ws_r_ptr), _ptr("ws_z_ptr", ws_z_ptr),
            _ptr("ws_p_search_ptr", ws_p_search_ptr), _ptr("ws_ap_ptr", ws_ap_ptr),
            _ptr("dot_workspace_ptr", dot_workspace_ptr), _ptr("dev_alpha_ptr", dev_alpha_ptr),
            _ptr("dev_beta_ptr", dev_beta_ptr), _ptr("dev_rzold_ptr", dev_rzold_ptr),
            _ptr("dev_rznew_ptr", dev_rznew_ptr), _ptr("dev_bsq_ptr", dev_bsq_ptr),
            _ptr("dev_rsq_ptr", dev_rsq_ptr), _ptr("dev_pap_ptr", dev_pap_ptr),
            ctypes.byref(iters),
        )
        return int(iters.value)

    @staticmethod
    def pcg_workspace_bytes(nx, ny, nz) -> int:
        return int(_lib.load().bs_pcg_workspace_bytes(_usize("nx", nx), _usize("ny", ny), _usize("nz", nz)))

    def pcg_solve(self, x_ptr, div_u_ptr, mat_ptr, nx, ny, nz, dx, dt, rho, tol, max_iter,
                  workspace_ptr, workspace_bytes, warm_start=False, max_restarts=4, mat_fluid=0,
                  mat_air=2, stream=0) -> dict:
        """Jacobi-PCG to a true relative residual ``tol`` (FP64 residual check,
        residual replacement).  Returns the solve statistics as a dict."""
        st = _lib.SolveStats()
        _lib.call(
            "bs_pcg_solve",
            _ptr("x_ptr", x_ptr), _ptr("div_u_ptr", div_u_ptr), _ptr("mat_ptr", mat_ptr),
            _usize("nx", nx), _usize("ny", ny), _usize("nz", nz),
            float(dx), float(dt), float(rho), _i32("mat_fluid", mat_fluid), _i32("mat_air", mat_air),
            float(tol), _usize("max_iter", max_iter), 1 if warm_start else 0, int(max_restarts),
            _ptr("workspace_ptr", workspace_ptr), int(workspace_bytes), _ptr("stream", stream),
            ctypes.byref(st),
        )
        return st.as_dict()

    @staticmethod
    def pcg_noalloc_scratch_bytes() -> int:
        return int(_lib.load().bs_pcg_noalloc_scratch_bytes())

    def pcg_solve_noalloc(self, x_ptr, div_u_ptr, mat_ptr, nx, ny, nz, dx, dt, rho, tol, max_iter,
                          warm_start=False, max_restarts=4, mat_fluid=0, mat_air=2, stream=0) -> dict:
        """``pcg_solve`` without a workspace and without any allocation: the
        library's fixed per-device scratch array is the workspace.  Grids whose
        workspace exceeds it raise ValueError; calls that overlap on one device
        race (they share that array)."""
        need = self.pcg_workspace_bytes(nx, ny, nz)
        have = self.pcg_noalloc_scratch_bytes()
        if need > have:
            raise ValueError(f"pcg_solve_noalloc: a {nx}x{ny}x{nz} grid needs {need} bytes of workspace; "
                             f"the fixed scratch array holds {have} bytes (use pcg_solve with a workspace)")
        st = _lib.SolveStats()
        _lib.call(
            "bs_pcg_solve_noalloc",
            _ptr("x_ptr", x_ptr), _ptr("div_u_ptr", div_u_ptr), _ptr("mat_ptr", mat_ptr),
            _usize("nx", nx), _usize("ny", ny), _usize("nz", nz),
            float(dx), float(dt), float(rho), _i32("mat_fluid", mat_fluid), _i32("mat_air", mat_air),
            float(tol), _usize("max_iter", max_iter), 1 if warm_start else 0, int(max_restarts),
            _ptr("stream", stream), ctypes.byref(st),
        )
        return st.as_dict()

    # -- flexible CG with the learned preconditioner (new) ------------------------
    @staticmethod
    def fcg_workspace_bytes(nx, ny, nz, precond) -> int:
        return int(_lib.load().bs_fcg_workspace_bytes(_usize("nx", nx), _usize("ny", ny), _usize("nz", nz),
                                                      int(precond)))

    def fcg_solve(self, x_ptr, div_u_ptr, mat_ptr, nx, ny, nz, dx, dt, rho, tol, max_iter, precond,
                  unet_params_ptr, unet_packed_ptr, workspace_ptr, workspace_bytes, warm_start=False,
                  mat_fluid=0, mat_air=2, stream=0) -> dict:
        """Flexible CG with preconditioner kind ``precond`` (1 U-Net, 2 V-cycle +
        U-Net coarse solve + Chebyshev polish, 3 the same V-cycle with a Jacobi
        sweep + the polish instead of the network, 4 Jacobi sweeps + U-Net on
        their residual, 5 the V-cycle with a Jacobi sweep + 60 Chebyshev steps
        on the coarsest level, 6 the V-cycle with two more levels and a bottom
        PCG) to a true relative residual ``tol``; falls back to plain PCG on a
        stall or a non-positive r.z.  Kinds 3, 5 and 6 take null network
        pointers."""
        st = _lib.FcgStats()
        _lib.call(
            "bs_fcg_solve",
            _ptr("x_ptr", x_ptr), _ptr("div_u_ptr", div_u_ptr), _ptr("mat_ptr", mat_ptr),
            _usize("nx", nx), _usize("ny", ny), _usize("nz", nz),
            float(dx), float(dt), float(rho), _i32("mat_fluid", mat_fluid), _i32("mat_air", mat_air),
            float(tol), _usize("max_iter", max_iter), 1 if warm_start else 0, int(precond),
            _ptr("unet_params_ptr", unet_params_ptr), _ptr("unet_packed_ptr", unet_packed_ptr),
            _ptr("workspace_ptr", workspace_ptr), int(workspace_bytes), _ptr("stream", stream),
            ctypes.byref(st),
        )
        return st.as_dict()

    # Fine-level passes of the learned flexible CG, in bs_fcg_pass_timing's order.
    TIMED_PASSES = ("fcg_spmv_dir_kernel", "pre2_restrict_kernel<true, 64>", "prolong_post_kernel<48>")

    @staticmethod
    def fcg_pass_timing_start() -> None:
        """Turn pass timing on (see bs_fcg_pass_timing): later flexible-CG
        solves launch eagerly and time their fine-level passes."""
        _lib.call("bs_fcg_pass_timing", 1, None, None, None)

    @classmethod
    def fcg_pass_timing_stop(cls) -> dict:
        """Turn pass timing off; {kernel: {"us_median", "us_mean", "launches"}}."""
        med = (ctypes.c_double * 3)()
        mean = (ctypes.c_double * 3)()
        cnt = (ctypes.c_int * 3)()
        _lib.call("bs_fcg_pass_timing", 0, ctypes.addressof(med), ctypes.addressof(mean), ctypes.addressof(cnt))
        return {k: {"us_median": med[i], "us_mean": mean[i], "launches": cnt[i]}
                for i, k in enumerate(cls.TIMED_PASSES)}

    def mg_smooth_restrict(self, code_ptr, r_ptr, z_ptr, rc_ptr, nx, ny, nz, fused=True, lo=0.0, stream=0) -> None:
        """The V-cycle's first pass on one level: z = two smoothing steps from
        zero (damped Jacobi for lo <= 0, Chebyshev over [lo, 2] otherwise), rc =
        0.5 x the children's sum of r - A z (see bs_mg_smooth_restrict); fused
        selects the single-pass kernel (True or 1: 256-thread tiles, 2: the fine
        level's taller tiles)."""
        _lib.call("bs_mg_smooth_restrict", _ptr("code_ptr", code_ptr), _ptr("r_ptr", r_ptr), _ptr("z_ptr", z_ptr),
                  _ptr("rc_ptr", rc_ptr), _usize("nx", nx), _usize("ny", ny), _usize("nz", nz), _fused(fused),
                  float(lo), _ptr("stream", stream))

    def mg_chebyshev(self, code_ptr, r_ptr, x_ptr, tmp_ptr, nx, ny, nz, steps, lo, stream=0) -> None:
        """`steps` Chebyshev steps over [lo, 2] on A x = r from and into x
        (see bs_mg_chebyshev)."""
        _lib.call("bs_mg_chebyshev", _ptr("code_ptr", code_ptr), _ptr("r_ptr", r_ptr), _ptr("x_ptr", x_ptr),
                  _ptr("tmp_ptr", tmp_ptr), _usize("nx", nx), _usize("ny", ny), _usize("nz", nz), int(steps),
                  float(lo), _ptr("stream", stream))

    def mg_prolong_post(self, code_ptr, z_in_ptr, ec_ptr, r_ptr, z_old_ptr, z_out_ptr, tmp_ptr, part_ptr, part_len,
                        nx, ny, nz, fused=True, w=1.0, stream=0) -> int:
        """The V-cycle's last fine-level pass (see bs_mg_prolong_post; fused as
        for mg_smooth_restrict); returns the number G of workgroup partials
        written per dot product."""
        g = ctypes.c_int(0)
        _lib.call("bs_mg_prolong_post", _ptr("code_ptr", code_ptr), _ptr("z_in_ptr", z_in_ptr),
                  _ptr("ec_ptr", ec_ptr), _ptr("r_ptr", r_ptr), z_old_ptr or None, _ptr("z_out_ptr", z_out_ptr),
                  _ptr("tmp_ptr", tmp_ptr), _ptr("part_ptr", part_ptr), int(part_len), ctypes.byref(g),
                  _usize("nx", nx), _usize("ny", ny), _usize("nz", nz), _fused(fused), float(w),
                  _ptr("stream", stream))
        return int(g.value)

    # ---- mixed-precision iterative refinement -------------------------------
    STORAGE = {"fp32": 0, "fp16": 1, "bf16": 2}
    # Residual reduction between replacements by storage type: 16-bit vectors
    # keep the recursive residual close to the true one only over about a
    # decade.
    INNER_TOL = {"fp32": 1e-2, "fp16": 1e-1, "bf16": 1e-1}

    @staticmethod
    def refine_workspace_bytes(nx, ny, nz) -> int:
        return int(_lib.load().bs_refine_workspace_bytes(_usize("nx", nx), _usize("ny", ny), _usize("nz", nz)))

    def refine_solve(self, x64_ptr, div_u_ptr, mat_ptr, nx, ny, nz, dx, dt, rho, tol, workspace_ptr,
                     workspace_bytes, storage="fp32", inner_tol=None, max_outer=40, max_inner=20000,
                     warm_start=False, x32_ptr=0, mat_fluid=0, mat_air=2, stream=0) -> dict:
        """FP64 x (``x64_ptr``, nx*ny*nz doubles) solving A x = b to a true
        relative residual ``tol`` by iterative refinement around a Jacobi-PCG
        whose vectors are stored as ``storage`` ("fp32", "fp16" or "bf16");
        the same A and b as ``pressure_solve_pcg``
        (crates/cuda-kernels/src/pressure_solve_pcg.cu:89-243)."""
        if storage not in self.STORAGE:
            raise ValueError(f"storage must be one of {sorted(self.STORAGE)}, got {storage!r}")
        if inner_tol is None:
            inner_tol = self.INNER_TOL[storage]
        st = _lib.RefineStats()
        _lib.call(
            "bs_refine_solve",
            _ptr("x64_ptr", x64_ptr), _ptr("x32_ptr", x32_ptr), _ptr("div_u_ptr", div_u_ptr), _ptr("mat_ptr", mat_ptr),
            _usize("nx", nx), _usize("ny", ny), _usize("nz", nz),
            float(dx), float(dt), float(rho), _i32("mat_fluid", mat_fluid), _i32("mat_air", mat_air),
            float(tol), _usize("max_outer", max_outer), self.STORAGE[storage], float(inner_tol),
            _usize("max_inner", max_inner), 1 if warm_start else 0,
            _ptr("workspace_ptr", workspace_ptr), int(workspace_bytes), _ptr("stream", stream), ctypes.byref(st),
        )
        return st.as_dict()

    @staticmethod
    def fcg_refine_workspace_bytes(nx, ny, nz, precond) -> int:
        return int(_lib.load().bs_fcg_refine_workspace_bytes(_usize("nx", nx), _usize("ny", ny), _usize("nz", nz),
                                                             int(precond)))

    def fcg_refine_solve(self, x64_ptr, div_u_ptr, mat_ptr, nx, ny, nz, dx, dt, rho, tol, precond,
                         unet_params_ptr, unet_packed_ptr, workspace_ptr, workspace_bytes, inner_tol=1e-3,
                         max_outer=20, max_inner=20000, warm_start=False, x32_ptr=0, mat_fluid=0, mat_air=2,
                         stream=0) -> dict:
        """FP64 x solving A x = b to a true relative residual ``tol`` by
        iterative refinement around the learned flexible CG (``fcg_solve``'s
        preconditioner kinds), at most ``max_inner`` inner iterations in all;
        a stalled inner flexible CG hands the remaining steps to plain
        Jacobi-PCG (``fallback``, ``pcg_iterations`` in the stats)."""
        st = _lib.RefineStats()
        _lib.call(
            "bs_fcg_refine_solve",
            _ptr("x64_ptr", x64_ptr), _ptr("x32_ptr", x32_ptr), _ptr("div_u_ptr", div_u_ptr), _ptr("mat_ptr", mat_ptr),
            _usize("nx", nx), _usize("ny", ny), _usize("nz", nz),
            float(dx), float(dt), float(rho), _i32("mat_fluid", mat_fluid), _i32("mat_air", mat_air),
            float(tol), _usize("max_outer", max_outer), float(inner_tol), _usize("max_inner", max_inner),
            1 if warm_start else 0, int(precond), _ptr("unet_params_ptr", unet_params_ptr),
            _ptr("unet_packed_ptr", unet_packed_ptr), _ptr("workspace_ptr", workspace_ptr), int(workspace_bytes),
            _ptr("stream", stream), ctypes.byref(st),
        )
        return st.as_dict()

    # ---- backward-Euler heat conduction (thermal.py:259-388, 631-665) -------
    def conduction_sweep(self, t_dst_ptr, t_iter_ptr, t_old_ptr, mat_ptr, k_ptr, rho_ptr, cp_ptr, n_mat, nx, ny, nz,
                         dx, dt, h_conv_air, t_amb_k, mat_air=2, mat_fluid=0, stream=0) -> None:
        """One apply_implicit_conduction_update sweep (asynchronous)."""
        _lib.call("bs_conduction_sweep_launch", _ptr("t_dst_ptr", t_dst_ptr), _ptr("t_iter_ptr", t_iter_ptr),
                  _ptr("t_old_ptr", t_old_ptr), _ptr("mat_ptr", mat_ptr), _ptr("k_ptr", k_ptr),
                  _ptr("rho_ptr", rho_ptr), _ptr("cp_ptr", cp_ptr), _i32("n_mat", n_mat), _usize("nx", nx),
                  _usize("ny", ny), _usize("nz", nz), float(dx), float(dt), float(h_conv_air), float(t_amb_k),
                  _i32("mat_air", mat_air), _i32("mat_fluid", mat_fluid), _ptr("stream", stream))

    def conduction_solve(self, t_ptr, t_alt_ptr, t_old_ptr, mat_ptr, k_ptr, rho_ptr, cp_ptr, n_mat, nx, ny, nz, dx,
                         dt, h_conv_air, t_amb_k, n_iter, mat_air=2, mat_fluid=0, stream=0) -> None:
        """The backward-Euler branch of conduct_one_step (asynchronous)."""
        _lib.call("bs_conduction_solve_launch", _ptr("t_ptr", t_ptr), _ptr("t_alt_ptr", t_alt_ptr),
                  _ptr("t_old_ptr", t_old_ptr), _ptr("mat_ptr", mat_ptr), _ptr("k_ptr", k_ptr),
                  _ptr("rho_ptr", rho_ptr), _ptr("cp_ptr", cp_ptr), _i32("n_mat", n_mat), _usize("nx", nx),
                  _usize("ny", ny), _usize("nz", nz), float(dx), float(dt), float(h_conv_air), float(t_amb_k),
                  _usize("n_iter", n_iter), _i32("mat_air", mat_air), _i32("mat_fluid", mat_fluid),
                  _ptr("stream", stream))
