"""GPU tests of the multi-step regression (the developer guide's 50-step
protocol, multiphysics_boiling_developer_guide.md §5.4) through the state-file
format, and of the zero-copy handoff of foreign device arrays
(python/boilingsim/fluid.py:646-651)."""

from __future__ import annotations

import os

import numpy as np
import pytest

from helpers import OracleRun, host

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reduced_steps_20x20x16.bsst")


@pytest.fixture
def jacobi(monkeypatch):
    monkeypatch.delenv("BOILINGSIM_USE_RUST_PRESSURE", raising=False)
    monkeypatch.delenv("BOILINGSIM_PRESSURE_SOLVER", raising=False)


def test_device_run_reproduces_golden_state(simcore, torch_gpu, jacobi, tmp_path):
    """5 steps on the device from the fixture's initial state, dumped and
    compared with the committed oracle state: every field bit-identical."""
    from boilsim import regression, statefile

    gold, meta = statefile.load_state(GOLDEN)
    cfg = regression.StepConfig(dt=meta["dt"], n_cond=meta["n_cond"])
    cfg.solver.pressure_max_iter = meta["pressure_iters"]
    out = tmp_path / "gpu.bsst"
    r = regression.run(regression.initial_state(meta["nx"], meta["ny"], meta["nz"], meta["dx"]), cfg, meta["steps"],
                       out)
    rep = regression.compare_files(out, GOLDEN)
    assert rep["passed"], rep
    for k in ("p", "ux", "uy", "uz", "T", "div_u"):
        assert rep[k]["bits_equal"], (k, rep[k])
    assert np.float32(r.last_max_div) == np.float32(meta["max_div"])
    assert np.float32(r.ws.u_max) == np.float32(meta["u_max"])


def test_fifty_step_protocol(simcore, torch_gpu, jacobi, tmp_path):
    """The protocol itself: 50 steps of the device path and of the oracle on
    the 40x40x32 pot, both dumped; per-field max relative error under the
    limits (pressure 1e-4, temperature 1e-5) — here every field is
    bit-identical; the dumps written at step 25 agree too."""
    from boilsim import regression, statefile

    state = regression.initial_state(40, 40, 32, 2.0e-3, seed=4)
    cfg = regression.StepConfig(dt=2.0e-3)
    cfg.solver.pressure_max_iter = 100
    gpu = tmp_path / "gpu.bsst"
    r = regression.run(state, cfg, 50, gpu, every=25)
    ref = OracleRun(state, cfg)
    for s in range(50):
        ref.step()
        if s + 1 == 25:
            statefile.save_state(tmp_path / "ref.bsst.step0025", ref.fields())
    statefile.save_state(tmp_path / "ref.bsst", ref.fields(), {"steps": 50})
    for a, b in ((gpu, tmp_path / "ref.bsst"), (f"{gpu}.step0025", tmp_path / "ref.bsst.step0025")):
        rep = regression.compare_files(a, b)
        print({k: (v["max_rel"], v["bits_equal"]) for k, v in rep.items() if isinstance(v, dict)})
        assert rep["passed"]
        assert all(v["bits_equal"] for v in rep.values() if isinstance(v, dict))
    assert np.float32(r.last_max_div) == np.float32(ref.max_div) and np.float32(r.ws.u_max) == np.float32(ref.u_max)
    # the run did something: heat moved into the water, the flow stayed bounded
    fl = state["mat"] == 0
    assert host(r.grid.T)[fl].mean() != state["T"][fl].mean() and r.ws.u_max < 1.0


class _Exported:
    """A foreign device array exposing only DLPack."""

    def __init__(self, t):
        self._t = t

    def __dlpack__(self, stream=None):
        return self._t.__dlpack__()

    def __dlpack_device__(self):
        return self._t.__dlpack_device__()


class _CaiOnly:
    """A foreign device array exposing only __cuda_array_interface__."""

    def __init__(self, t):
        self._t = t
        self.__cuda_array_interface__ = {"data": (t.data_ptr(), False), "shape": tuple(t.shape),
                                         "typestr": "<f4" if t.dtype.is_floating_point else "<i4",
                                         "strides": None, "version": 3}


def test_zero_copy_handoff(simcore, torch_gpu, jacobi):
    """A grid built over DLPack-exported arrays shares their memory: a
    projection step through it changes the original tensors; raw addresses
    from an array-interface object are the tensors' own."""
    torch = torch_gpu
    from boilsim import fluid, interop
    from boilsim.fixtures import pot_mask, seeded_velocities

    mat = pot_mask(32, 32, 24)
    ux, uy, uz = seeded_velocities(32, 32, 24, seed=2)
    t = {k: torch.from_numpy(np.ascontiguousarray(v)).cuda() for k, v in
         (("mat", mat), ("ux", ux), ("uy", uy), ("uz", uz))}
    t["p"] = torch.zeros((32, 32, 24), device="cuda")
    ref = fluid.make_grid(mat, ux, uy, uz, 2e-3)
    g = interop.grid_from_arrays(*(_Exported(t[k]) for k in ("mat", "ux", "uy", "uz", "p")), dx=2e-3)
    assert g.ux.data_ptr() == t["ux"].data_ptr() and g.p.data_ptr() == t["p"].data_ptr()
    sc = fluid.ScenarioConfig()
    m_ref = fluid.projection_step(ref, fluid.allocate_fluid_workspace(ref), sc, 1e-3, 958.0)
    m = fluid.projection_step(g, fluid.allocate_fluid_workspace(g), sc, 1e-3, 958.0)
    assert m == m_ref
    for k in ("ux", "uy", "uz", "p"):
        assert np.array_equal(host(t[k]), host(getattr(ref, k)))  # the originals hold the result
    # array-interface objects: the address is the tensor's own, checked against the layout
    c = _CaiOnly(t["uy"])
    assert interop.device_ptr(c, (32, 33, 24), np.float32) == t["uy"].data_ptr()
    with pytest.raises(interop.LayoutError):
        interop.device_ptr(c, (33, 32, 24), np.float32)
    # and a SimCore call through that address works like the tensor's
    simcore.enforce_no_slip(t["ux"].data_ptr(), interop.device_ptr(c), t["uz"].data_ptr(), t["mat"].data_ptr(),
                            32, 32, 24)
    torch.cuda.synchronize()
    assert np.array_equal(host(t["uy"]), host(ref.uy))
