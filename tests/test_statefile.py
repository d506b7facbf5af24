"""CPU tests of the state-file format, the regression protocol's comparison
(multiphysics_boiling_developer_guide.md §5.4; metric of
python/tests/test_pressure_parity.py:208-224), the golden multi-step fixture,
and the zero-copy handoff's layout checks (fluid.py:646-651)."""

from __future__ import annotations

import os

import numpy as np
import pytest

from boilsim import interop, statefile

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reduced_steps_20x20x16.bsst")


def test_round_trip(tmp_path):
    rng = np.random.default_rng(0)
    fields = {"p": rng.standard_normal((5, 6, 7)).astype(np.float32), "mat": rng.integers(0, 4, (5, 6, 7), np.int32),
              "x64": rng.standard_normal(33), "empty": np.zeros((0, 3), np.float32),
              "be": np.arange(5, dtype=">f4")}
    path = tmp_path / "s.bsst"
    statefile.save_state(path, fields, {"step": 7, "note": "x"})
    for mmap in (False, True):
        got, meta = statefile.load_state(path, mmap=mmap)
        assert meta == {"step": 7, "note": "x"}
        for k, v in fields.items():
            assert got[k].dtype == v.dtype.newbyteorder("<") and got[k].shape == v.shape
            assert np.array_equal(np.asarray(got[k]), v)
    header, base = statefile.read_header(path)
    assert base % statefile.ALIGN == 0
    assert all(e["offset"] % statefile.ALIGN == 0 for e in header["fields"])


def test_corruption_is_detected(tmp_path):
    path = tmp_path / "s.bsst"
    statefile.save_state(path, {"a": np.arange(100, dtype=np.float32)})
    raw = bytearray(path.read_bytes())
    _, base = statefile.read_header(path)
    assert len(raw) == base + 400  # the file ends with the last field
    raw[base + 123] ^= 0x40
    bad = tmp_path / "bad.bsst"
    bad.write_bytes(bytes(raw))
    with pytest.raises(statefile.StateFileError, match="CRC"):
        statefile.load_state(bad)
    statefile.load_state(bad, verify=False)  # readable when asked not to verify
    (tmp_path / "trunc.bsst").write_bytes(bytes(raw[:-40]))
    with pytest.raises(statefile.StateFileError):
        statefile.load_state(tmp_path / "trunc.bsst")
    (tmp_path / "magic.bsst").write_bytes(b"NOTSTATE" + bytes(raw[8:]))
    with pytest.raises(statefile.StateFileError, match="magic"):
        statefile.load_state(tmp_path / "magic.bsst")


def test_protocol_comparison():
    rng = np.random.default_rng(1)
    mat = rng.choice([0, 1, 2], size=(6, 6, 6)).astype(np.int32)
    ref = {"mat": mat, "p": rng.standard_normal((6, 6, 6)).astype(np.float32),
           "T": (300 + rng.random((6, 6, 6))).astype(np.float32),
           "ux": rng.standard_normal((7, 6, 6)).astype(np.float32)}
    same = {k: v.copy() for k, v in ref.items()}
    rep = statefile.compare(same, ref)
    assert rep["passed"] and all(rep[k]["bits_equal"] for k in ("p", "T", "ux"))
    # pressure: 2e-4 relative in a fluid cell fails, anything in a non-fluid cell is ignored
    fl = np.argwhere(mat == 0)[0]
    nf = np.argwhere(mat != 0)[0]
    bad = {k: v.copy() for k, v in ref.items()}
    bad["p"][tuple(nf)] += 100.0
    assert statefile.compare(bad, ref)["passed"]
    bad["p"][tuple(fl)] *= np.float32(1 + 2e-4)
    rep = statefile.compare(bad, ref)
    assert not rep["passed"] and not rep["p"]["passes"] and 1.5e-4 < rep["p"]["max_rel"] < 2.5e-4
    # temperature: limit 1e-5
    t = {k: v.copy() for k, v in ref.items()}
    t["T"][0, 0, 0] *= np.float32(1 + 5e-6)
    assert statefile.compare(t, ref)["passed"]
    t["T"][0, 0, 0] *= np.float32(1 + 2e-5)
    assert not statefile.compare(t, ref)["T"]["passes"]
    # velocity has no limit in the protocol: reported, never failing
    v = {k: x.copy() for k, x in ref.items()}
    v["ux"] += 1.0
    r = statefile.compare(v, ref)
    assert r["passed"] and r["ux"]["limit"] is None and r["ux"]["max_abs"] == pytest.approx(1.0, rel=1e-6)
    with pytest.raises(ValueError, match="geometry"):
        statefile.compare({**ref, "mat": mat + 1}, ref)


def test_protocol_comparison_missing_or_mismatched_fields():
    """A state that lacks a reference field, or holds it with another shape or
    dtype, fails the protocol; an empty state fails too."""
    rng = np.random.default_rng(2)
    mat = rng.choice([0, 1, 2], size=(5, 4, 6)).astype(np.int32)
    ref = {"mat": mat, "p": rng.standard_normal((5, 4, 6)).astype(np.float32),
           "T": (300 + rng.random((5, 4, 6))).astype(np.float32),
           "ux": rng.standard_normal((6, 4, 6)).astype(np.float32)}
    for drop in ("p", "T", "ux"):
        st = {k: v.copy() for k, v in ref.items() if k != drop}
        rep = statefile.compare(st, ref)
        assert not rep["passed"] and not rep[drop]["passes"] and "missing" in rep[drop]["error"]
    assert not statefile.compare({}, ref)["passed"]
    assert not statefile.compare({"mat": mat}, ref)["passed"]
    assert not statefile.compare(ref, {"mat": mat})["passed"]  # nothing compared
    wrong_shape = {**ref, "T": ref["T"][:, :, :5].copy()}
    rep = statefile.compare(wrong_shape, ref)
    assert not rep["passed"] and "(5, 4, 5)" in rep["T"]["error"]
    wrong_dtype = {**ref, "p": ref["p"].astype(np.float64)}
    rep = statefile.compare(wrong_dtype, ref)
    assert not rep["passed"] and "float64" in rep["p"]["error"]


def test_golden_fixture_reproduced_by_oracle():
    """The committed 5-step state equals what the oracle computes today."""
    import sys

    sys.path.insert(0, os.path.dirname(GOLDEN))
    import make_state_fixture as mk

    gold, meta = statefile.load_state(GOLDEN)
    assert meta["steps"] == mk.STEPS and (meta["nx"], meta["ny"], meta["nz"]) == mk.SHAPE
    run = mk.compute()
    rep = statefile.compare(run.fields(), gold)
    assert rep["passed"] and all(rep[k]["bits_equal"] for k in ("p", "ux", "uy", "uz", "T", "div_u"))
    # the state has moved: velocity advected and projected, heat conducted
    from boilsim.regression import initial_state

    s0 = initial_state(*mk.SHAPE, mk.DX)
    assert not np.array_equal(gold["T"], s0["T"]) and not np.array_equal(gold["uz"], s0["uz"])


class _CAI:
    def __init__(self, ptr, shape, typestr="<f4", strides=None, version=3, readonly=False):
        self.__cuda_array_interface__ = {"data": (ptr, readonly), "shape": shape, "typestr": typestr,
                                         "strides": strides, "version": version}


def test_interop_layout_checks():
    assert interop.device_ptr(123456) == 123456
    a = _CAI(0x7F0000001000, (4, 5, 6))
    assert interop.device_ptr(a, (4, 5, 6), np.float32) == 0x7F0000001000
    assert interop.device_ptr(_CAI(64, (4, 5, 6), strides=(120, 24, 4)), (4, 5, 6), np.float32) == 64
    with pytest.raises(interop.LayoutError, match="shape"):
        interop.device_ptr(a, (5, 5, 6), np.float32)
    with pytest.raises(interop.LayoutError, match="dtype"):
        interop.device_ptr(a, (4, 5, 6), np.int32)
    with pytest.raises(interop.LayoutError, match="C-contiguous"):
        interop.device_ptr(_CAI(64, (4, 5, 6), strides=(4, 16, 80)))
    with pytest.raises(interop.LayoutError, match="typestr"):
        interop.device_ptr(_CAI(64, (4,), typestr=">f4"))
    with pytest.raises(interop.LayoutError, match="version"):
        interop.device_ptr(_CAI(64, (4,), version=1))
    with pytest.raises(interop.LayoutError, match="read-only"):
        interop.device_ptr(_CAI(64, (4,), readonly=True), writable=True)
    with pytest.raises(interop.LayoutError, match="protocol"):
        interop.device_ptr(object())
    # a host torch tensor is refused (the kernels need device memory)
    import torch

    with pytest.raises(interop.LayoutError, match="device"):
        interop.device_ptr(torch.zeros(3))
    assert interop.grid_shapes(4, 5, 6)["uy"] == (4, 6, 6)
