"""CPU checks of the drop-in boundary: the C-ABI library loads and exports every
symbol include/boilsim_hip.h declares, the Python shim mirrors the reference's
argument conversion and error behaviour, and solver selection follows
fluid.py:585-609."""

from __future__ import annotations

import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "boilsim_hip.h")
LIB = os.path.join(ROOT, "boiling-sim_amd", "sim_core", "libboilsim_hip.so")


def header_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(bs_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from sim_core import _lib

    lib = _lib.load()
    declared = header_symbols()
    assert len(declared) >= 20
    for name in declared:
        assert hasattr(lib, name), f"{name} declared in the header but not exported"
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True, check=True).stdout
    exported = set(re.findall(r" T (bs_[a-z0-9_]+)$", out, flags=re.M))
    assert set(declared) <= exported
    # The ctypes table covers exactly the declared entry points.
    assert set(_lib.SIGNATURES) == set(declared)


def test_header_compiles_as_c():
    src = f'#include "{HEADER}"\nint main(void) {{ return (int)sizeof(bs_solve_stats); }}\n'
    r = subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-x", "c", "-", "-o", "/dev/null"],
                       input=src, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_library_carries_gfx950_code_object():
    """The fat binary built here targets gfx950 (its bundle id names it)."""
    blob = open(LIB, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in blob


def test_simcore_without_device_raises_runtime_error():
    import torch

    import sim_core

    if torch.cuda.is_available():
        pytest.skip("a HIP device is present")
    assert sim_core.cuda is not None
    with pytest.raises(RuntimeError, match="unavailable"):
        sim_core.cuda.SimCore(0)


def test_argument_conversion_mirrors_pyo3():
    from sim_core import cuda

    with pytest.raises(OverflowError):
        cuda._usize("nx", -1)
    with pytest.raises(RuntimeError, match="overflows i32"):
        cuda._usize("nx", 2**31)
    assert cuda._usize("nx", 2**31 - 1) == 2**31 - 1
    with pytest.raises(OverflowError):
        cuda._ptr("p", -5)
    with pytest.raises(OverflowError):
        cuda._i32("mat_fluid", 2**31)


def test_error_string_roundtrip():
    from sim_core import _lib

    assert "no error" in _lib.error_string(0).lower() or _lib.error_string(0)
    with pytest.raises(RuntimeError, match="hipError_t 1"):
        _lib.check("bs_dummy", 1)


@pytest.mark.parametrize("env,want", [
    ({}, "jacobi"),
    ({"BOILINGSIM_USE_RUST_PRESSURE": "1"}, "rust-jacobi"),
    ({"BOILINGSIM_PRESSURE_SOLVER": "cg-rust"}, "cg-rust"),
    ({"BOILINGSIM_PRESSURE_SOLVER": "CG-RUST"}, "cg-rust"),
    ({"BOILINGSIM_PRESSURE_SOLVER": "pcg-hip"}, "pcg-hip"),
    ({"BOILINGSIM_PRESSURE_SOLVER": "lpcg-hip"}, "lpcg-hip"),
    ({"BOILINGSIM_PRESSURE_SOLVER": "mgpcg-hip"}, "mgpcg-hip"),
    ({"BOILINGSIM_USE_RUST_PRESSURE": "0", "BOILINGSIM_PRESSURE_SOLVER": "jacobi"}, "jacobi"),
])
def test_solver_selection(monkeypatch, env, want):
    from boilsim.fluid import _select_pressure_solver

    monkeypatch.delenv("BOILINGSIM_USE_RUST_PRESSURE", raising=False)
    monkeypatch.delenv("BOILINGSIM_PRESSURE_SOLVER", raising=False)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    assert _select_pressure_solver() == want


@pytest.mark.parametrize("env", [
    {"BOILINGSIM_PRESSURE_SOLVER": "sor"},
    {"BOILINGSIM_USE_RUST_PRESSURE": "1", "BOILINGSIM_PRESSURE_SOLVER": "cg-rust"},
    {"BOILINGSIM_USE_RUST_PRESSURE": "1", "BOILINGSIM_PRESSURE_SOLVER": "mgpcg-hip"},
])
def test_solver_selection_errors(monkeypatch, env):
    from boilsim.fluid import _select_pressure_solver

    monkeypatch.delenv("BOILINGSIM_USE_RUST_PRESSURE", raising=False)
    monkeypatch.delenv("BOILINGSIM_PRESSURE_SOLVER", raising=False)
    for k, v in env.items():
        monkeypatch.setenv(k, v)
    with pytest.raises(ValueError):
        _select_pressure_solver()


def test_fastdiv_exhaustive_small():
    """The multiply-high divider used for thread -> (j, k) mapping, compiled
    for the host from the library's own header and checked exhaustively."""
    src = r'''
#include "common.hpp"
#include <cstdio>
int main() {
    for (uint32_t d = 1; d <= 3000; ++d) {
        bs::FastDiv f = bs::make_fastdiv(d);
        for (uint32_t n = 0; n < (1u << 17); ++n)
            if (bs::fdiv(n, f) != n / d) { printf("bad %u %u\n", n, d); return 1; }
        for (uint32_t n = (1u << 24) - 4096; n < (1u << 24); ++n)
            if (bs::fdiv(n, f) != n / d) { printf("bad %u %u\n", n, d); return 1; }
    }
    return 0;
}
'''
    exe = "/tmp/bs_fastdiv_check"
    cpp = "/tmp/bs_fastdiv_check.hip"
    with open(cpp, "w") as f:
        f.write(src)
    csrc = os.path.join(ROOT, "boiling-sim_amd", "csrc")
    r = subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O2", f"-I{csrc}", cpp, "-o", exe],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
