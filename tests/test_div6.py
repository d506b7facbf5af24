"""The Jacobi kernels divide by 6 as fma(x, zh, x*zl) (zh = RN(1/6),
zl = RN(1/6 - zh)) for finite |x| >= 2^-96 and by the IEEE divide otherwise;
bit parity with the reference's (s - rhs) / 6 (python/boilingsim/fluid.py:276)
rests on that being the IEEE quotient for every float.  Checked here over all
2^32 bit patterns on the host's FMA (tests/div6_check.c); the GPU suite runs
the same check on the device (test_gpu_jacobi.py)."""

from __future__ import annotations

import ctypes
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


def test_div6_equals_ieee_quotient_for_every_float(tmp_path):
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    with open("/proc/cpuinfo") as f:
        flags = next((line.split() for line in f if line.startswith("flags")), [])
    if "avx2" not in flags or "fma" not in flags:
        pytest.skip("the host CPU lacks AVX2/FMA")
    so = tmp_path / "libdiv6.so"
    subprocess.run(["gcc", "-O3", "-mavx2", "-mfma", "-fopenmp", "-ffp-contract=off", "-fno-fast-math", "-shared", "-fPIC",
                    os.path.join(HERE, "div6_check.c"), "-o", str(so), "-lm"], check=True)
    counts = (ctypes.c_longlong * 4)()
    ctypes.CDLL(str(so)).div6_check(counts)
    mismatches, fma_path, one_term, unguarded = list(counts)
    print(f"div6: {mismatches} mismatches; {fma_path} floats on the fma path; x*zh alone differs on {one_term}; "
          f"the fma form without the 2^-96 guard differs on {unguarded}")
    assert mismatches == 0
    # The fma path covers almost every finite float (2^32 minus the NaN/inf
    # and tiny-magnitude patterns) ...
    assert fma_path == 7 * 2**29  # |x| in [2^-96, 2^128): 224 of the 256 exponents, both signs
    # ... and the check has teeth: the obvious one-term form and the unguarded
    # two-term form are both wrong somewhere.
    assert one_term > 0 and unguarded > 0
