"""Shared pytest set-up: import paths, the ``gpu`` marker and device fixtures."""

from __future__ import annotations

import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "boiling-sim_amd")
for p in (ROOT, PKG):
    if p not in sys.path:
        sys.path.insert(0, p)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (HIP device) and the built HIP library")


@pytest.fixture(scope="session")
def torch_gpu():
    import torch

    if not torch.cuda.is_available():
        pytest.fail("GPU test collected on a host without a HIP device")
    return torch


@pytest.fixture(scope="session")
def simcore(torch_gpu):
    import sim_core

    # The HIP library must be the code under test: fail loudly if it is absent.
    assert sim_core.cuda is not None, "libboilsim_hip.so failed to load"
    return sim_core.cuda.SimCore(0)
