"""The frozen synthetic inputs and the oracle's config-1 outputs match the
committed sha256 digests (tests/golden/config_digests.json, written by
tests/golden/make_digests.py): the pot masks of the 64x64x48, 110x110x70 and
440x440x280 grids, the seed-7 velocities, and the divergence, 200 Jacobi
sweeps and projection of config 1."""

from __future__ import annotations

import json
import os
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))
import make_digests  # noqa: E402

GOLD = json.load(open(os.path.join(HERE, "golden", "config_digests.json")))


@pytest.mark.parametrize("name", ["small", "dev", "prod"])
def test_mask_digests(name):
    from boilsim.fixtures import CONFIGS, pot_mask

    assert make_digests.sha(pot_mask(*CONFIGS[name].shape)) == GOLD["masks"][name]


def test_config1_oracle_digests():
    c1 = make_digests.config1()
    assert c1.pop("max_div_after") == GOLD["config1"]["max_div_after"]
    for k, v in c1.items():
        assert make_digests.sha(v) == GOLD["config1"]["sha256"][k], k


@pytest.mark.parametrize("name", ["dev", "prod"])
def test_config2_oracle_digests(name):
    """Config 2 at 110x110x70 and 440x440x280: the seed-7 velocities, the
    oracle's divergence and its 200 sweeps (p, p_tmp) still hash to the
    committed digests the GPU suite checks the device against."""
    c2 = make_digests.config2(name)
    for k, v in c2.items():
        assert make_digests.sha(v) == GOLD["config2"][name]["sha256"][k], k


def test_reference_gate_oracle_digests():
    """The reference's 200-step gate case (single_carrot.yaml at dx = 4 mm,
    test_pressure_parity.py:159-224): mask, velocities and the oracle's
    projection still hash to the committed digests."""
    g = make_digests.reference_gate()
    assert g["max_div_after"] == GOLD["reference_gate"]["max_div_after"]
    for k, v in g["sha256"].items():
        assert v == GOLD["reference_gate"]["sha256"][k], k
