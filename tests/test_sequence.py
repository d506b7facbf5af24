"""CPU tests of the correlated right-hand-side sequence: the AR(1) recursion
and the unit variance the frame definition states, checked on the generated
fields (plume subtracted)."""

from __future__ import annotations

import math

import numpy as np
import torch

from boilsim.fixtures import pot_mask
from boilsim.sequence import STEP_CORR, ar1_frames, plume_source


def test_ar1_frames_statistics():
    shape = (48, 48, 36)
    mat = torch.from_numpy(pot_mask(*shape))
    fl = (mat == 0).numpy()
    fs = []
    # A short correlation length gives enough independent modes on a small
    # grid for the sample statistics to be tight.
    for t, b in ar1_frames(torch, mat, 4, seed=3, ell=1.5):
        f = (b - plume_source(torch, shape, t, device="cpu")).numpy()[fl]
        fs.append(f)
        assert not b.numpy()[~fl].any()
    for f in fs:
        assert abs(f.std() - 1.0) < 0.05 and abs(f.mean()) < 0.05
    # f_{t+1} = a f_t + sqrt(1-a^2) xi with xi independent of f_t: the
    # regression coefficient of f_{t+1} on f_t is a.
    for f0, f1 in zip(fs, fs[1:]):
        coef = float(np.dot(f0, f1) / np.dot(f0, f0))
        assert abs(coef - STEP_CORR) < 0.06, coef
    # A dropped innovation term (coef 1) or a wrong coefficient would fail above;
    # the innovation itself has variance 1 - a^2.
    innov = fs[1] - STEP_CORR * fs[0]
    assert abs(innov.var() - (1 - STEP_CORR ** 2)) < 0.03


def test_plume_drifts_around_axis():
    shape = (64, 64, 48)
    a = plume_source(torch, shape, 0, device="cpu")
    b = plume_source(torch, shape, 25, device="cpu")
    ia = np.unravel_index(int(torch.argmax(a)), shape)
    ib = np.unravel_index(int(torch.argmax(b)), shape)
    # quarter period: the centre moved 90 degrees around (32, 32) at radius 13
    ra = math.hypot(ia[0] + 0.5 - 32, ia[1] + 0.5 - 32)
    rb = math.hypot(ib[0] + 0.5 - 32, ib[1] + 0.5 - 32)
    assert abs(ra - 13) < 1 and abs(rb - 13) < 1 and ia != ib
    assert abs(float(a.max()) - 2.0) < 0.05
