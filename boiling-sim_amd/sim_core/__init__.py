"""Drop-in replacement for the reference's ``sim_core`` package on MI355X.

The reference ships ``sim_core`` as a PyO3 extension with a ``cuda`` submodule
that is ``None`` when no CUDA device or driver is present
(python/sim_core/__init__.py:33-39).  Here ``sim_core.cuda`` wraps the hand-
written gfx950 HIP library ``libboilsim_hip.so`` through ctypes and keeps the
same contract: ``sim_core.cuda`` is ``None`` when the library cannot be loaded,
and ``sim_core.cuda.SimCore(device_id)`` raises ``RuntimeError`` when the device
is unavailable.  ``props`` (the reference's CPU material table) is out of scope.
"""

from __future__ import annotations

try:
    from . import _lib as _libmod

    _libmod.load()
    from . import cuda  # noqa: F401
except (OSError, RuntimeError):
    cuda = None  # type: ignore[assignment]
