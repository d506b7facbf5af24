"""Zero-copy handoff of device arrays to the HIP library.

The reference hands Warp arrays to its native layer as raw addresses taken
from ``__cuda_array_interface__["data"][0]`` (python/boilingsim/fluid.py:
646-651), relying on the Warp layout: C-contiguous float32 / int32 arrays,
cell fields (nx, ny, nz) with k fastest, faces ux (nx+1, ny, nz),
uy (nx, ny+1, nz), uz (nx, ny, nz+1) (geometry.py:254-315).  This module
accepts any of

* a Python int (an address, passed through unchecked, as the reference does),
* a torch tensor on the HIP device,
* an object exposing ``__cuda_array_interface__`` (version 2 or 3: Warp,
  CuPy and the like; on ROCm the same protocol carries HIP addresses),
* an object exposing ``__dlpack__`` (imported through torch, no copy),

checks dtype, shape and C-contiguity against what the kernel expects, and
returns the address.  Nothing is copied.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np

_TYPESTR = {"f4": np.float32, "i4": np.int32, "f8": np.float64, "u1": np.uint8, "i8": np.int64}


@dataclass(frozen=True)
class ArrayView:
    """What the library needs to know about one device array."""

    ptr: int
    shape: tuple
    dtype: np.dtype
    readonly: bool = False


class LayoutError(ValueError):
    """A device array whose dtype, shape or strides the kernels cannot use."""


def _from_cai(cai: dict) -> ArrayView:
    if cai.get("version", 0) < 2:
        raise LayoutError(f"__cuda_array_interface__ version {cai.get('version')} is not supported (need >= 2)")
    typestr = cai["typestr"]
    if typestr[0] not in "<|" or typestr[1:] not in _TYPESTR:
        raise LayoutError(f"unsupported typestr {typestr!r} (little-endian f4/i4/f8/u1/i8 only)")
    dtype = np.dtype(_TYPESTR[typestr[1:]])
    shape = tuple(int(s) for s in cai["shape"])
    strides = cai.get("strides")
    if strides is not None:
        want = tuple(int(s) for s in np.empty(shape, dtype).strides) if shape else ()
        if any(n > 1 and st != w for n, st, w in zip(shape, strides, want)):
            raise LayoutError(f"strides {tuple(strides)} are not C-contiguous for shape {shape} ({want})")
    if cai.get("mask") is not None:
        raise LayoutError("masked arrays are not supported")
    ptr, readonly = cai["data"]
    return ArrayView(int(ptr), shape, dtype, bool(readonly))


def _from_torch(t) -> ArrayView:
    if not t.is_cuda:
        raise LayoutError(f"tensor on {t.device}, not on the HIP device")
    if not t.is_contiguous():
        raise LayoutError(f"tensor with strides {tuple(t.stride())} is not C-contiguous")
    dtype = np.dtype(str(t.dtype).replace("torch.", ""))
    return ArrayView(int(t.data_ptr()), tuple(t.shape), dtype)


def view(obj) -> ArrayView:
    """Describe a device array without copying it."""
    try:
        import torch
    except ImportError:  # pragma: no cover - torch is part of the runtime
        torch = None
    if torch is not None and isinstance(obj, torch.Tensor):
        return _from_torch(obj)
    cai = getattr(obj, "__cuda_array_interface__", None)
    if cai is not None:
        return _from_cai(cai)
    if hasattr(obj, "__dlpack__"):
        return _from_torch(torch.from_dlpack(obj))
    raise LayoutError(f"{type(obj).__name__} exposes no device-array protocol "
                      "(__cuda_array_interface__, __dlpack__) and is not a torch tensor")


def device_ptr(obj, shape=None, dtype=None, writable: bool = False) -> int:
    """The device address of `obj`, after checking it against the expected
    shape and dtype.  An int is returned as given (a raw address)."""
    if isinstance(obj, (int, np.integer)):
        return int(obj)
    v = view(obj)
    if dtype is not None and v.dtype != np.dtype(dtype):
        raise LayoutError(f"dtype {v.dtype}, expected {np.dtype(dtype)}")
    if shape is not None and v.shape != tuple(shape):
        raise LayoutError(f"shape {v.shape}, expected {tuple(shape)}")
    if writable and v.readonly:
        raise LayoutError("array is read-only but the kernel writes it")
    return v.ptr


def grid_shapes(nx: int, ny: int, nz: int) -> dict:
    """Expected shapes of the grid's fields (geometry.py:254-282)."""
    c = (nx, ny, nz)
    return {"p": c, "mat": c, "T": c, "div_u": c, "p_tmp": c,
            "ux": (nx + 1, ny, nz), "uy": (nx, ny + 1, nz), "uz": (nx, ny, nz + 1)}


def as_torch(obj):
    """A torch view of a device array (no copy): the tensor itself, a
    DLPack import, or torch.as_tensor over the array interface."""
    import torch

    if isinstance(obj, torch.Tensor):
        return obj
    if hasattr(obj, "__dlpack__"):
        return torch.from_dlpack(obj)
    if hasattr(obj, "__cuda_array_interface__"):
        return torch.as_tensor(obj, device="cuda")
    raise LayoutError(f"{type(obj).__name__} is not a device array")


def grid_from_arrays(mat, ux, uy, uz, p, dx: float, T=None, origin=(0.0, 0.0, 0.0)):
    """A boilsim.fluid.Grid over existing device arrays (e.g. the reference's
    Warp grid.mat, grid.ux, ...), sharing their memory: results written by
    the library are visible through the original arrays."""
    from .fluid import Grid

    mv = view(mat)
    if len(mv.shape) != 3:
        raise LayoutError(f"mat must be 3-D, got shape {mv.shape}")
    nx, ny, nz = mv.shape
    shp = grid_shapes(nx, ny, nz)
    fields = {"mat": (mat, np.int32), "ux": (ux, np.float32), "uy": (uy, np.float32), "uz": (uz, np.float32),
              "p": (p, np.float32)}
    if T is not None:
        fields["T"] = (T, np.float32)
    out = {}
    for name, (a, dt) in fields.items():
        device_ptr(a, shp[name], dt)
        out[name] = as_torch(a)
    return Grid(nx, ny, nz, float(dx), out["p"], out["mat"], out["ux"], out["uy"], out["uz"], out.get("T"),
                tuple(float(c) for c in origin))


__all__ = ["ArrayView", "LayoutError", "view", "device_ptr", "grid_shapes", "as_torch", "grid_from_arrays"]
