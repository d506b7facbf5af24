"""On-disk simulation state for multi-step regression (the dumps of the
developer guide's 50-step protocol, multiphysics_boiling_developer_guide.md
§5.4, which the reference writes as HDF5 in Simulation.run,
python/boilingsim/pipeline.py:893-990).

h5py is not part of this runtime, so the state goes to a small
self-describing binary file instead:

    bytes 0-7     magic  b"BSSTATE1"
    bytes 8-15    little-endian uint64 H, the length of the header
    bytes 16-16+H UTF-8 JSON header:
                  {"format": "boilsim-state", "version": 1, "meta": {...},
                   "fields": [{"name", "dtype" (numpy typestr, little-endian),
                               "shape", "offset", "nbytes", "crc32"}, ...]}
    then each field's C-contiguous bytes at its offset (64-byte aligned).

Fields can be memory-mapped (np.memmap) for states too large to read whole;
every field carries a CRC-32 of its bytes.  ``compare`` computes the
protocol's metrics per field — max absolute and max relative error, the
relative error |a - b| / max(|ref|, 1e-6) as in the reference's parity gate
(python/tests/test_pressure_parity.py:208-224), pressure over fluid cells —
and applies its acceptance limits (pressure 1e-4, temperature 1e-5,
concentration 1e-5).
"""

from __future__ import annotations

import json
import os
import struct
import zlib

import numpy as np

MAGIC = b"BSSTATE1"
VERSION = 1
ALIGN = 64

# Acceptance limits of the regression protocol (max relative error).
PROTOCOL_LIMITS = {"p": 1e-4, "T": 1e-5, "C": 1e-5}


class StateFileError(ValueError):
    """A state file that is not one, is truncated, or fails its checksum."""


def _host(a) -> np.ndarray:
    if hasattr(a, "detach"):  # torch tensor (device or host)
        a = a.detach().cpu().numpy()
    a = np.ascontiguousarray(a)
    if a.dtype.byteorder == ">":
        a = a.astype(a.dtype.newbyteorder("<"))
    return a


def save_state(path, fields: dict, meta: dict | None = None) -> None:
    """Write `fields` (name -> numpy array or torch tensor) and `meta` (any
    JSON-serialisable dict) to `path`."""
    arrays = {k: _host(v) for k, v in fields.items()}
    entries = []
    offset = 0
    for name, a in arrays.items():
        entries.append({"name": name, "dtype": a.dtype.str, "shape": list(a.shape), "offset": offset,
                        "nbytes": int(a.nbytes), "crc32": zlib.crc32(a.view(np.uint8).ravel()) & 0xFFFFFFFF})
        offset += (a.nbytes + ALIGN - 1) // ALIGN * ALIGN
    header = {"format": "boilsim-state", "version": VERSION, "meta": meta or {}, "fields": entries}
    hb = json.dumps(header, sort_keys=True).encode()
    base = (16 + len(hb) + ALIGN - 1) // ALIGN * ALIGN
    hb += b" " * (base - 16 - len(hb))  # pad the header so the data starts aligned
    tmp = f"{path}.tmp{os.getpid()}"
    with open(tmp, "wb") as f:
        f.write(MAGIC)
        f.write(struct.pack("<Q", len(hb)))
        f.write(hb)
        for e, a in zip(entries, arrays.values()):
            f.seek(base + e["offset"])
            f.write(a.view(np.uint8).ravel().tobytes())
        end = base + (entries[-1]["offset"] + entries[-1]["nbytes"] if entries else 0)
        f.truncate(end)
    os.replace(tmp, path)


def read_header(path) -> tuple[dict, int]:
    """The JSON header and the file offset where field data starts."""
    with open(path, "rb") as f:
        head = f.read(16)
        if len(head) < 16 or head[:8] != MAGIC:
            raise StateFileError(f"{path}: not a boilsim state file (magic {head[:8]!r})")
        (h,) = struct.unpack("<Q", head[8:16])
        hb = f.read(h)
    if len(hb) != h:
        raise StateFileError(f"{path}: truncated header")
    header = json.loads(hb.decode())
    if header.get("format") != "boilsim-state" or header.get("version") != VERSION:
        raise StateFileError(f"{path}: unsupported format {header.get('format')} v{header.get('version')}")
    return header, 16 + h


def load_state(path, mmap: bool = False, verify: bool = True) -> tuple[dict, dict]:
    """(fields, meta).  With mmap=True the arrays are read-only memory maps
    of the file; verify=True checks every field's CRC-32."""
    header, base = read_header(path)
    size = os.path.getsize(path)
    out = {}
    for e in header["fields"]:
        dt = np.dtype(e["dtype"])
        shape = tuple(e["shape"])
        start = base + e["offset"]
        if start + e["nbytes"] > size:
            raise StateFileError(f"{path}: field {e['name']} runs past the end of the file")
        if mmap:
            a = np.memmap(path, dtype=dt, mode="r", offset=start, shape=shape)
        else:
            with open(path, "rb") as f:
                f.seek(start)
                a = np.frombuffer(f.read(e["nbytes"]), dtype=dt).reshape(shape)
        if verify and (zlib.crc32(np.asarray(a).view(np.uint8).ravel()) & 0xFFFFFFFF) != e["crc32"]:
            raise StateFileError(f"{path}: field {e['name']} fails its CRC-32")
        out[e["name"]] = a
    return out, header["meta"]


def grid_fields(grid, ws=None) -> dict:
    """The state of a boilsim.fluid.Grid (and its workspace's divergence)."""
    f = {"mat": grid.mat, "p": grid.p, "ux": grid.ux, "uy": grid.uy, "uz": grid.uz}
    if getattr(grid, "T", None) is not None:
        f["T"] = grid.T
    if ws is not None and getattr(ws, "div_u", None) is not None:
        f["div_u"] = ws.div_u
    return f


def dump_grid(path, grid, ws=None, **meta) -> None:
    """save_state of a grid, with nx, ny, nz, dx and origin in the meta."""
    m = {"nx": grid.nx, "ny": grid.ny, "nz": grid.nz, "dx": grid.dx, "origin": list(grid.origin)}
    m.update(meta)
    save_state(path, grid_fields(grid, ws), m)


def field_error(a, ref, mask=None, floor: float = 1e-6) -> dict:
    """max |a - ref|, max |a - ref| / max(|ref|, floor) and whether the bits
    are identical, over `mask` (all elements if None)."""
    a = np.asarray(a)
    ref = np.asarray(ref)
    if a.shape != ref.shape:
        raise ValueError(f"shape {a.shape} vs {ref.shape}")
    bits_equal = bool(np.array_equal(a.view(np.uint8), ref.view(np.uint8)))
    if mask is not None:
        a, ref = a[mask], ref[mask]
    if a.size == 0:
        return {"max_abs": 0.0, "max_rel": 0.0, "bits_equal": bits_equal}
    a64, r64 = a.astype(np.float64), ref.astype(np.float64)
    diff = np.abs(a64 - r64)
    rel = diff / np.maximum(np.abs(r64), floor)
    return {"max_abs": float(diff.max()), "max_rel": float(rel.max()), "bits_equal": bits_equal}


def compare(state: dict, ref: dict, fluid_value: int = 0, limits: dict | None = None) -> dict:
    """The protocol's per-field report between two states (name -> array):
    max_abs, max_rel, bits_equal, the limit (None where the protocol sets
    none) and whether it passes.  Pressure is compared over fluid cells
    (mat == fluid_value) when both states hold `mat`; other fields over all
    elements.  A field of `ref` that `state` lacks, or holds with another
    shape or dtype, fails (its entry carries "error" instead of the metrics).
    report["passed"] is True when at least one field was compared and every
    field passes."""
    limits = PROTOCOL_LIMITS if limits is None else limits
    fluid = None
    if "mat" in state and "mat" in ref:
        if not np.array_equal(np.asarray(state["mat"]), np.asarray(ref["mat"])):
            raise ValueError("the two states have different geometry (mat differs)")
        fluid = np.asarray(ref["mat"]) == fluid_value
    report = {}
    for name in ref:
        if name == "mat":
            continue
        lim = limits.get(name)
        if name not in state:
            report[name] = {"error": "missing from the state", "limit": lim, "passes": False}
            continue
        a, b = np.asarray(state[name]), np.asarray(ref[name])
        if a.shape != b.shape or a.dtype != b.dtype:
            report[name] = {"error": f"{a.dtype}{a.shape} against the reference's {b.dtype}{b.shape}",
                            "limit": lim, "passes": False}
            continue
        r = field_error(a, b, fluid if name == "p" else None)
        r["limit"] = lim
        r["passes"] = True if lim is None else r["max_rel"] < lim
        report[name] = r
    entries = [v for v in report.values() if isinstance(v, dict)]
    report["passed"] = bool(entries) and all(v["passes"] for v in entries)
    return report


__all__ = ["MAGIC", "PROTOCOL_LIMITS", "StateFileError", "save_state", "load_state", "read_header", "grid_fields",
           "dump_grid", "field_error", "compare"]
