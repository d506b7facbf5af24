"""ORACLE — test infrastructure only.

Independent numpy restatement of the reference's Warp kernels, written with
padded whole-array operations instead of the per-cell loops of cpu_ref.c, so
the two can cross-check each other bit for bit.  Float32 scalars are cast
before use, exactly as Warp converts its ``float`` kernel arguments.

* jacobi_step: python/boilingsim/fluid.py:222-276
* divergence:  fluid.py:201-219
* gradient:    fluid.py:279-330
* enforce_no_slip: fluid.py:111-150, as masks of padded arrays
* laplacian_csr: the operator of crates/cuda-kernels/src/laplacian_spmv.cu:60-143
  as a scipy sparse matrix (FP64), assembled from the class rules.
"""

from __future__ import annotations

import numpy as np

f32 = np.float32


def _pad(a, value):
    return np.pad(a, 1, mode="constant", constant_values=value)


def jacobi_step(p_old, div, mat, dx, dt, rho, mat_fluid=0, mat_air=2):
    p_old = np.asarray(p_old, dtype=f32)
    div = np.asarray(div, dtype=f32)
    mat = np.asarray(mat, dtype=np.int32)
    pp = _pad(p_old, 0.0).astype(f32)
    mp = _pad(mat, mat_air).astype(np.int32)  # out of bounds = air
    ps = p_old
    sl = {
        "L": (slice(0, -2), slice(1, -1), slice(1, -1)),
        "R": (slice(2, None), slice(1, -1), slice(1, -1)),
        "D": (slice(1, -1), slice(0, -2), slice(1, -1)),
        "U": (slice(1, -1), slice(2, None), slice(1, -1)),
        "B": (slice(1, -1), slice(1, -1), slice(0, -2)),
        "F": (slice(1, -1), slice(1, -1), slice(2, None)),
    }
    terms = []
    for key in ("L", "R", "D", "U", "B", "F"):
        m = mp[sl[key]]
        v = pp[sl[key]]
        t = np.where(m == mat_fluid, v, np.where(m == mat_air, f32(0.0), ps)).astype(f32)
        terms.append(t)
    s = terms[0] + terms[1]
    for t in terms[2:]:
        s = (s + t).astype(f32)
    c = f32(f32(f32(rho) * f32(dx)) * f32(dx))
    rhs = ((c * div).astype(f32) / f32(dt)).astype(f32)
    out = ((s - rhs).astype(f32) / f32(6.0)).astype(f32)
    return np.where(mat == mat_fluid, out, f32(0.0)).astype(f32)


def divergence(ux, uy, uz, mat, dx, mat_fluid=0):
    dxf = f32(dx)
    dudx = ((ux[1:] - ux[:-1]).astype(f32) / dxf).astype(f32)
    dvdy = ((uy[:, 1:] - uy[:, :-1]).astype(f32) / dxf).astype(f32)
    dwdz = ((uz[:, :, 1:] - uz[:, :, :-1]).astype(f32) / dxf).astype(f32)
    d = ((dudx + dvdy).astype(f32) + dwdz).astype(f32)
    return np.where(mat == mat_fluid, d, f32(0.0)).astype(f32)


def gradient(ux, uy, uz, p, mat, dx, dt, rho, mat_fluid=0):
    """Returns corrected copies of (ux, uy, uz)."""
    ux, uy, uz = ux.copy(), uy.copy(), uz.copy()
    dtr = f32(f32(dt) / f32(rho))
    dxf = f32(dx)
    fl = mat == mat_fluid
    both = fl[1:] & fl[:-1]
    g = ((dtr * (p[1:] - p[:-1]).astype(f32)).astype(f32) / dxf).astype(f32)
    ux[1:-1] = np.where(both, (ux[1:-1] - g).astype(f32), ux[1:-1])
    both = fl[:, 1:] & fl[:, :-1]
    g = ((dtr * (p[:, 1:] - p[:, :-1]).astype(f32)).astype(f32) / dxf).astype(f32)
    uy[:, 1:-1] = np.where(both, (uy[:, 1:-1] - g).astype(f32), uy[:, 1:-1])
    both = fl[:, :, 1:] & fl[:, :, :-1]
    g = ((dtr * (p[:, :, 1:] - p[:, :, :-1]).astype(f32)).astype(f32) / dxf).astype(f32)
    uz[:, :, 1:-1] = np.where(both, (uz[:, :, 1:-1] - g).astype(f32), uz[:, :, 1:-1])
    return ux, uy, uz


def laplacian_csr(mat, mat_fluid=0, mat_air=2):
    """A = 6I - S as a scipy CSR matrix over all n cells (non-fluid rows empty)."""
    import scipy.sparse as sp

    mat = np.asarray(mat, dtype=np.int32)
    nx, ny, nz = mat.shape
    n = mat.size
    idx = np.arange(n).reshape(mat.shape)
    fl = mat == mat_fluid
    rows, cols, vals = [], [], []
    diag = np.where(fl, 6.0, 0.0)
    mp = _pad(mat, mat_air)
    ip = _pad(idx, -1)
    for sl in (
        (slice(0, -2), slice(1, -1), slice(1, -1)),
        (slice(2, None), slice(1, -1), slice(1, -1)),
        (slice(1, -1), slice(0, -2), slice(1, -1)),
        (slice(1, -1), slice(2, None), slice(1, -1)),
        (slice(1, -1), slice(1, -1), slice(0, -2)),
        (slice(1, -1), slice(1, -1), slice(2, None)),
    ):
        m = mp[sl]
        j = ip[sl]
        nb_fluid = fl & (m == mat_fluid)
        nb_solid = fl & (m != mat_fluid) & (m != mat_air)
        rows.append(idx[nb_fluid])
        cols.append(j[nb_fluid])
        vals.append(-np.ones(int(nb_fluid.sum())))
        diag = diag - nb_solid  # Neumann ghost = self: -1 on the diagonal
    rows.append(idx[fl])
    cols.append(idx[fl])
    vals.append(diag[fl])
    A = sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(n, n))
    return A


def enforce_no_slip(ux, uy, uz, mat, mat_fluid=0):
    """Returns new (ux, uy, uz): a face keeps its value only when the cells on
    both sides exist and are fluid.  The mat array is padded with a non-fluid
    ring, so the domain-edge faces fall out of the same rule."""
    mat = np.asarray(mat, dtype=np.int32)
    fl = np.pad(mat == mat_fluid, 1, mode="constant", constant_values=False)
    keep_x = fl[:-1, 1:-1, 1:-1] & fl[1:, 1:-1, 1:-1]
    keep_y = fl[1:-1, :-1, 1:-1] & fl[1:-1, 1:, 1:-1]
    keep_z = fl[1:-1, 1:-1, :-1] & fl[1:-1, 1:-1, 1:]
    z = f32(0.0)
    return (np.where(keep_x, np.asarray(ux, f32), z), np.where(keep_y, np.asarray(uy, f32), z),
            np.where(keep_z, np.asarray(uz, f32), z))
