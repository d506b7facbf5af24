"""ORACLE — test infrastructure only.

Independent numpy restatement of the reference's Warp kernels, written with
padded whole-array operations instead of the per-cell loops of cpu_ref.c, so
the two can cross-check each other bit for bit.  Float32 scalars are cast
before use, exactly as Warp converts its ``float`` kernel arguments.

* jacobi_step: python/boilingsim/fluid.py:222-276
* divergence:  fluid.py:201-219
* gradient:    fluid.py:279-330
* enforce_no_slip: fluid.py:111-150, as masks of padded arrays
* extend_temperature / advect: fluid.py:33-103, 338-492, 737-775, whole
  output arrays at a time
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


def extend_temperature(t, mat, mat_fluid=0):
    """Non-fluid cells take T of the first fluid neighbour in the order
    i-1, i+1, j-1, j+1, k-1, k+1 (none: their own T)."""
    t = np.asarray(t, f32)
    fl = np.asarray(mat) == mat_fluid
    out = t.copy()
    done = fl.copy()
    tp = np.pad(t, 1)
    fp = np.pad(fl, 1, constant_values=False)
    c = (slice(1, -1),) * 3
    for ax, d in ((0, -1), (0, 1), (1, -1), (1, 1), (2, -1), (2, 1)):
        sl = list(c)
        sl[ax] = slice(1 + d, tp.shape[ax] - 1 + d)
        sl = tuple(sl)
        take = ~done & fp[sl]
        out[take] = tp[sl][take]
        done |= take
    return out


def _tri(field, fx, fy, fz):
    n0, n1, n2 = field.shape
    fx = np.fmin(np.fmax(f32(0.0), fx), f32(n0 - 1) - f32(1.0e-6))
    fy = np.fmin(np.fmax(f32(0.0), fy), f32(n1 - 1) - f32(1.0e-6))
    fz = np.fmin(np.fmax(f32(0.0), fz), f32(n2 - 1) - f32(1.0e-6))
    i0, j0, k0 = fx.astype(np.int64), fy.astype(np.int64), fz.astype(np.int64)
    tx, ty, tz = fx - i0.astype(f32), fy - j0.astype(f32), fz - k0.astype(f32)
    i1, j1, k1 = np.minimum(i0 + 1, n0 - 1), np.minimum(j0 + 1, n1 - 1), np.minimum(k0 + 1, n2 - 1)
    one = f32(1.0)
    c00 = field[i0, j0, k0] * (one - tx) + field[i1, j0, k0] * tx
    c10 = field[i0, j1, k0] * (one - tx) + field[i1, j1, k0] * tx
    c01 = field[i0, j0, k1] * (one - tx) + field[i1, j0, k1] * tx
    c11 = field[i0, j1, k1] * (one - tx) + field[i1, j1, k1] * tx
    c0 = c00 * (one - ty) + c10 * ty
    c1 = c01 * (one - ty) + c11 * ty
    return c0 * (one - tz) + c1 * tz


def _coords(px, py, pz, o, dx, off):
    dx = f32(dx)
    return ((px - o[0]) / dx - f32(off[0]), (py - o[1]) / dx - f32(off[1]), (pz - o[2]) / dx - f32(off[2]))


_OFFS = {"ux": (0.0, 0.5, 0.5), "uy": (0.5, 0.0, 0.5), "uz": (0.5, 0.5, 0.0), "t": (0.5, 0.5, 0.5)}


def advect(ux, uy, uz, t, mat, origin, dx, dt, mat_fluid=0):
    """Returns (ux_new, uy_new, uz_new, t_new)."""
    ux, uy, uz, t = (np.asarray(a, f32) for a in (ux, uy, uz, t))
    mat = np.asarray(mat)
    nx, ny, nz = mat.shape
    o = tuple(f32(c) for c in origin)
    dx, dt = f32(dx), f32(dt)
    t_ext = extend_temperature(t, mat, mat_fluid)
    fl = mat == mat_fluid
    fields = {"ux": ux, "uy": uy, "uz": uz, "t": t_ext}
    out = {}
    for name, shape in (("ux", ux.shape), ("uy", uy.shape), ("uz", uz.shape), ("t", t.shape)):
        off = _OFFS[name]
        I, J, K = (g.astype(f32) for g in np.indices(shape))
        # the output point: integer index on the face axis, +0.5 elsewhere
        px = o[0] + (I + f32(off[0])) * dx
        py = o[1] + (J + f32(off[1])) * dx
        pz = o[2] + (K + f32(off[2])) * dx
        u = _tri(ux, *_coords(px, py, pz, o, dx, _OFFS["ux"]))
        v = _tri(uy, *_coords(px, py, pz, o, dx, _OFFS["uy"]))
        w = _tri(uz, *_coords(px, py, pz, o, dx, _OFFS["uz"]))
        bx, by, bz = px - u * dt, py - v * dt, pz - w * dt
        val = _tri(fields[name], *_coords(bx, by, bz, o, dx, off))
        if name == "t":
            out[name] = np.where(fl, val, t)
            continue
        ax = "xyz".index(name[1])
        src = fields[name]
        fpad = np.pad(fl, [(1, 1) if a == ax else (0, 0) for a in range(3)], constant_values=False)
        lo = fpad[tuple(slice(0, -1) if a == ax else slice(None) for a in range(3))]
        hi = fpad[tuple(slice(1, None) if a == ax else slice(None) for a in range(3))]
        res = np.where(lo | hi, val, src)
        edge = np.zeros(shape, bool)
        edge[tuple(slice(0, 1) if a == ax else slice(None) for a in range(3))] = True
        edge[tuple(slice(-1, None) if a == ax else slice(None) for a in range(3))] = True
        out[name] = np.where(edge, f32(0.0), res).astype(f32)
    return out["ux"], out["uy"], out["uz"], out["t"]
