"""Input generators and comparison helpers shared by the tests."""
import numpy as np

from vpic_amd import types as T


def voxel_coords(g, ncell):
    """(x,y,z) of interior cell k in x-fastest order, 1-based."""
    k = np.arange(ncell, dtype=np.int64)
    x = 1 + k % g.nx
    y = 1 + (k // g.nx) % g.ny
    z = 1 + k // (g.nx * g.ny)
    return x, y, z


def thermal_particles(g, ppc, uth, seed, w=1.0, drift=(0.0, 0.0, 0.0),
                      shuffle=False):
    """ppc particles per interior cell, positions uniform on [-1,1) in the
    cell, momenta N(drift, uth) per axis, equal weights (float32)."""
    rng = np.random.default_rng(seed)
    ncell = g.nx * g.ny * g.nz
    n = ncell * ppc
    p = np.zeros(n, dtype=T.PARTICLE)
    x, y, z = voxel_coords(g, ncell)
    vox = (x + (g.nx + 2) * (y + (g.ny + 2) * z)).astype(np.int32)
    p["i"] = np.repeat(vox, ppc)
    for c in ("dx", "dy", "dz"):
        p[c] = rng.uniform(-1.0, 1.0, n).astype(np.float32)
    uth = np.broadcast_to(np.asarray(uth, dtype=np.float64), (3,))
    for c, s, d in zip(("ux", "uy", "uz"), uth, drift):
        p[c] = (d + s * rng.standard_normal(n)).astype(np.float32)
    p["w"] = np.float32(w)
    if shuffle:
        p = p[rng.permutation(n)]
    return p


def random_fields(g, seed, e_amp=0.05, b_amp=0.05):
    """Random E and cB on every voxel (ghosts included), other fields zero."""
    rng = np.random.default_rng(seed)
    f = np.zeros(g.nv, dtype=T.FIELD)
    for c in ("ex", "ey", "ez"):
        f[c] = (e_amp * rng.standard_normal(g.nv)).astype(np.float32)
    for c in ("cbx", "cby", "cbz"):
        f[c] = (b_amp * rng.standard_normal(g.nv)).astype(np.float32)
    return f


def random_full_fields(g, seed, amp=1.0):
    """Every float component of field_t random (for the FDTD stencils)."""
    rng = np.random.default_rng(seed)
    f = np.zeros(g.nv, dtype=T.FIELD)
    for c in T.FIELD_FLOATS:
        f[c] = (amp * rng.standard_normal(g.nv)).astype(np.float32)
    return f


def acc_array(g):
    return np.zeros((g.nv, 12), dtype=np.float32)


def interp_array(g):
    return np.zeros(g.nv, dtype=T.INTERPOLATOR)


def same_bits(a, b):
    """Bitwise equality of two equal-dtype arrays (NaN-safe, +0 != -0)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b)
    return a.shape == b.shape and a.tobytes() == b.tobytes()


def particle_ulp_diff(a, b):
    """Max ulp distance over the float members of two particle arrays."""
    worst = 0
    for c in ("dx", "dy", "dz", "ux", "uy", "uz", "w"):
        x = a[c].astype(np.float32).view(np.int32).astype(np.int64)
        y = b[c].astype(np.float32).view(np.int32).astype(np.int64)
        # map sign-magnitude to a monotone integer line
        x = np.where(x < 0, -(x & 0x7FFFFFFF), x)
        y = np.where(y < 0, -(y & 0x7FFFFFFF), y)
        if len(x):
            worst = max(worst, int(np.max(np.abs(x - y))))
    return worst


def acc_rel_err(a, b, floor_frac=1e-3):
    """Norm-wise relative error of accumulators with an absolute floor:
    max|a-b| / max(max|b|, tiny).  Components can cancel to ~0, so an
    element-wise relative error is not meaningful for deposit sums."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    scale = max(float(np.max(np.abs(b))) if b.size else 0.0, 1e-30)
    return float(np.max(np.abs(a - b))) / scale if b.size else 0.0
