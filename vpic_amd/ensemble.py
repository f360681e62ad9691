"""The batched ensemble: many small independent periodic boxes advanced in
one persistent, graph-captured step (vpic_hip_ensemble_*), and the
counter-streaming beam workload drawn from a seed.
"""
import ctypes

import numpy as np

from . import _lib
from . import types as T

SIDES = (32, 64, 96, 128)


class Ensemble:
    """A packed set of boxes, one pushed species each (charge q, mass m)."""

    def __init__(self, grids, nps, q=-1.0, m=1.0, stream=None):
        n = len(grids)
        self.grids = list(grids)
        self.np = [int(v) for v in nps]
        ga = (T.Grid * n)(*self.grids)
        npa = (ctypes.c_int32 * n)(*self.np)
        qa = (ctypes.c_float * n)(*([q] * n if np.isscalar(q) else q))
        ma = (ctypes.c_float * n)(*([m] * n if np.isscalar(m) else m))
        self.q = list(qa)
        self.m = list(ma)
        h = ctypes.c_void_p()
        _lib.call("vpic_hip_ensemble_create", ctypes.byref(h), n, ga, npa, qa,
                  ma, stream)
        self.h = h

    def close(self):
        if self.h:
            _lib.call("vpic_hip_ensemble_destroy", self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def box(self, b):
        """Device pointers (particles, fields, interpolators, accumulators)
        of box b."""
        p, f, fi, a = (ctypes.c_void_p() for _ in range(4))
        npb = ctypes.c_int32()
        _lib.call("vpic_hip_ensemble_box", self.h, b, ctypes.byref(p),
                  ctypes.byref(npb), ctypes.byref(f), ctypes.byref(fi),
                  ctypes.byref(a))
        return p.value, f.value, fi.value, a.value

    def set_particles(self, b, p):
        assert len(p) == self.np[b] and p.dtype == T.PARTICLE
        p = np.ascontiguousarray(p)
        _lib.call("vpic_hip_copy", self.box(b)[0], p.ctypes.data, p.nbytes,
                  None)

    def particles(self, b):
        out = np.empty(self.np[b], dtype=T.PARTICLE)
        _lib.call("vpic_hip_copy", out.ctypes.data, self.box(b)[0],
                  out.nbytes, None)
        return out

    def fields(self, b):
        out = np.empty(self.grids[b].nv, dtype=T.FIELD)
        _lib.call("vpic_hip_copy", out.ctypes.data, self.box(b)[1],
                  out.nbytes, None)
        return out

    def set_fields(self, b, f):
        f = np.ascontiguousarray(f)
        assert f.dtype == T.FIELD and len(f) == self.grids[b].nv
        _lib.call("vpic_hip_copy", self.box(b)[1], f.ctypes.data, f.nbytes,
                  None)

    def interpolators(self, b):
        out = np.empty(self.grids[b].nv, dtype=T.INTERPOLATOR)
        _lib.call("vpic_hip_copy", out.ctypes.data, self.box(b)[2],
                  out.nbytes, None)
        return out

    def accumulators(self, b):
        out = np.empty((self.grids[b].nv, 12), dtype=np.float32)
        _lib.call("vpic_hip_copy", out.ctypes.data, self.box(b)[3],
                  out.nbytes, None)
        return out

    def prepare(self):
        _lib.call("vpic_hip_ensemble_prepare", self.h)

    def step(self, n=1, graph=True):
        _lib.call("vpic_hip_ensemble_step", self.h, int(n), int(bool(graph)))

    def shape(self, push_bs=None, push_chunk=None, field_bs=None,
              field_chunk=None, blocks_per_cu=None):
        cur = (ctypes.c_int32 * 5)()
        _lib.call("vpic_hip_ensemble_get_shape", self.h, cur)
        new = [v if v is not None else cur[k] for k, v in
               enumerate((push_bs, push_chunk, field_bs, field_chunk,
                          blocks_per_cu))]
        if new != list(cur):
            _lib.call("vpic_hip_ensemble_shape", self.h, *new)
        return tuple(new)

    def sort_interval(self, n):
        """Sort every box's particles by voxel before every n-th step (0:
        never; default 20)."""
        _lib.call("vpic_hip_ensemble_sort_interval", self.h, int(n))

    def tile(self, tx=None, ty=None, tz=None, halo=None):
        """The tiled push's tile (cells per axis; 0,0,0: the flat push) and
        its window halo; returns (tx, ty, tz, halo) in use."""
        cur = (ctypes.c_int32 * 4)()
        _lib.call("vpic_hip_ensemble_get_tile_window", self.h, cur)
        new = [v if v is not None else cur[k] for k, v in
               enumerate((tx, ty, tz, halo))]
        if new != list(cur):
            _lib.call("vpic_hip_ensemble_tile_window", self.h, *new)
        return tuple(new)

    def tune(self, trial_steps=3):
        ms = ctypes.c_double()
        _lib.call("vpic_hip_ensemble_tune", self.h, int(trial_steps),
                  ctypes.byref(ms))
        return self.shape(), ms.value


def beam_config(nbox, seed, sides=SIDES, ppc=(16, 256), drift=(0.05, 0.3)):
    """The counter-streaming beam ensemble drawn from a seed: per box, the
    sides (x, y) uniformly from `sides`, particles per cell uniformly from
    ppc[0]..ppc[1], and the beam drift uniformly in drift (units of c)."""
    rng = np.random.default_rng(seed)
    nx = rng.choice(sides, nbox)
    ny = rng.choice(sides, nbox)
    ppcs = rng.integers(ppc[0], ppc[1] + 1, nbox)
    drifts = rng.uniform(drift[0], drift[1], nbox)
    return [dict(nx=int(a), ny=int(b), ppc=int(c), drift=float(d))
            for a, b, c, d in zip(nx, ny, ppcs, drifts)]


def beam_grid(nx, ny, dx=0.5):
    """A 2D periodic box (nz = 1) of square cells dx (in d_e), dt = 0.95 of
    the Courant limit."""
    lx, ly = nx * dx, ny * dx
    dt = 0.95 * T.courant_length(lx, ly, dx, nx, ny, 1)
    return T.grid_periodic(0, 0, 0, lx, ly, dx, nx, ny, 1, dt)


def beam_ensemble(cfg, seed, uth=0.01, dx=0.5, stream=None):
    """An Ensemble of two counter-streaming electron beams per box (half the
    particles of every cell drifting +x, half -x, thermal spread uth),
    loaded on the device, weights n0 dV / ppc; interpolators prepared."""
    grids = [beam_grid(c["nx"], c["ny"], dx) for c in cfg]
    nps = [c["ppc"] * c["nx"] * c["ny"] for c in cfg]
    e = Ensemble(grids, nps, -1.0, 1.0, stream)
    for b, c in enumerate(cfg):
        p = e.box(b)[0]
        g = grids[b]
        ncell = c["nx"] * c["ny"]
        half = c["ppc"] // 2
        w = float(dx ** 3 / c["ppc"])
        for k, (pp, sgn, off) in enumerate(((half, 1.0, 0),
                                            (c["ppc"] - half, -1.0,
                                             half * ncell))):
            if pp == 0:
                continue
            _lib.call("vpic_hip_load_uniform_plasma",
                      p + off * T.PARTICLE.itemsize, ctypes.byref(g), pp, uth,
                      uth, uth, sgn * c["drift"], 0.0, 0.0, w,
                      (seed * 1000003 + 2 * b + k) & ((1 << 64) - 1), None)
    e.prepare()
    return e
