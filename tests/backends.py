"""Two interchangeable executors of the hot-path operations for the tests:
the oracle (CPU restatement) and the HIP library through the C-ABI.  Both
take and return host numpy arrays so a test can run the same deck on
either."""
import ctypes

import numpy as np

from vpic_amd import types as T


class OracleBackend:
    name = "oracle"

    def __init__(self):
        from oracle.oracle import Oracle, host_grid
        self.o = Oracle.get()
        self._hg = host_grid

    def advance_p(self, p, q, m, acc, fi, g, max_nm=None, tiled=False):
        return self.o.advance_p(p, q, m, acc, fi, self._hg(g), max_nm=max_nm)

    def clear_accumulator(self, acc, g):
        acc[...] = 0

    def load_interpolator(self, fi, f, g):
        self.o.load_interpolator(fi, f, g)

    def clear_jf(self, f, g):
        self.o.clear_jf(f, g)

    def clear_rhof(self, f, g):
        self.o.clear_rhof(f, g)

    def unload_accumulator(self, f, acc, g):
        self.o.unload_accumulator(f, acc, g)

    def synchronize_jf(self, f, g):
        self.o.synchronize_jf(f, g)

    def accumulate_rho_p(self, f, p, q, g):
        self.o.accumulate_rho_p(f, p, q, g)

    def synchronize_rho(self, f, g):
        self.o.synchronize_rho(f, g)


class GpuBackend:
    name = "gpu"

    def __init__(self):
        from vpic_amd import _lib, device
        _lib.call("vpic_hip_init", 0)
        self.lib = _lib
        self.dev = device

    def _up(self, a):
        return self.dev.DeviceArray.from_host(a)

    @staticmethod
    def _back(d, a):
        a[...] = d.download()

    def advance_p(self, p, q, m, acc, fi, g, max_nm=None, tiled=False):
        """tiled=True sorts the particles on the device first (p comes back
        in sorted order) and hands the partition to advance_p, which then
        runs the cell-tiled push."""
        if max_nm is None:
            max_nm = max(len(p), 16)
        dp, da, dfi = self._up(p), self._up(acc), self._up(fi)
        dpm = self.dev.movers(max_nm)
        part = None
        if tiled:
            aux = self.dev.particles(len(p))
            dpart = self.dev.DeviceArray((g.nv + 1,), np.int32)
            self.lib.call("vpic_hip_sort_p", dp.ptr, aux.ptr, len(p),
                          dpart.ptr, ctypes.byref(g), 1, None)
            part = dpart.ptr
        nm, ign = ctypes.c_int32(), ctypes.c_int32()
        self.lib.call("vpic_hip_advance_p", dp.ptr, len(p), q, m, dpm.ptr,
                      max_nm, ctypes.byref(nm), ctypes.byref(ign), da.ptr,
                      dfi.ptr, ctypes.byref(g), part, None)
        self._back(dp, p)
        self._back(da, acc)
        return nm.value, ign.value, dpm.download()[:nm.value].copy()

    def clear_accumulator(self, acc, g):
        da = self._up(acc)
        self.lib.call("vpic_hip_clear_accumulator_array", da.ptr,
                      ctypes.byref(g), None)
        self._back(da, acc)

    def _field_op(self, name, f, g, *extra):
        df = self._up(f)
        self.lib.call(name, df.ptr, *extra, ctypes.byref(g), None)
        self._back(df, f)

    def load_interpolator(self, fi, f, g):
        dfi, df = self._up(fi), self._up(f)
        self.lib.call("vpic_hip_load_interpolator_array", dfi.ptr, df.ptr,
                      ctypes.byref(g), None)
        self._back(dfi, fi)

    def clear_jf(self, f, g):
        self._field_op("vpic_hip_clear_jf", f, g)

    def clear_rhof(self, f, g):
        self._field_op("vpic_hip_clear_rhof", f, g)

    def unload_accumulator(self, f, acc, g):
        da = self._up(acc)
        self._field_op("vpic_hip_unload_accumulator_array", f, g, da.ptr)

    def synchronize_jf(self, f, g):
        self._field_op("vpic_hip_synchronize_jf", f, g)

    def accumulate_rho_p(self, f, p, q, g):
        df, dp = self._up(f), self._up(p)
        self.lib.call("vpic_hip_accumulate_rho_p", df.ptr, dp.ptr, len(p), q,
                      ctypes.byref(g), None)
        self._back(df, f)

    def synchronize_rho(self, f, g):
        self._field_op("vpic_hip_synchronize_rho", f, g)


def make(name):
    return OracleBackend() if name == "oracle" else GpuBackend()


def inject(g, x, y, z, ux, uy, uz, w, lo=(0.0, 0.0, 0.0), hi=None):
    """inject_particle's position -> (voxel, cell offset) mapping, in double
    then rounded to float (src/vpic/misc.cc:43-92).  Arrays in, particles
    out.  A one-rank box from `lo` to `hi` (default: lo + n*d per axis)."""
    x0, y0, z0 = (float(v) for v in lo)
    if hi is None:
        hi = (x0 + g.nx * float(g.dx), y0 + g.ny * float(g.dy),
              z0 + g.nz * float(g.dz))
    x1, y1, z1 = (float(v) for v in hi)
    out = np.zeros(len(x), dtype=T.PARTICLE)
    coords = []
    for r, lo, hi, n in ((x, x0, x1, g.nx), (y, y0, y1, g.ny),
                         (z, z0, z1, g.nz)):
        r = n * ((np.asarray(r, dtype=np.float64) - lo) / (hi - lo))
        i = r.astype(np.int64)
        r = r - i
        r = (r + r) - 1
        r = np.where(i == n, 1.0, r)
        i = np.where(i == n, n - 1, i)
        coords.append((i + 1, r))
    (ix, dx), (iy, dy), (iz, dz) = coords
    out["dx"], out["dy"], out["dz"] = dx, dy, dz
    out["i"] = ix + (g.nx + 2) * (iy + (g.ny + 2) * iz)
    out["ux"], out["uy"], out["uz"] = ux, uy, uz
    out["w"] = w
    return out
