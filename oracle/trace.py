"""The oracle's step loop, FP32 or FP64 (TEST INFRASTRUCTURE ONLY).

`OracleRun` runs vpic_simulation::advance (src/vpic/advance.cc:15-207)
restricted to the hot path -- clear_accumulator_array, advance_p of every
species, unload_accumulator_array with clear_jf / synchronize_jf,
advance_b(0.5), vacuum_advance_e, advance_b(0.5), load_interpolator_array --
on the oracle's CPU restatement, in FP32 (liboracle's orc_* functions, which
are pinned bit for bit against the reference's scalar kernels) or in FP64
(orc64_*, the same source compiled over double).  The FP64 run is the
reference against which the device's FP32 energy and charge traces are
bounded; its initial state is the FP32 state widened exactly, so the two
runs differ only in their arithmetic.

The set-up follows vpic_simulation::initialize (src/vpic/initialize.cc:
33-60) for what the hot path needs: accumulate_rho_p of every species,
synchronize_rho, compute_rhob, load_interpolator_array, uncenter_p.

advance_p is run on particle blocks in threads with a private accumulator
each (as the reference's pipelines do), summed in block order.  Only
tests/ and bench.py's cpu_baseline may import this module.
"""
import ctypes
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

from vpic_amd import types as T

from .oracle import ORACLE_SO, Oracle, neighbor_table

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_f64 = ctypes.c_double


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_vp)


# FP64 twins of the reference's types (oracle/oracle.h: orc64_*).
PARTICLE64 = np.dtype({"names": ["dx", "dy", "dz", "i", "pad", "ux", "uy",
                                 "uz", "w"],
                       "formats": ["<f8", "<f8", "<f8", "<i4", "<i4", "<f8",
                                   "<f8", "<f8", "<f8"],
                       "offsets": [0, 8, 16, 24, 28, 32, 40, 48, 56],
                       "itemsize": 64})
MOVER64 = np.dtype({"names": ["dispx", "dispy", "dispz", "i", "pad"],
                    "formats": ["<f8", "<f8", "<f8", "<i4", "<i4"],
                    "offsets": [0, 8, 16, 24, 28], "itemsize": 32})
INTERPOLATOR64 = np.dtype([(n, "<f8") for n in T.INTERP_FIELDS])
FIELD64 = np.dtype([(n, "<f8") for n in T.FIELD_FLOATS] +
                   [(n, "<i2") for n in ("ematx", "ematy", "ematz", "nmat",
                                         "fmatx", "fmaty", "fmatz", "cmat")])
assert PARTICLE64.itemsize == 64 and MOVER64.itemsize == 32
assert INTERPOLATOR64.itemsize == 160 and FIELD64.itemsize == 144


class Grid64(ctypes.Structure):
    """orc64_grid_t."""
    _fields_ = [(n, ctypes.c_double) for n in
                ("dt", "cvac", "eps0", "dx", "dy", "dz", "dV", "rdx", "rdy",
                 "rdz", "r8V")] + \
               [(n, ctypes.c_int32) for n in
                ("nx", "ny", "nz", "sx", "sy", "sz", "nv")] + \
               [("rangel", ctypes.c_int64), ("rangeh", ctypes.c_int64),
                ("neighbor", ctypes.c_void_p),
                ("pbc", ctypes.c_int32 * 6), ("fbc", ctypes.c_int32 * 6)]


class FieldParams64(ctypes.Structure):
    """orc64_field_params_t."""
    _fields_ = [(n, ctypes.c_double) for n in
                ("decayx", "drivex", "decayy", "drivey", "decayz", "drivez",
                 "rmux", "rmuy", "rmuz", "epsx", "epsy", "epsz", "damp",
                 "nonconductive")]


def widen_particles(p):
    """FP32 particles as FP64 particles (exact)."""
    q = np.zeros(len(p), dtype=PARTICLE64)
    for c in ("dx", "dy", "dz", "i", "ux", "uy", "uz", "w"):
        q[c] = p[c]
    return q


def widen_fields(f):
    q = np.zeros(len(f), dtype=FIELD64)
    for c in f.dtype.names:
        q[c] = f[c]
    return q


def widen_params(mp):
    out = FieldParams64()
    for n, _ in FieldParams64._fields_:
        setattr(out, n, float(np.float32(getattr(mp, n))))
    return out


class _Lib:
    """The orc_* (FP32) or orc64_* (FP64) step functions of liboracle."""

    def __init__(self, fp64):
        lib = Oracle.get().lib if not fp64 else ctypes.CDLL(ORACLE_SO)
        pre = "orc64_" if fp64 else "orc_"
        real = _f64 if fp64 else ctypes.c_float
        pg = _vp
        sig = {
            "advance_p": ([_vp, _i32, real, real, _vp, _i32, _vp, _vp, _vp,
                           pg], ctypes.c_int),
            "load_interpolator": ([_vp, _vp, pg], None),
            "unload_accumulator": ([_vp, _vp, pg], None),
            "clear_jf": ([_vp, pg], None),
            "clear_rhof": ([_vp, pg], None),
            "synchronize_jf": ([_vp, pg], None),
            "synchronize_rho": ([_vp, pg], None),
            "advance_b": ([_vp, pg, real], None),
            "vacuum_advance_e": ([_vp, pg, _vp], None),
            "energy_p": ([_vp, _i32, real, real, _vp, pg], _f64),
            "energy_f": ([_vp, _vp, pg, _vp], None),
            "accumulate_rho_p": ([_vp, _vp, _i32, real, pg], None),
            "compute_div_e_err": ([_vp, pg, _vp], None),
            "compute_rhob": ([_vp, pg, _vp], None),
            "compute_rms_div_e_err": ([_vp, pg], _f64),
            "center_p": ([_vp, _i32, real, real, _vp, pg, ctypes.c_int],
                         None),
        }
        for name, (args, res) in sig.items():
            fn = getattr(lib, pre + name)
            fn.argtypes = args
            fn.restype = res
            setattr(self, name, fn)
        if fp64:
            lib.orc64_grid_from.argtypes = [_vp, _vp]
            lib.orc64_grid_from.restype = None
            self.grid_from = lib.orc64_grid_from


class OracleRun:
    """One simulation on the oracle.

    g: the FP32 grid (vpic_amd.types.Grid, its pbc/fbc set);
    species: [(name, q, m, FP32 particles)]; fields: an optional FP32
    field array (initial E and cB); mp: FieldParams (vacuum).
    """

    def __init__(self, g, species, fp64=False, fields=None, mp=None,
                 threads=None):
        self.fp64 = fp64
        self.L = _Lib(fp64)
        self.g32 = g.copy()
        self.nb = neighbor_table(g)
        self.g32.neighbor = self.nb.ctypes.data
        mp = mp or T.vacuum_params()
        if fp64:
            self.g = Grid64()
            self.L.grid_from(ctypes.byref(self.g), ctypes.byref(self.g32))
            self.mp = widen_params(mp)
            self.f = widen_fields(fields) if fields is not None else \
                np.zeros(g.nv, dtype=FIELD64)
            self.fi = np.zeros(g.nv, dtype=INTERPOLATOR64)
            self.acc_dtype = np.float64
            self.species = [(n, float(q), float(m), widen_particles(p))
                            for n, q, m, p in species]
        else:
            self.g = self.g32
            self.mp = mp
            self.f = fields.copy() if fields is not None else \
                np.zeros(g.nv, dtype=T.FIELD)
            self.fi = np.zeros(g.nv, dtype=T.INTERPOLATOR)
            self.acc_dtype = np.float32
            self.species = [(n, float(np.float32(q)), float(np.float32(m)),
                             p.copy()) for n, q, m, p in species]
        self.G = ctypes.byref(self.g)
        self.MP = ctypes.byref(self.mp)
        self.nv = g.nv
        self.acc = np.zeros((self.nv, 12), dtype=self.acc_dtype)
        self.threads = threads or min(16, os.cpu_count() or 1)
        self.pool = ThreadPoolExecutor(self.threads) if self.threads > 1 \
            else None
        self.step = 0

    # -- the reference-named operations -----------------------------------
    def _blocks(self, n):
        """Contiguous particle blocks, multiples of 16 long (the
        reference's DISTRIBUTE(np, 16) granularity)."""
        k = max(1, min(self.threads, n // 4096))
        edges = [16 * ((n // 16) * j // k) for j in range(k)] + [n]
        return [(edges[j], edges[j + 1]) for j in range(k)
                if edges[j + 1] > edges[j]]

    def advance_p(self, p, q, m):
        blocks = self._blocks(len(p))
        accs = [np.zeros_like(self.acc) for _ in blocks]
        mdt = MOVER64 if self.fp64 else T.MOVER

        def run(j):
            a, b = blocks[j]
            pm = np.zeros(max(b - a, 1), dtype=mdt)
            sub = p[a:b]
            nm = self.L.advance_p(_ptr(sub), b - a, q, m, _ptr(pm), b - a,
                                  None, _ptr(accs[j]), _ptr(self.fi), self.G)
            return nm

        if self.pool and len(blocks) > 1:
            nms = list(self.pool.map(run, range(len(blocks))))
        else:
            nms = [run(j) for j in range(len(blocks))]
        if sum(nms):
            raise RuntimeError("oracle run: %d movers (this driver handles "
                               "periodic and reflecting faces only)" % sum(nms))
        for a in accs:
            self.acc += a

    def load_interpolator(self):
        self.L.load_interpolator(_ptr(self.fi), _ptr(self.f), self.G)

    def rho(self):
        """rhof of every species (accumulate_rho_p + synchronize_rho)."""
        self.L.clear_rhof(_ptr(self.f), self.G)
        for _, q, _, p in self.species:
            self.L.accumulate_rho_p(_ptr(self.f), _ptr(p), len(p), q, self.G)
        self.L.synchronize_rho(_ptr(self.f), self.G)

    def initialize(self):
        """The charge and interpolator set-up of initialize.cc:33-60:
        rhob from the initial E and charge, the interpolators, and the
        momenta uncentred to t = -dt/2."""
        self.rho()
        self.L.compute_rhob(_ptr(self.f), self.G, self.MP)
        self.load_interpolator()
        for _, q, m, p in self.species:
            self.L.center_p(_ptr(p), len(p), q, m, _ptr(self.fi), self.G, 1)

    def advance(self):
        """One step (advance.cc:25-185, hot path)."""
        self.acc[...] = 0
        for _, q, m, p in self.species:
            self.advance_p(p, q, m)
        self.L.clear_jf(_ptr(self.f), self.G)
        self.L.unload_accumulator(_ptr(self.f), _ptr(self.acc), self.G)
        self.L.synchronize_jf(_ptr(self.f), self.G)
        self.L.advance_b(_ptr(self.f), self.G, 0.5)
        self.L.vacuum_advance_e(_ptr(self.f), self.G, self.MP)
        self.L.advance_b(_ptr(self.f), self.G, 0.5)
        self.load_interpolator()
        self.step += 1

    # -- diagnostics --------------------------------------------------------
    def energies(self):
        """(energy_f[6], [energy_p per species]) as dump_energies forms
        them (dump.cc:38-77), FP64 sums."""
        en = np.zeros(6, dtype=np.float64)
        self.L.energy_f(_ptr(en), _ptr(self.f), self.G, self.MP)
        ep = [self.L.energy_p(_ptr(p), len(p), q, m, _ptr(self.fi), self.G)
              for _, q, m, p in self.species]
        return en, ep

    def div_e_rms(self):
        """compute_rms_div_e_err after accumulate_rho_p + synchronize_rho
        + compute_div_e_err: the charge-conservation diagnostic
        (advance.cc:143-160)."""
        self.rho()
        self.L.compute_div_e_err(_ptr(self.f), self.G, self.MP)
        return self.L.compute_rms_div_e_err(_ptr(self.f), self.G)

    def close(self):
        if self.pool:
            self.pool.shutdown()
