"""Deterministic small dumps written by libvpichip's writers (the inputs of
the dump format fixtures).  Used by make_dump_vectors.py, which records how
the reference's own reader (scripts/binary_reader/vpic_binary_reader.py)
parses them, and by tests/test_dump.py, which rewrites them and compares."""
import ctypes
import os

import numpy as np

from vpic_amd import _lib
from vpic_amd import types as T

STEP = 168
ORIGIN = (-1.5, 0.25, 2.0)
SP_ID, Q_M = 1, -1.0 / 25.0


def grid():
    g = T.grid_periodic(ORIGIN[0], ORIGIN[1], ORIGIN[2], ORIGIN[0] + 3.0,
                        ORIGIN[1] + 2.0, ORIGIN[2] + 2.0, 3, 2, 4, 0.25)
    g.fbc[0] = g.fbc[3] = T.BC_PEC_FIELDS
    return g


def info():
    return T.DumpInfo(STEP, ORIGIN[0], ORIGIN[1], ORIGIN[2], 0, 1)


def arrays(g):
    """Field, hydro and particle arrays with distinct, exact values."""
    f = np.zeros(g.nv, dtype=T.FIELD)
    for k, c in enumerate(T.FIELD_FLOATS):
        f[c] = np.arange(g.nv, dtype=np.float32) * 0.5 + k
    f["cmat"] = 3
    h = np.zeros(g.nv, dtype=T.HYDRO)
    for k, c in enumerate(T.HYDRO_FIELDS):
        h[c] = -np.arange(g.nv, dtype=np.float32) * 0.25 - k
    p = np.zeros(7, dtype=T.PARTICLE)
    for k, c in enumerate(("dx", "dy", "dz", "ux", "uy", "uz", "w")):
        p[c] = (np.arange(7, dtype=np.float32) - 3) * 0.125 + k
    p["i"] = g.voxel(1, 1, 1) + np.arange(7, dtype=np.int32)
    return f, h, p


# The strided dumps (field_dump / hydro_dump): name -> (array, stride,
# format, output_vars).  Bit 23 of a field mask is the fourth word of the
# next voxel's record, as the reference reads it.
STRIDED = {
    "field_band": ("field", (1, 1, 1), T.DUMP_BAND, 0xffffffff),
    "field_band_s": ("field", (3, 1, 2), T.DUMP_BAND, 0x00808007),
    "field_il": ("field", (1, 1, 1), T.DUMP_BAND_INTERLEAVE, 0xffffffff),
    "field_il_s": ("field", (1, 2, 2), T.DUMP_BAND_INTERLEAVE, 0xffffffff),
    "hydro_band_s": ("hydro", (3, 2, 1), T.DUMP_BAND, 0x00003fff),
    "hydro_il": ("hydro", (1, 1, 1), T.DUMP_BAND_INTERLEAVE, 0xffffffff),
    "hydro_il_s": ("hydro", (3, 1, 2), T.DUMP_BAND_INTERLEAVE, 0xffffffff),
}
# global_header's dump parameters: the fields, then two species
HEADER_MASKS = (0xffffffff, 0x00003fff, 0x00000009)


def strided_name(outdir, name):
    return os.path.join(outdir, name, "T.%d" % STEP, "%s.%d.0" % (name, STEP))


def write_strided(outdir):
    """Writes the STRIDED dumps and the global header; returns their file
    names."""
    g = grid()
    f, h, _ = arrays(g)
    inf = ctypes.byref(info())
    G = ctypes.byref(g)
    names = {}
    for name, (kind, stride, fmt, mask) in STRIDED.items():
        base = os.path.join(outdir, name)
        os.makedirs(base, exist_ok=True)
        prm = T.dump_params(base, name, stride, fmt, mask)
        if kind == "field":
            _lib.call("vpic_hip_field_dump", f.ctypes.data, G, inf,
                      ctypes.byref(prm), -1)
        else:
            _lib.call("vpic_hip_hydro_dump", h.ctypes.data, G, inf, SP_ID,
                      Q_M, ctypes.byref(prm), -1)
        names[name] = strided_name(outdir, name)
    prms = (T.DumpParams * 3)(*[
        T.dump_params("fields_dir" if k == 0 else "hydro_dir",
                      ("fields", "ehydro", "Hhydro")[k], (1, 1, 1),
                      T.DUMP_BAND, m) for k, m in enumerate(HEADER_MASKS)])
    gi = T.GlobalInfo(ORIGIN[0], ORIGIN[0] + 3.0, ORIGIN[1], ORIGIN[1] + 2.0,
                      ORIGIN[2], ORIGIN[2] + 2.0, 1, 1, 1, 0)
    base = os.path.join(outdir, "global")
    _lib.call("vpic_hip_global_header", base.encode(), G, ctypes.byref(gi),
              prms, 3)
    names["global"] = base + ".vpc"
    return names


def write_all(outdir):
    """Writes grid, fields, hydro, particles and energies dumps; returns
    their file names."""
    g = grid()
    f, h, p = arrays(g)
    names = {k: os.path.join(outdir, k) for k in
             ("grid", "fields", "hydro", "particles", "energies")}
    inf = ctypes.byref(info())
    G = ctypes.byref(g)
    bc = np.zeros(27, dtype=np.int32)
    bc[12] = bc[14] = T.BC_PEC_FIELDS
    rng = np.array([0, g.nv], dtype=np.int64)
    nb = np.zeros(6 * g.nv, dtype=np.int64)
    _lib.call("vpic_hip_grid_neighbor_table", G, nb.ctypes.data)
    _lib.call("vpic_hip_write_grid", names["grid"].encode(), G, inf,
              bc.ctypes.data, rng.ctypes.data, nb.ctypes.data)
    _lib.call("vpic_hip_write_fields", names["fields"].encode(),
              f.ctypes.data, G, inf)
    _lib.call("vpic_hip_write_hydro", names["hydro"].encode(), h.ctypes.data,
              G, inf, SP_ID, Q_M)
    _lib.call("vpic_hip_write_particles", names["particles"].encode(),
              p.ctypes.data, len(p), len(p), G, inf, SP_ID, Q_M)
    en_f = (ctypes.c_double * 6)(1.5e-7, 2.25, 0.0, 1e10, -3.5, 7.125e-3)
    en_p = (ctypes.c_double * 2)(119.1102, 0.0650305)
    sp = (ctypes.c_char_p * 2)(b"electron", b"ion")
    _lib.call("vpic_hip_write_energies", names["energies"].encode(), 0, inf,
              g.dt, en_f, en_p, sp, 2)
    return names
