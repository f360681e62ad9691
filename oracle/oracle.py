"""ctypes front end of the oracle (TEST INFRASTRUCTURE ONLY).

`Oracle` wraps liboracle.so, the CPU restatement in oracle.c.  `Ref` wraps
_ref/libvpicref.so, the reference's own scalar kernels compiled from the
lanl/vpic sources (present only where those sources were available when it
was built).  Only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline may import this module, and only as the checker.

All arrays are numpy arrays with the dtypes of vpic_amd.types; they are
modified in place exactly as the reference functions modify theirs.
"""
import ctypes
import os

import numpy as np

from vpic_amd import types as T

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libvpicref.so")

_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_f32 = ctypes.c_float
_pg = ctypes.POINTER(T.Grid)


def _ptr(a):
    return None if a is None else a.ctypes.data_as(_vp)


def neighbor_table(g):
    """Host neighbor table of size_grid + join_grid (6*nv int64)."""
    nb = np.empty(6 * g.nv, dtype=np.int64)
    Oracle.get().lib.orc_neighbor_table(ctypes.byref(g), _ptr(nb))
    return nb


class _GridWithTable:
    """A grid copy whose `neighbor` points at a host table kept alive here."""

    def __init__(self, g):
        self.g = g.copy()
        self.nb = neighbor_table(g)
        self.g.neighbor = self.nb.ctypes.data


def host_grid(g):
    return _GridWithTable(g)


class Oracle:
    _inst = None

    def __init__(self, path=ORACLE_SO):
        if not os.path.exists(path):
            raise RuntimeError("oracle library not built: %s" % path)
        lib = ctypes.CDLL(path)
        self.lib = lib
        lib.orc_neighbor_table.argtypes = [_pg, _vp]
        lib.orc_move_p.argtypes = [_vp, _vp, _vp, _pg, _f32]
        lib.orc_move_p.restype = ctypes.c_int
        lib.orc_advance_p.argtypes = [_vp, _i32, _f32, _f32, _vp, _i32, _vp,
                                      _vp, _vp, _pg]
        lib.orc_advance_p.restype = ctypes.c_int
        lib.orc_advance_p_threaded.argtypes = [_vp, _i32, _f32, _f32, _vp,
                                               _i32, _vp, _vp, _pg, _i32]
        lib.orc_advance_p_threaded.restype = ctypes.c_int
        lib.orc_load_interpolator.argtypes = [_vp, _vp, _pg]
        lib.orc_unload_accumulator.argtypes = [_vp, _vp, _pg]
        lib.orc_clear_jf.argtypes = [_vp, _pg]
        lib.orc_clear_rhof.argtypes = [_vp, _pg]
        lib.orc_synchronize_jf.argtypes = [_vp, _pg]
        lib.orc_advance_b.argtypes = [_vp, _pg, _f32]
        lib.orc_vacuum_advance_e.argtypes = [_vp, _pg,
                                             ctypes.POINTER(T.FieldParams)]
        lib.orc_sort_p.argtypes = [_vp, _vp, _i32, _vp, _pg]
        lib.orc_energy_p.argtypes = [_vp, _i32, _f32, _f32, _vp, _pg]
        lib.orc_center_p.argtypes = [_vp, _i32, _f32, _f32, _vp, _pg,
                                     ctypes.c_int]
        lib.orc_center_p.restype = None
        lib.orc_energy_p.restype = ctypes.c_double
        lib.orc_energy_f.argtypes = [_vp, _vp, _pg,
                                     ctypes.POINTER(T.FieldParams)]
        lib.orc_accumulate_rho_p.argtypes = [_vp, _vp, _i32, _f32, _pg]
        lib.orc_synchronize_rho.argtypes = [_vp, _pg]
        for name in ("orc_compute_div_e_err", "orc_compute_rhob",
                     "orc_clean_div_e", "orc_compute_curl_b"):
            getattr(lib, name).argtypes = [_vp, _pg,
                                           ctypes.POINTER(T.FieldParams)]
            getattr(lib, name).restype = None
        lib.orc_accumulate_hydro_p.argtypes = [_vp, _vp, _i32, _f32, _f32,
                                               _vp, _pg]
        lib.orc_accumulate_hydro_p.restype = None
        lib.orc_accumulate_rhob.argtypes = [_vp, _vp, _f32, _pg]
        lib.orc_accumulate_rhob.restype = None
        lib.orc_boundary_p.argtypes = [_vp, _i32, _vp, _i32, _vp, _f32, _pg,
                                       _vp, _vp]
        lib.orc_boundary_p.restype = ctypes.c_int
        lib.orc_inject_p.argtypes = [_vp, _vp, _i32, _vp, _vp, _i32, _vp,
                                     _i32, _vp, _f32, _pg, _vp]
        lib.orc_inject_p.restype = None
        lib.orc_philox4x32.argtypes = [_vp, _vp, _vp]
        lib.orc_philox4x32.restype = None
        lib.orc_takizuka_abe.argtypes = [_vp, _vp, _f32, _f32, _vp, _vp,
                                         _f32, _f32, ctypes.c_double, _i32,
                                         ctypes.c_uint64, ctypes.c_uint64,
                                         _pg]
        lib.orc_takizuka_abe.restype = None
        lib.orc_clear_hydro.argtypes = [_vp, _pg]
        lib.orc_clear_hydro.restype = None
        lib.orc_synchronize_hydro.argtypes = [_vp, _pg]
        lib.orc_synchronize_hydro.restype = None
        for name in ("orc_compute_div_b_err", "orc_clean_div_b"):
            getattr(lib, name).argtypes = [_vp, _pg]
            getattr(lib, name).restype = None
        for name in ("orc_compute_rms_div_e_err", "orc_compute_rms_div_b_err",
                     "orc_synchronize_tang_e_norm_b"):
            getattr(lib, name).argtypes = [_vp, _pg]
            getattr(lib, name).restype = ctypes.c_double

    @classmethod
    def get(cls):
        if cls._inst is None:
            cls._inst = Oracle()
        return cls._inst

    # --- species advance ---------------------------------------------------
    def advance_p(self, p, q, m, acc, interp, hg, max_nm=None):
        """Returns (nm, n_ignored, movers[nm])."""
        if max_nm is None:
            max_nm = max(len(p), 16)
        pm = np.zeros(max(max_nm, 1), dtype=T.MOVER)
        ign = np.zeros(1, dtype=np.int32)
        nm = self.lib.orc_advance_p(_ptr(p), len(p), q, m, _ptr(pm), max_nm,
                                    _ptr(ign), _ptr(acc), _ptr(interp),
                                    ctypes.byref(hg.g))
        return nm, int(ign[0]), pm[:nm].copy()

    def advance_p_threaded(self, p, q, m, acc_copies, interp, hg, n_thread,
                           max_nm):
        pm = np.zeros(max(max_nm, 1), dtype=T.MOVER)
        return self.lib.orc_advance_p_threaded(
            _ptr(p), len(p), q, m, _ptr(pm), max_nm, _ptr(acc_copies),
            _ptr(interp), ctypes.byref(hg.g), n_thread)

    def move_p(self, p, mover, acc, hg, q):
        """mover: 1-element MOVER array, modified in place."""
        return self.lib.orc_move_p(_ptr(p), _ptr(mover), _ptr(acc),
                                   ctypes.byref(hg.g), q)

    def sort_p(self, p, g):
        aux = np.empty_like(p)
        part = np.zeros(g.nv + 1, dtype=np.int32)
        self.lib.orc_sort_p(_ptr(p), _ptr(aux), len(p), _ptr(part),
                            ctypes.byref(g))
        return part

    def center_p(self, p, q, m, interp, g, uncenter=False):
        self.lib.orc_center_p(_ptr(p), len(p), q, m, _ptr(interp),
                              ctypes.byref(g), int(uncenter))

    def energy_p(self, p, q, m, interp, g):
        return self.lib.orc_energy_p(_ptr(p), len(p), q, m, _ptr(interp),
                                     ctypes.byref(g))

    def accumulate_rho_p(self, f, p, q, g):
        self.lib.orc_accumulate_rho_p(_ptr(f), _ptr(p), len(p), q,
                                      ctypes.byref(g))

    # --- fields -------------------------------------------------------------
    def load_interpolator(self, fi, f, g):
        self.lib.orc_load_interpolator(_ptr(fi), _ptr(f), ctypes.byref(g))

    def unload_accumulator(self, f, acc, g):
        self.lib.orc_unload_accumulator(_ptr(f), _ptr(acc), ctypes.byref(g))

    def clear_jf(self, f, g):
        self.lib.orc_clear_jf(_ptr(f), ctypes.byref(g))

    def clear_rhof(self, f, g):
        self.lib.orc_clear_rhof(_ptr(f), ctypes.byref(g))

    def synchronize_jf(self, f, g):
        self.lib.orc_synchronize_jf(_ptr(f), ctypes.byref(g))

    def synchronize_rho(self, f, g):
        self.lib.orc_synchronize_rho(_ptr(f), ctypes.byref(g))

    def advance_b(self, f, g, frac):
        self.lib.orc_advance_b(_ptr(f), ctypes.byref(g), frac)

    def vacuum_advance_e(self, f, g, mp):
        self.lib.orc_vacuum_advance_e(_ptr(f), ctypes.byref(g),
                                      ctypes.byref(mp))

    def energy_f(self, f, g, mp):
        en = np.zeros(6, dtype=np.float64)
        self.lib.orc_energy_f(_ptr(en), _ptr(f), ctypes.byref(g),
                              ctypes.byref(mp))
        return en

    # hydro moments
    def accumulate_hydro_p(self, h, p, q, m, fi, g):
        self.lib.orc_accumulate_hydro_p(_ptr(h), _ptr(p), len(p), q, m,
                                        _ptr(fi), ctypes.byref(g))

    # particle boundaries
    def accumulate_rhob(self, f, p, q, hg):
        """accumulate_rhob of each particle of p in index order."""
        for k in range(len(p)):
            self.lib.orc_accumulate_rhob(_ptr(f), _ptr(p[k:k + 1]), q,
                                         ctypes.byref(hg.g))

    def boundary_p(self, p, np_, pm, f, q, hg):
        """boundary_p's mover loop: returns (new np, hits, counts)."""
        hits = np.zeros(max(len(pm), 1), dtype=T.BOUNDARY_HIT)
        counts = np.zeros(2, dtype=np.int32)
        n = self.lib.orc_boundary_p(_ptr(p), np_, _ptr(pm), len(pm), _ptr(f),
                                    q, ctypes.byref(hg.g), _ptr(hits),
                                    _ptr(counts))
        return n, hits[:len(pm)], counts

    def philox4x32(self, ctr, key):
        c = np.asarray(ctr, dtype=np.uint32)
        k = np.asarray(key, dtype=np.uint32)
        out = np.zeros(4, dtype=np.uint32)
        self.lib.orc_philox4x32(_ptr(c), _ptr(k), _ptr(out))
        return out

    def takizuka_abe(self, pi, part_i, qi, mi, pj, part_j, qj, mj, cvar0,
                     interval, seed, call, g):
        """Collisions of species i with j (pj is pi for one species), in
        place; part_* are the sort partitions (int32, nv+1)."""
        same = pj is pi
        self.lib.orc_takizuka_abe(_ptr(pi), _ptr(part_i), qi, mi,
                                  _ptr(pi if same else pj),
                                  _ptr(part_i if same else part_j), qj, mj,
                                  cvar0, interval, seed, call,
                                  ctypes.byref(g))

    def inject_p(self, p, np_, pm, nm, pi, acc, q, hg):
        """Injection loop: returns (np, nm, counts); p and pm are the
        capacity."""
        npc = np.array([np_], dtype=np.int32)
        nmc = np.array([nm], dtype=np.int32)
        counts = np.zeros(2, dtype=np.int32)
        self.lib.orc_inject_p(_ptr(p), _ptr(npc), len(p), _ptr(pm), _ptr(nmc),
                              len(pm), _ptr(pi), len(pi), _ptr(acc), q,
                              ctypes.byref(hg.g), _ptr(counts))
        return int(npc[0]), int(nmc[0]), counts

    def synchronize_hydro(self, h, g):
        self.lib.orc_synchronize_hydro(_ptr(h), ctypes.byref(g))

    # divergence cleaning / initialization kernels of the vacuum table
    def compute_div_e_err(self, f, g, mp):
        self.lib.orc_compute_div_e_err(_ptr(f), ctypes.byref(g),
                                       ctypes.byref(mp))

    def compute_rhob(self, f, g, mp):
        self.lib.orc_compute_rhob(_ptr(f), ctypes.byref(g), ctypes.byref(mp))

    def clean_div_e(self, f, g, mp):
        self.lib.orc_clean_div_e(_ptr(f), ctypes.byref(g), ctypes.byref(mp))

    def compute_curl_b(self, f, g, mp):
        self.lib.orc_compute_curl_b(_ptr(f), ctypes.byref(g),
                                    ctypes.byref(mp))

    def compute_div_b_err(self, f, g):
        self.lib.orc_compute_div_b_err(_ptr(f), ctypes.byref(g))

    def clean_div_b(self, f, g):
        self.lib.orc_clean_div_b(_ptr(f), ctypes.byref(g))

    def compute_rms_div_e_err(self, f, g):
        return self.lib.orc_compute_rms_div_e_err(_ptr(f), ctypes.byref(g))

    def compute_rms_div_b_err(self, f, g):
        return self.lib.orc_compute_rms_div_b_err(_ptr(f), ctypes.byref(g))

    def synchronize_tang_e_norm_b(self, f, g):
        return self.lib.orc_synchronize_tang_e_norm_b(_ptr(f),
                                                      ctypes.byref(g))


class Ref:
    """The reference's own scalar kernels (oracle/_ref)."""
    _inst = None

    def __init__(self, path=REF_SO):
        if not os.path.exists(path):
            raise RuntimeError("reference build absent: %s" % path)
        lib = ctypes.CDLL(path)
        self.lib = lib
        lib.ref_advance_p.argtypes = [_vp, ctypes.c_int, _f32, _f32, _vp,
                                      ctypes.c_int, _vp, _vp, _vp, _pg,
                                      ctypes.c_int]
        lib.ref_advance_p.restype = ctypes.c_int
        lib.ref_advance_p_pipelines.argtypes = [_vp, ctypes.c_int, _f32, _f32,
                                                _vp, ctypes.c_int, _vp, _vp,
                                                _pg, ctypes.c_int]
        lib.ref_advance_p_pipelines.restype = ctypes.c_int
        lib.ref_move_p.argtypes = [_vp, _vp, _vp, _pg, _f32]
        lib.ref_move_p.restype = ctypes.c_int
        lib.ref_energy_p.argtypes = [_vp, ctypes.c_int, _f32, _f32, _vp, _pg]
        lib.ref_energy_p.restype = ctypes.c_double
        lib.ref_center_p.argtypes = [_vp, ctypes.c_int, _f32, _f32, _vp, _pg,
                                     ctypes.c_int]
        lib.ref_center_p.restype = None
        for name in ("ref_div_e_err_bulk", "ref_rhob_bulk",
                     "ref_clean_div_e_bulk", "ref_curl_b_bulk"):
            getattr(lib, name).argtypes = [_vp, _pg,
                                           ctypes.POINTER(T.FieldParams)]
            getattr(lib, name).restype = None
        for name in ("ref_div_b_err", "ref_clean_div_b_bulk"):
            getattr(lib, name).argtypes = [_vp, _pg]
            getattr(lib, name).restype = None
        lib.ref_accumulate_hydro_p.argtypes = [_vp, _vp, ctypes.c_int, _f32,
                                               _f32, _vp, _pg]
        lib.ref_accumulate_hydro_p.restype = None
        lib.ref_accumulate_rhob.argtypes = [_vp, _vp, ctypes.c_int, _f32, _pg]
        lib.ref_accumulate_rhob.restype = None
        for name in ("ref_rms_div_b_err", "ref_rms_div_e_err_bulk"):
            getattr(lib, name).argtypes = [_vp, _pg]
            getattr(lib, name).restype = ctypes.c_double
        lib.ref_sort_p.argtypes = [_vp, ctypes.c_int, _vp, _pg]
        lib.ref_takizuka_abe.argtypes = [_vp, _vp, _f32, _f32, _vp, _vp,
                                         _f32, _f32, ctypes.c_double, _i32,
                                         ctypes.c_uint64, ctypes.c_uint64,
                                         _pg]
        lib.ref_takizuka_abe.restype = None
        lib.ref_load_interpolator.argtypes = [_vp, _vp, _pg]
        lib.ref_unload_accumulator.argtypes = [_vp, _vp, _pg]
        lib.ref_advance_b_bulk.argtypes = [_vp, _pg, _f32]
        lib.ref_vacuum_advance_e_bulk.argtypes = [
            _vp, _pg, ctypes.POINTER(T.FieldParams)]
        lib.ref_energy_f.argtypes = [_vp, _vp, _pg,
                                     ctypes.POINTER(T.FieldParams)]

    @classmethod
    def available(cls):
        return os.path.exists(REF_SO)

    @classmethod
    def get(cls):
        if cls._inst is None:
            cls._inst = Ref()
        return cls._inst

    def advance_p(self, p, q, m, acc, interp, hg, max_nm=None, n_pipeline=1):
        if max_nm is None:
            max_nm = max(len(p), 16)
        pm = np.zeros(max(max_nm, 1), dtype=T.MOVER)
        ign = np.zeros(1, dtype=np.int32)
        nm = self.lib.ref_advance_p(_ptr(p), len(p), q, m, _ptr(pm), max_nm,
                                    _ptr(ign), _ptr(acc), _ptr(interp),
                                    ctypes.byref(hg.g), n_pipeline)
        return nm, int(ign[0]), pm[:nm].copy()

    def advance_p_pipelines(self, p, q, m, acc_copies, interp, hg, n_pipeline,
                            pm):
        """Timed part of the reference's advance_p on n_pipeline threads;
        acc_copies holds (n_pipeline+1)*POW2_CEIL(nv,2) accumulators."""
        return self.lib.ref_advance_p_pipelines(
            _ptr(p), len(p), q, m, _ptr(pm), len(pm), _ptr(acc_copies),
            _ptr(interp), ctypes.byref(hg.g), n_pipeline)

    def move_p(self, p, mover, acc, hg, q):
        return self.lib.ref_move_p(_ptr(p), _ptr(mover), _ptr(acc),
                                   ctypes.byref(hg.g), q)

    def center_p(self, p, q, m, interp, g, uncenter=False):
        self.lib.ref_center_p(_ptr(p), len(p), q, m, _ptr(interp),
                              ctypes.byref(g), int(uncenter))

    def energy_p(self, p, q, m, interp, g):
        return self.lib.ref_energy_p(_ptr(p), len(p), q, m, _ptr(interp),
                                     ctypes.byref(g))

    def sort_p(self, p, g):
        part = np.zeros(g.nv + 1, dtype=np.int32)
        self.lib.ref_sort_p(_ptr(p), len(p), _ptr(part), ctypes.byref(g))
        return part

    def takizuka_abe(self, pi, part_i, qi, mi, pj, part_j, qj, mj, cvar0,
                     interval, seed, call, g):
        """The reference's takizuka_abe_pipeline_scalar on the oracle's
        per-voxel Philox stream (oracle/ref/ref_ta.cc), in place."""
        same = pj is pi
        self.lib.ref_takizuka_abe(_ptr(pi), _ptr(part_i), qi, mi,
                                  _ptr(pi if same else pj),
                                  _ptr(part_i if same else part_j), qj, mj,
                                  cvar0, interval, seed, call,
                                  ctypes.byref(g))

    def load_interpolator(self, fi, f, g):
        self.lib.ref_load_interpolator(_ptr(fi), _ptr(f), ctypes.byref(g))

    def unload_accumulator(self, f, acc, g):
        self.lib.ref_unload_accumulator(_ptr(f), _ptr(acc), ctypes.byref(g))

    def advance_b_bulk(self, f, g, frac):
        self.lib.ref_advance_b_bulk(_ptr(f), ctypes.byref(g), frac)

    def vacuum_advance_e_bulk(self, f, g, mp):
        self.lib.ref_vacuum_advance_e_bulk(_ptr(f), ctypes.byref(g),
                                           ctypes.byref(mp))

    def energy_f(self, f, g, mp):
        en = np.zeros(6, dtype=np.float64)
        self.lib.ref_energy_f(_ptr(en), _ptr(f), ctypes.byref(g),
                              ctypes.byref(mp))
        return en

    # bulk (pipelined interior) parts of the divergence-cleaning kernels
    def div_e_err_bulk(self, f, g, mp):
        self.lib.ref_div_e_err_bulk(_ptr(f), ctypes.byref(g), ctypes.byref(mp))

    def rhob_bulk(self, f, g, mp):
        self.lib.ref_rhob_bulk(_ptr(f), ctypes.byref(g), ctypes.byref(mp))

    def clean_div_e_bulk(self, f, g, mp):
        self.lib.ref_clean_div_e_bulk(_ptr(f), ctypes.byref(g),
                                      ctypes.byref(mp))

    def curl_b_bulk(self, f, g, mp):
        self.lib.ref_curl_b_bulk(_ptr(f), ctypes.byref(g), ctypes.byref(mp))

    def div_b_err(self, f, g):
        self.lib.ref_div_b_err(_ptr(f), ctypes.byref(g))

    def clean_div_b_bulk(self, f, g):
        self.lib.ref_clean_div_b_bulk(_ptr(f), ctypes.byref(g))

    def rms_div_b_err(self, f, g):
        return self.lib.ref_rms_div_b_err(_ptr(f), ctypes.byref(g))

    def rms_div_e_err_bulk(self, f, g):
        return self.lib.ref_rms_div_e_err_bulk(_ptr(f), ctypes.byref(g))

    def accumulate_hydro_p(self, h, p, q, m, fi, g):
        self.lib.ref_accumulate_hydro_p(_ptr(h), _ptr(p), len(p), q, m,
                                        _ptr(fi), ctypes.byref(g))

    def accumulate_rhob(self, f, p, q, hg):
        """The reference's accumulate_rhob of each particle in index order."""
        self.lib.ref_accumulate_rhob(_ptr(f), _ptr(p), len(p), q,
                                     ctypes.byref(hg.g))
