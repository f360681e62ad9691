"""The reference's output formats (src/vpic/dump.cc, dumpmacros.h).

Pins: the energies line format against the reference's own golden file
(test/unit/energy_comparison/energies_gold.3d_test, copied under
tests/golden/): rewriting its numbers reproduces it byte for byte; the
binary dumps against how the reference's own reader parses them
(tests/golden/dump_vectors.json, made by tests/golden/make_dump_vectors.py):
the files are rewritten identically and the reader's values are the ones
written.  The GPU test checks the Simulation's dump_* against the oracle
(particles time-centred with center_p, hydro moments, fields).
"""
import ctypes
import hashlib
import json
import os
import sys

import numpy as np
import pytest

from vpic_amd import _lib
from vpic_amd import types as T

import helpers as H

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
sys.path.insert(0, GOLD)
import dump_cases  # noqa: E402


def test_energies_lines_reproduce_reference_gold(tmp_path):
    """Each line of the reference's energies file, parsed and written again
    with append = 1 (the deck's dump_energies(name, 1)), gives the same
    bytes: %li step, six %e field energies, one %e per species."""
    gold = open(os.path.join(GOLD, "energies_gold.3d_test"), "rb").read()
    out = str(tmp_path / "energies")
    for line in gold.decode().splitlines():
        v = line.split()
        info = T.DumpInfo(int(v[0]), 0.0, 0.0, 0.0, 0, 1)
        en_f = (ctypes.c_double * 6)(*[float(x) for x in v[1:7]])
        en_p = (ctypes.c_double * 2)(*[float(x) for x in v[7:9]])
        _lib.call("vpic_hip_write_energies", out.encode(), 1,
                  ctypes.byref(info), 0.5, en_f, en_p, None, 2)
    assert open(out, "rb").read() == gold


def test_energies_new_file_layout(tmp_path):
    """A new file opens with the layout comment (dump.cc:51-57)."""
    out = str(tmp_path / "energies")
    info = T.DumpInfo(0, 0.0, 0.0, 0.0, 0, 1)
    en_f = (ctypes.c_double * 6)(0, 0, 0, 0, 0, 0)
    en_p = (ctypes.c_double * 2)(1.0, 2.0)
    names = (ctypes.c_char_p * 2)(b"electron", b"ion")
    _lib.call("vpic_hip_write_energies", out.encode(), 0, ctypes.byref(info),
              np.float32(0.5485), en_f, en_p, names, 2)
    lines = open(out).read().splitlines()
    assert lines[0] == "% Layout"
    assert lines[1] == '% step ex ey ez bx by bz "electron" "ion"'
    assert lines[2] == "%% timestep = %e" % float(np.float32(0.5485))
    assert lines[3] == ("0 " + " ".join(["0.000000e+00"] * 6) +
                        " 1.000000e+00 2.000000e+00")


def test_binary_dumps_as_the_reference_reader_parses_them(tmp_path):
    vec = json.load(open(os.path.join(GOLD, "dump_vectors.json")))
    names = dump_cases.write_all(str(tmp_path))
    for kind, fn in names.items():
        digest = hashlib.sha256(open(fn, "rb").read()).hexdigest()
        assert digest == vec[kind]["sha256"], kind
    g = dump_cases.grid()
    f, h, p = dump_cases.arrays(g)
    types = {"grid": 0, "fields": 1, "hydro": 2, "particles": 3}
    for kind, t in types.items():
        hd = vec[kind]["header"]
        assert hd["dump_type"] == t
        assert hd["step()"] == dump_cases.STEP
        assert (hd["nxout"], hd["nyout"], hd["nzout"]) == (g.nx, g.ny, g.nz)
        assert hd["grid->dt"] == float(np.float32(g.dt))
        assert hd["dxout"] == float(np.float32(g.dx))
        assert (hd["grid->x0"], hd["grid->y0"], hd["grid->z0"]) == tuple(
            float(np.float32(v)) for v in dump_cases.ORIGIN)
        assert (hd["rank()"], hd["nproc()"]) == (0, 1)
        species = kind in ("hydro", "particles")
        assert hd["sp_id"] == (dump_cases.SP_ID if species else -1)
        assert hd["q_m"] == (float(np.float32(dump_cases.Q_M))
                             if species else 0.0)
    arr = {k: vec[k]["array"] for k in types}
    assert arr["grid"]["dim"] == [3, 3, 3] and arr["grid"]["sizeof(p[0])"] == 4
    assert arr["fields"]["dim"] == [g.nx + 2, g.ny + 2, g.nz + 2]
    assert arr["fields"]["sizeof(p[0])"] == T.FIELD.itemsize
    assert arr["hydro"]["sizeof(p[0])"] == T.HYDRO.itemsize
    assert arr["particles"]["dim"] == [len(p)]
    assert arr["particles"]["sizeof(p[0])"] == T.PARTICLE.itemsize
    # the payload as the reader sees it: the float components of each record
    fsum = f.view(np.float32).reshape(g.nv, -1)[:, :16].sum(axis=0,
                                                           dtype=np.float64)
    assert np.array_equal(vec["fields"]["field_sums"][:16], fsum)
    hsum = h.view(np.float32).reshape(g.nv, -1)[:, :14].sum(axis=0,
                                                           dtype=np.float64)
    assert np.array_equal(vec["hydro"]["field_sums"][:14], hsum)
    for k, rec in enumerate(vec["particles"]["particles"]):
        ux, uy, uz, w, voxel = rec[3], rec[4], rec[5], rec[6], rec[7]
        assert (ux, uy, uz, w) == (float(p["ux"][k]), float(p["uy"][k]),
                                   float(p["uz"][k]), float(p["w"][k]))
        assert voxel == int(p["i"][k])
        assert rec[11:14] == [float(p["dx"][k]), float(p["dy"][k]),
                              float(p["dz"][k])]


@pytest.mark.gpu
def test_simulation_dumps_gpu(tmp_path, orc):
    """Simulation.dump_particles writes center_p of the particles (the
    species unchanged), dump_hydro the synchronized moments, dump_fields
    the field array, dump_energies a line of the energies."""
    import torch
    from vpic_amd import sim as S
    assert torch.cuda.is_available()
    _lib.call("vpic_hip_init", 0)
    g = T.grid_periodic(0, 0, 0, 6.0, 5.0, 4.0, 6, 5, 4,
                        0.95 * T.courant_length(6.0, 5.0, 4.0, 6, 5, 4))
    s = S.Simulation(g)
    f = H.random_full_fields(g, 4, 0.1)
    f["rhob"] = 0
    s.fields.upload(f)
    s.load_interpolator_array()
    sp = S.Species("electron", -1.0, 1.0, 6 * 5 * 4 * 5, g, sort_interval=1)
    p0 = H.thermal_particles(g, 5, 0.2, 12, shuffle=True)
    sp.p.upload(p0)
    s.add_species(sp)
    s.sort_p(sp)
    p_sorted = sp.p.download().copy()
    fi = s.interp.download().copy()
    print("dumps", flush=True)
    s.dump_particles(sp, str(tmp_path / "eparticle"))
    s.dump_hydro(sp, str(tmp_path / "ehydro"))
    s.dump_fields(str(tmp_path / "fields"))
    s.dump_energies(str(tmp_path / "energies"), 0)
    assert H.same_bits(sp.p.download(), p_sorted)
    raw = open(str(tmp_path / "eparticle.0.0"), "rb").read()
    got = np.frombuffer(raw[-len(p0) * 32:], dtype=T.PARTICLE)
    want = p_sorted.copy()
    orc.center_p(want, -1.0, 1.0, fi, g)
    assert H.same_bits(got, want)
    raw = open(str(tmp_path / "ehydro.0.0"), "rb").read()
    hgot = np.frombuffer(raw[-g.nv * 64:], dtype=T.HYDRO)
    hw = np.zeros(g.nv, dtype=T.HYDRO)
    orc.accumulate_hydro_p(hw, p_sorted, -1.0, 1.0, fi, g)
    orc.synchronize_hydro(hw, g)
    for c in T.HYDRO_FIELDS:
        den = float(np.max(np.abs(hw[c]))) or 1.0
        assert float(np.max(np.abs(hgot[c] - hw[c]))) <= 1e-5 * den, c
    raw = open(str(tmp_path / "fields.0.0"), "rb").read()
    assert raw[-g.nv * 80:] == f.tobytes()
    lines = open(str(tmp_path / "energies")).read().splitlines()
    assert len(lines) == 4 and lines[3].startswith("0 ")
