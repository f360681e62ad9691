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
    # the strided dumps of the same state and their global header
    print("field_dump / hydro_dump", flush=True)
    os.makedirs(str(tmp_path / "fd"))
    os.makedirs(str(tmp_path / "hd"))
    fp = T.dump_params(str(tmp_path / "fd"), "fields", (1, 1, 1),
                       T.DUMP_BAND_INTERLEAVE)
    hp = T.dump_params(str(tmp_path / "hd"), "ehydro", (1, 1, 1),
                       T.DUMP_BAND, 0x3fff)
    s.field_dump(fp)
    s.hydro_dump(sp, hp)
    s.global_header(str(tmp_path / "global"), fp, [hp])
    raw = open(str(tmp_path / "fd" / "T.0" / "fields.0.0"), "rb").read()
    assert raw[123:] == f.tobytes()
    raw = open(str(tmp_path / "hd" / "T.0" / "ehydro.0.0"), "rb").read()
    hb = np.frombuffer(raw[123:], dtype=np.float32).reshape(14, g.nv)
    for k, c in enumerate(T.HYDRO_FIELDS[:14]):
        den = float(np.max(np.abs(hw[c]))) or 1.0
        assert float(np.max(np.abs(hb[k] - hw[c]))) <= 1e-5 * den, c
    hdr = open(str(tmp_path / "global.vpc")).read()
    assert "NUM_OUTPUT_SPECIES 1\n" in hdr and "HYDRO_DATA_VARIABLES 5\n" in hdr


def _out_index(nout, n, s, unit, ghosts=True):
    """dump.cc:643-645: the voxels field_dump / hydro_dump write along an
    axis (unit strides: all of them as they are)."""
    if unit:
        return np.arange(nout + 2 if ghosts else nout)
    idx = [0] + [k * s - 1 for k in range(1, nout + 1)] + [n + 1]
    return np.array(idx if ghosts else idx[:nout])


def _strided_expect(arr, g, stride, fmt, mask):
    """The data block of a strided dump, restated with numpy indexing."""
    rec = arr.dtype.itemsize
    words = np.frombuffer(arr.tobytes() + b"\0" * 16, dtype=np.uint32)
    n = (g.nx, g.ny, g.nz)
    nout = [n[a] // stride[a] for a in range(3)]
    unit = tuple(stride) == (1, 1, 1)
    field = rec == T.FIELD.itemsize
    nwords = 24 if field else 14
    if fmt == T.DUMP_BAND_INTERLEAVE and unit:
        nrec = g.nv if field else nout[0] * nout[1] * nout[2]
        return arr.tobytes()[:nrec * rec]
    ghosts = field or fmt == T.DUMP_BAND
    ix, iy, iz = (_out_index(nout[a], n[a], stride[a], unit, ghosts)
                  for a in range(3))
    vox = (iz[:, None, None] * (g.nx + 2) * (g.ny + 2) +
           iy[None, :, None] * (g.nx + 2) + ix[None, None, :])
    if fmt == T.DUMP_BAND_INTERLEAVE:
        return arr[vox.reshape(-1)].tobytes()
    out = []
    for v in range(nwords):
        if mask >> v & 1:
            at = vox * (rec // 4) + v
            w = words[at]
            w[at >= len(arr) * rec // 4] = 0  # past the array's last record
            out.append(w.astype(np.uint32).tobytes())
    return b"".join(out)


def test_strided_dumps_as_the_reference_reader_parses_them(tmp_path):
    """field_dump / hydro_dump (dump.cc:533-880): the files are rewritten
    byte-identically to the ones the reference's reader parsed
    (tests/golden/dump_vectors.json); the header the reader saw carries the
    strided extents and spacing, the array header the record size and the
    output extents, and the data block is the band / band_interleave
    restatement of the arrays above."""
    vec = json.load(open(os.path.join(GOLD, "dump_vectors.json")))
    names = dump_cases.write_strided(str(tmp_path))
    g = dump_cases.grid()
    f, h, _ = dump_cases.arrays(g)
    for name, (kind, stride, fmt, mask) in dump_cases.STRIDED.items():
        raw = open(names[name], "rb").read()
        ref = vec["strided_" + name]
        assert hashlib.sha256(raw).hexdigest() == ref["sha256"], name
        hd, ah = ref["header"], ref["array"]
        nout = (g.nx // stride[0], g.ny // stride[1], g.nz // stride[2])
        assert (hd["nxout"], hd["nyout"], hd["nzout"]) == nout, name
        for a, d in enumerate(("dxout", "dyout", "dzout")):
            sp = float(np.float32((g.dx, g.dy, g.dz)[a]) *
                       np.float32(stride[a]))
            assert hd[d] == sp, (name, d)
        assert hd["dump_type"] == (1 if kind == "field" else 2), name
        assert hd["sp_id"] == (-1 if kind == "field" else dump_cases.SP_ID)
        assert hd["step()"] == dump_cases.STEP
        arr = f if kind == "field" else h
        assert ah["sizeof(p[0])"] == arr.dtype.itemsize, name
        ghosts = kind == "field" or fmt == T.DUMP_BAND
        assert ah["dim"] == [x + 2 if ghosts else x for x in nout], name
        want = _strided_expect(arr, g, stride, fmt, mask)
        assert raw[123:] == want, name
        nwords = (bin(mask & ((1 << (24 if kind == "field" else 14)) - 1))
                  .count("1") if fmt == T.DUMP_BAND
                  else arr.dtype.itemsize // 4)
        assert ah["no_fields"] == nwords, name


def test_global_header(tmp_path):
    """global_header (dump.cc:405-532): the text the reference writes for
    this grid, the field variables of a full mask and each species' hydro
    variables of its mask (groups selected by their first word, names from
    the reference's tables)."""
    vec = json.load(open(os.path.join(GOLD, "dump_vectors.json")))
    names = dump_cases.write_strided(str(tmp_path))
    raw = open(names["global"], "rb").read()
    assert hashlib.sha256(raw).hexdigest() == vec["global_header"]["sha256"]
    lines = raw.decode().splitlines()
    bar = "#" * 80
    assert lines[0] == bar and lines[1] == "# Header version information"
    kv = {}
    for ln in lines:
        if ln and not ln.startswith("#") and not ln.startswith('"'):
            k, _, v = ln.partition(" ")
            kv.setdefault(k, []).append(v)
    g = dump_cases.grid()
    o = dump_cases.ORIGIN
    assert kv["VPIC_HEADER_VERSION"] == ["1.0.0"]
    assert kv["DATA_HEADER_SIZE"] == ["123"]
    assert kv["GRID_DELTA_T"] == ["%f" % g.dt]
    assert kv["GRID_EXTENTS_X"] == ["%f %f" % (o[0], o[0] + 3.0)]
    assert kv["GRID_DELTA_Z"] == ["%f" % g.dz]
    assert kv["GRID_TOPOLOGY_Y"] == ["1"]
    assert kv["NUM_OUTPUT_SPECIES"] == ["2"]
    assert kv["FIELD_DATA_VARIABLES"] == ["12"]
    assert kv["HYDRO_DATA_VARIABLES"] == ["5", "2"]
    assert kv["SPECIES_DATA_BASE_FILENAME"] == ["ehydro", "Hhydro"]
    quoted = [ln for ln in lines if ln.startswith('"')]
    assert quoted[0] == '"Electric Field" VECTOR 3 FLOATING_POINT 4'
    assert quoted[11] == '"Cell Material" SCALAR 1 INTEGER 2'
    assert quoted[-2:] == ['"Current Density" VECTOR 3 FLOATING_POINT 4',
                           '"Charge Density" SCALAR 1 FLOATING_POINT 4']
    assert not raw.endswith(b"\n\n")


def test_strided_dump_errors(tmp_path):
    g = dump_cases.grid()
    f, _, _ = dump_cases.arrays(g)
    inf = ctypes.byref(dump_cases.info())
    bad = [T.dump_params(str(tmp_path), "f", (2, 1, 1)),  # 2 does not divide 3
           T.dump_params(str(tmp_path), "f", (1, 1, 0)),
           # bit 30 without bits 1..23: the reference's list runs short
           T.dump_params(str(tmp_path), "f", (1, 1, 1), T.DUMP_BAND,
                         (1 << 30) | 1)]
    for prm in bad:
        with pytest.raises(_lib.VpicError):
            _lib.call("vpic_hip_field_dump", f.ctypes.data, ctypes.byref(g),
                      inf, ctypes.byref(prm), -1)
    one = (T.DumpParams * 1)(T.dump_params("a", "b"))
    gi = T.GlobalInfo(0, 1, 0, 1, 0, 1, 1, 1, 1, 0)
    with pytest.raises(_lib.VpicError):
        _lib.call("vpic_hip_global_header", str(tmp_path / "x").encode(),
                  ctypes.byref(g), ctypes.byref(gi), one, 1)
