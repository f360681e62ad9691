"""CPU checks of the C-ABI library: it loads, exports every entry point the
headers declare, and rejects bad arguments before touching a device (the
reference's "Bad args" paths, e.g. advance_p_pipeline.cc:263-270)."""
import ctypes
import os
import re
import subprocess

import pytest

from vpic_amd import _lib
from vpic_amd import types as T

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header(fn):
    src = open(os.path.join(ROOT, "include", fn)).read()
    return re.sub(r"/\*.*?\*/", "", src, flags=re.S)


def _declared(fn="vpic_hip.h"):
    return sorted(set(re.findall(r"\b(vpic_hip_\w+)\s*\(", _header(fn))))


# The reference-named functions of vpic_dropin.h (function declarations:
# a return type, the name, an argument list ending in ");").
DROPIN_NAMES = sorted(set(re.findall(
    r"^(?:void|int|double)\s+(\w+)\s*\([^;]*\);", _header("vpic_dropin.h"),
    flags=re.M)))


def test_library_loads():
    lib = _lib.load()
    assert b"gfx950" in lib.vpic_hip_version()


@pytest.mark.parametrize("name", _declared() + DROPIN_NAMES)
def test_symbol_exported(name):
    lib = _lib.load()
    assert hasattr(lib, name), name


def test_dropin_declares_reference_kernels():
    for name in ("advance_p", "sort_p", "move_p", "energy_p",
                 "accumulate_rho_p", "center_p", "uncenter_p",
                 "load_interpolator_array", "clear_accumulator_array",
                 "reduce_accumulator_array", "unload_accumulator_array",
                 "vpic_hip_install_field_kernels"):
        assert name in DROPIN_NAMES, name


def test_binding_covers_header():
    assert set(_declared()) <= set(_lib.EXPORTS) | {"vpic_hip_version"}


def test_dropin_header_in_c(tmp_path):
    """vpic_dropin.h compiles as C99, and the library's layout table agrees
    with the header's own types (no device needed)."""
    src = tmp_path / "t.c"
    src.write_text(
        '#include "vpic_dropin.h"\n#include "vpic_dropin_layout.h"\n'
        "#include <stdio.h>\n"
        "int main(void) {\n"
        "  const int64_t mine[] = {VPIC_DROPIN_LAYOUT(VPIC_DROPIN_SIZE,"
        " VPIC_DROPIN_OFFSET)};\n"
        "  int64_t lib[512];\n"
        "  int n = (int)(sizeof mine / sizeof mine[0]);\n"
        "  int c = vpic_hip_dropin_layout(lib, 512);\n"
        "  if (c != n) return 2;\n"
        "  for (int k = 0; k < n; k++) if (lib[k] != mine[k]) return 3;\n"
        "  printf(\"%d\\n\", n);\n  return 0;\n}\n")
    exe = tmp_path / "t"
    libdir = os.path.dirname(_lib.LIB_PATH)
    subprocess.run(["gcc", "-std=c99", "-Wall", "-Werror", "-I",
                    os.path.join(ROOT, "include"), str(src), "-o", str(exe),
                    "-L" + libdir, "-l:libvpichip.so",
                    "-Wl,-rpath," + libdir, "-Wl,-rpath,/opt/rocm/lib"],
                   check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True)
    assert out.returncode == 0, out
    assert int(out.stdout) > 150


REF_CHECK = os.path.join(ROOT, "oracle", "_ref", "dropin_check")


@pytest.mark.ref
@pytest.mark.skipif(not os.path.exists(REF_CHECK),
                    reason="oracle/_ref (reference build) not present")
def test_dropin_layout_matches_reference_headers():
    """A program compiled against lanl/vpic's own headers sees the same
    sizes and offsets as the library."""
    out = subprocess.run([REF_CHECK, "--layout"], capture_output=True,
                         text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert "ok layout" in out.stdout


def test_grid_helper_matches_reference_math():
    g1 = T.Grid()
    _lib.call("vpic_hip_grid_periodic", ctypes.byref(g1), 0.0, 0.0, 0.0,
              32.0, 16.0, 8.0, 32, 16, 8, 0.5, 1.0, 1.0)
    g2 = T.grid_periodic(0, 0, 0, 32, 16, 8, 32, 16, 8, 0.5)
    assert bytes(g1) == bytes(g2)


def test_bad_args_rejected():
    lib = _lib.load()
    g = T.grid_periodic(0, 0, 0, 4, 4, 4, 4, 4, 4, 0.5)
    nm = ctypes.c_int32()
    # NULL accumulator / interpolator
    rc = lib.vpic_hip_advance_p(None, 0, -1.0, 1.0, None, 0,
                                ctypes.byref(nm), None, None, None,
                                ctypes.byref(g), None, None)
    assert rc < 0 and b"NULL" in lib.vpic_hip_last_error()
    # inconsistent strides
    bad = g.copy()
    bad.sy += 1
    rc = lib.vpic_hip_advance_b(ctypes.c_void_p(16), ctypes.byref(bad), 0.5,
                                None)
    assert rc < 0 and b"strides" in lib.vpic_hip_last_error()
    # advance_e frac != 1 (vacuum_advance_e_pipeline.cc:58-61)
    mp = T.vacuum_params()
    rc = lib.vpic_hip_vacuum_advance_e(ctypes.c_void_p(16), ctypes.byref(g),
                                       ctypes.byref(mp), 0.5, None)
    assert rc < 0 and b"frac" in lib.vpic_hip_last_error()
    # absorbing field faces are not supported
    ab = g.copy()
    ab.fbc[0] = ab.fbc[3] = T.BC_ABSORB_FIELDS
    rc = lib.vpic_hip_advance_b(ctypes.c_void_p(16), ctypes.byref(ab), 0.5,
                                None)
    assert rc < 0 and b"absorbing" in lib.vpic_hip_last_error()
    # one-sided periodic axis
    one = g.copy()
    one.pbc[0] = T.BC_REFLECT_PARTICLES
    rc = lib.vpic_hip_sort_p(ctypes.c_void_p(16), ctypes.c_void_p(32), 1,
                             ctypes.c_void_p(64), ctypes.byref(one), 1, None)
    assert rc < 0 and b"partner" in lib.vpic_hip_last_error()


def test_species_push_args_rejected():
    """vpic_hip_advance_p_species checks its species before any device
    work: a count outside 1..2, and two species sharing a per-species
    output (dev_nm counters, movers, keys)."""
    lib = _lib.load()
    g = T.grid_periodic(0, 0, 0, 4, 4, 4, 4, 4, 4, 0.5)
    v = ctypes.c_void_p
    arr = (T.PushSpecies * 2)(
        T.PushSpecies(16, 8, -1.0, 1.0, 4096, 8, 8192, None, None),
        T.PushSpecies(32, 8, 1.0, 25.0, 4096 * 3, 8, 8192, None, None))
    rc = lib.vpic_hip_advance_p_species(3, arr, v(64), v(128),
                                        ctypes.byref(g), None)
    assert rc < 0 and b"species in one advance_p" in lib.vpic_hip_last_error()
    rc = lib.vpic_hip_advance_p_species(2, arr, v(64), v(128),
                                        ctypes.byref(g), None)
    assert rc < 0 and b"dev_nm" in lib.vpic_hip_last_error()
    arr[1].dev_nm = 8192 + 64
    arr[1].pm = 4096
    rc = lib.vpic_hip_advance_p_species(2, arr, v(64), v(128),
                                        ctypes.byref(g), None)
    assert rc < 0 and b"mover array" in lib.vpic_hip_last_error()
    arr[1].pm = 4096 * 3
    arr[0].keys = arr[1].keys = 1 << 20
    rc = lib.vpic_hip_advance_p_species(2, arr, v(64), v(128),
                                        ctypes.byref(g), None)
    assert rc < 0 and b"key array" in lib.vpic_hip_last_error()


def test_neighbor_table_matches_oracle(orc):
    import numpy as np
    from oracle.oracle import neighbor_table
    g = T.grid_periodic(0, 0, 0, 4, 5, 6, 4, 5, 6, 0.5)
    g.pbc[0] = g.pbc[3] = T.BC_REFLECT_PARTICLES
    g.pbc[2] = g.pbc[5] = T.BC_ABSORB_PARTICLES
    nb = np.zeros(6 * g.nv, dtype=np.int64)
    _lib.call("vpic_hip_grid_neighbor_table", ctypes.byref(g),
              nb.ctypes.data_as(ctypes.c_void_p))
    assert np.array_equal(nb, neighbor_table(g))
