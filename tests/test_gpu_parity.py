"""GPU parity of the HIP path (through the C-ABI) against the oracle.

Bars (from the path's specification):
  * cell indices, cell-crosser counts, movers: bit-exact
  * FP32 positions and momenta: <= 2 ulp per step (the kernels reproduce the
    reference's operation order, so the observed difference is 0 ulp and
    the tests assert 0)
  * accumulated currents: <= 1e-5 relative (max-norm, floor = max |ref|):
    only the summation order of the deposits differs
  * field stencils, interpolator load, jf unload/fold, sort permutation and
    partition: bit-exact
  * FP64 energies: <= 1e-12 relative (summation order)
"""
import ctypes
import sys

import numpy as np
import pytest

from oracle.oracle import host_grid
from vpic_amd import _lib
from vpic_amd import types as T

import helpers as H

pytestmark = pytest.mark.gpu

ACC_TOL = 1e-5


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    _lib.call("vpic_hip_init", 0)
    from vpic_amd import device
    return device


def _box(nx, ny, nz, cell=1.0, frac=0.95):
    L = (nx * cell, ny * cell, nz * cell)
    dt = frac * T.courant_length(L[0], L[1], L[2], nx, ny, nz)
    return T.grid_periodic(0, 0, 0, L[0], L[1], L[2], nx, ny, nz, dt)


def _inputs(orc, g, ppc, uth, seed, e_amp=0.05, b_amp=0.05, shuffle=False):
    f = H.random_fields(g, seed, e_amp, b_amp)
    fi = H.interp_array(g)
    orc.load_interpolator(fi, f, g)
    p = H.thermal_particles(g, ppc, uth, seed + 1, shuffle=shuffle)
    return p, fi


def _gpu_advance_p(dev, p, q, m, fi, g, max_nm=None, acc=None,
                   partition=None):
    if max_nm is None:
        max_nm = max(len(p), 16)
    dp = dev.DeviceArray.from_host(p)
    dfi = dev.DeviceArray.from_host(fi)
    da = dev.DeviceArray.from_host(acc if acc is not None else H.acc_array(g))
    dpm = dev.movers(max_nm)
    dpart = (dev.DeviceArray.from_host(partition.astype(np.int32))
             if partition is not None else None)
    nm = ctypes.c_int32()
    ign = ctypes.c_int32()
    _lib.call("vpic_hip_advance_p", dp.ptr, len(p), q, m, dpm.ptr, max_nm,
              ctypes.byref(nm), ctypes.byref(ign), da.ptr, dfi.ptr,
              ctypes.byref(g), dpart.ptr if dpart is not None else None, None)
    pm = dpm.download()[:nm.value].copy()
    return dp.download(), da.download(), nm.value, ign.value, pm


CASES = [
    ("electrons uth=0.1", -1.0, 1.0, 0.1, 8, (8, 8, 8)),
    ("ions mi/me=25 uth=0.02", 1.0, 25.0, 0.02, 8, (8, 8, 8)),
    ("hot uth=0.5, heavy crossing", -1.0, 1.0, 0.5, 13, (8, 8, 8)),
    ("ragged np, 1-cell-thick box", -1.0, 1.0, 0.3, 7, (5, 1, 3)),
]


@pytest.mark.parametrize("name,q,m,uth,ppc,n", CASES, ids=[c[0] for c in CASES])
def test_advance_p_parity(dev, orc, name, q, m, uth, ppc, n):
    g = _box(*n)
    p, fi = _inputs(orc, g, ppc, uth, 21)
    hg = host_grid(g)
    po, ao = p.copy(), H.acc_array(g)
    nmo, igo, pmo = orc.advance_p(po, q, m, ao, fi, hg)
    say("advance_p case:", name, "np =", len(p))
    pg, ag, nmg, igg, pmg = _gpu_advance_p(dev, p, q, m, fi, g)
    assert np.array_equal(pg["i"], po["i"])
    assert H.particle_ulp_diff(pg, po) == 0
    assert (nmg, igg) == (nmo, igo) == (0, 0)
    assert H.acc_rel_err(ag, ao) <= ACC_TOL
    assert np.any(po["i"] != p["i"])  # crossers exercised


def test_advance_p_unsorted_and_np_edges(dev, orc):
    g = _box(6, 5, 4)
    p, fi = _inputs(orc, g, 9, 0.2, 5, shuffle=True)
    hg = host_grid(g)
    for npart in (0, 1, 63, 64, 65, 257, len(p)):
        sub = p[:npart].copy()
        po, ao = sub.copy(), H.acc_array(g)
        orc.advance_p(po, -1.0, 1.0, ao, fi, hg)
        say("advance_p np =", npart)
        pg, ag, nmg, _, _ = _gpu_advance_p(dev, sub, -1.0, 1.0, fi, g)
        assert H.same_bits(pg, po)
        assert nmg == 0
        assert H.acc_rel_err(ag, ao) <= ACC_TOL or npart == 0


def test_advance_p_walls_movers(dev, orc):
    """Reflecting x walls, absorbing z walls (movers returned in particle
    order), and mover-capacity overflow with the p->i >>= 3 rollback."""
    g = _box(4, 5, 6)
    g.pbc[0] = g.pbc[3] = T.BC_REFLECT_PARTICLES
    g.pbc[2] = g.pbc[5] = T.BC_ABSORB_PARTICLES
    p, fi = _inputs(orc, g, 7, 0.9, 5, e_amp=0.2, b_amp=0.2)
    hg = host_grid(g)
    po, ao = p.copy(), H.acc_array(g)
    nmo, igo, pmo = orc.advance_p(po, -1.0, 1.0, ao, fi, hg, max_nm=len(p))
    assert nmo > 0
    say("advance_p walls/movers, nm =", nmo)
    pg, ag, nmg, igg, pmg = _gpu_advance_p(dev, p, -1.0, 1.0, fi, g,
                                           max_nm=len(p))
    assert (nmg, igg) == (nmo, 0)
    assert H.same_bits(pmg, pmo)
    assert H.same_bits(pg, po)
    assert H.acc_rel_err(ag, ao) <= ACC_TOL
    # overflow: keep the first max_nm movers in particle order, roll back
    # the voxel index of the rest
    cap = nmo // 3
    say("advance_p mover overflow, max_nm =", cap)
    pg, ag, nmg, igg, pmg = _gpu_advance_p(dev, p, -1.0, 1.0, fi, g,
                                           max_nm=cap)
    assert (nmg, igg) == (cap, nmo - cap)
    assert H.same_bits(pmg, pmo[:cap])
    expect = po.copy()
    lost = pmo["i"][cap:]
    expect["i"][lost] = expect["i"][lost] >> 3
    assert H.same_bits(pg, expect)


def test_advance_p_bad_voxel_reported(dev, orc):
    g = _box(4, 4, 4)
    p, fi = _inputs(orc, g, 4, 0.1, 8)
    p["i"][10] = g.nv + 5
    p["i"][11] = -3
    say("advance_p with out-of-range voxel indices")
    with pytest.raises(_lib.VpicError, match="outside"):
        _gpu_advance_p(dev, p, -1.0, 1.0, fi, g)


def test_move_p_api(dev, orc):
    g = _box(4, 5, 6)
    g.pbc[1] = g.pbc[4] = T.BC_REFLECT_PARTICLES
    hg = host_grid(g)
    p = H.thermal_particles(g, 2, 0.1, 8)
    rng = np.random.default_rng(3)
    mv = np.zeros(50, dtype=T.MOVER)
    mv["i"] = rng.choice(len(p), 50, replace=False)
    for c in ("dispx", "dispy", "dispz"):
        mv[c] = rng.uniform(-3, 3, 50).astype(np.float32)
    po, ao, mo = p.copy(), H.acc_array(g), mv.copy()
    ret_o = []
    for k in range(len(mo)):
        one = mo[k:k + 1].copy()
        ret_o.append(orc.move_p(po, one, ao, hg, -1.0))
        mo[k] = one[0]
    dp = dev.DeviceArray.from_host(p)
    dm = dev.DeviceArray.from_host(mv)
    da = dev.DeviceArray.from_host(H.acc_array(g))
    dr = dev.DeviceArray((len(mv),), np.int32)
    say("move_p API on", len(mv), "movers")
    _lib.call("vpic_hip_move_p", dp.ptr, dm.ptr, len(mv), da.ptr,
              ctypes.byref(g), -1.0, dr.ptr, None)
    assert np.array_equal(dr.download(), np.array(ret_o, dtype=np.int32))
    assert H.same_bits(dp.download(), po)
    assert H.same_bits(dm.download(), mo)
    assert H.acc_rel_err(da.download(), ao) <= ACC_TOL


def test_load_interpolator_bitexact(dev, orc):
    g = _box(7, 5, 6)
    f = H.random_full_fields(g, 2)
    fo = H.interp_array(g)
    orc.load_interpolator(fo, f, g)
    df = dev.DeviceArray.from_host(f)
    di = dev.DeviceArray.from_host(H.interp_array(g))
    say("load_interpolator")
    _lib.call("vpic_hip_load_interpolator_array", di.ptr, df.ptr,
              ctypes.byref(g), None)
    got = di.download()
    for c in T.INTERP_FIELDS[:18]:
        assert H.same_bits(got[c], fo[c]), c


def _field_bc_grids():
    g = _box(7, 5, 6)
    yield "periodic", g
    g2 = g.copy()
    g2.fbc[0] = g2.fbc[3] = T.BC_PEC_FIELDS
    g2.fbc[2] = g2.fbc[5] = T.BC_SYMMETRIC_FIELDS
    yield "pec-x sym-z", g2
    g3 = g.copy()
    g3.fbc[1] = g3.fbc[4] = T.BC_PMC_FIELDS
    yield "pmc-y", g3


def test_field_kernels_bitexact(dev, orc):
    mp = T.vacuum_params(damp=0.0)
    mp2 = T.vacuum_params(damp=0.02)
    for name, g in _field_bc_grids():
        f = H.random_full_fields(g, 3)
        acc = np.random.default_rng(1).standard_normal((g.nv, 12)).astype(
            np.float32)
        # the reference assumes zero ghost accumulators
        v = np.arange(g.nv)
        x, y, z = v % (g.nx + 2), (v // (g.nx + 2)) % (g.ny + 2), v // (
            (g.nx + 2) * (g.ny + 2))
        ghost = ((x == 0) | (x == g.nx + 1) | (y == 0) | (y == g.ny + 1) |
                 (z == 0) | (z == g.nz + 1))
        acc[ghost] = 0
        fo = f.copy()
        orc.clear_jf(fo, g)
        orc.unload_accumulator(fo, acc, g)
        orc.synchronize_jf(fo, g)
        orc.advance_b(fo, g, 0.5)
        orc.vacuum_advance_e(fo, g, mp)
        orc.advance_b(fo, g, 0.5)
        orc.vacuum_advance_e(fo, g, mp2)
        df = dev.DeviceArray.from_host(f)
        da = dev.DeviceArray.from_host(acc)
        say("field kernels,", name)
        G = ctypes.byref(g)
        _lib.call("vpic_hip_clear_jf", df.ptr, G, None)
        _lib.call("vpic_hip_unload_accumulator_array", df.ptr, da.ptr, G, None)
        _lib.call("vpic_hip_synchronize_jf", df.ptr, G, None)
        _lib.call("vpic_hip_advance_b", df.ptr, G, 0.5, None)
        _lib.call("vpic_hip_vacuum_advance_e", df.ptr, G, ctypes.byref(mp),
                  1.0, None)
        _lib.call("vpic_hip_advance_b", df.ptr, G, 0.5, None)
        _lib.call("vpic_hip_vacuum_advance_e", df.ptr, G, ctypes.byref(mp2),
                  1.0, None)
        got = df.download()
        for c in T.FIELD_FLOATS:
            assert H.same_bits(got[c], fo[c]), (name, c)
        # energy_f
        en = (ctypes.c_double * 6)()
        _lib.call("vpic_hip_energy_f", en, df.ptr, G, ctypes.byref(mp), None)
        eo = orc.energy_f(fo, g, mp)
        assert np.allclose(np.array(en[:]), eo, rtol=1e-12, atol=0)


def test_clear_jf_unload_fused(dev):
    """vpic_hip_clear_jf_unload_accumulator_array is the reference's
    clear_jf + unload_accumulator_array pair (advance.cc:107-109) in one
    pass: every field component bit-identical to the two calls', with
    nonzero ghost accumulators, rhof and the other fields left as they
    were."""
    for name, g in _field_bc_grids():
        f = H.random_full_fields(g, 5)
        acc = np.random.default_rng(2).standard_normal((g.nv, 12)).astype(
            np.float32)
        G = ctypes.byref(g)
        d2 = dev.DeviceArray.from_host(f)
        d1 = dev.DeviceArray.from_host(f)
        da = dev.DeviceArray.from_host(acc)
        say("clear_jf + unload fused,", name)
        _lib.call("vpic_hip_clear_jf", d2.ptr, G, None)
        _lib.call("vpic_hip_unload_accumulator_array", d2.ptr, da.ptr, G, None)
        _lib.call("vpic_hip_clear_jf_unload_accumulator_array", d1.ptr,
                  da.ptr, G, None)
        want, got = d2.download(), d1.download()
        for c in T.FIELD_FLOATS:
            assert H.same_bits(got[c], want[c]), (name, c)
        assert H.same_bits(got["rhof"], f["rhof"]), name
        assert not H.same_bits(got["jfx"], f["jfx"]), name


CLEANING = ["compute_div_e_err", "compute_rhob", "clean_div_e",
            "compute_curl_b", "compute_div_b_err", "clean_div_b",
            "synchronize_tang_e_norm_b", "rms"]


@pytest.mark.parametrize("op", CLEANING)
def test_divergence_cleaning_bitexact(dev, orc, op):
    """The divergence-cleaning / initialization kernels of the vacuum field
    table against the oracle on every boundary mix: fields bit-exact, the
    FP64 sums (rms, err) to 1e-12 relative."""
    mp = T.vacuum_params(damp=0.0)
    for name, g in _field_bc_grids():
        f = H.random_full_fields(g, 11)
        fo = f.copy()
        df = dev.DeviceArray.from_host(f)
        G = ctypes.byref(g)
        say(op + ",", name)
        if op == "rms":
            for kind in ("e", "b"):
                r = ctypes.c_double()
                _lib.call("vpic_hip_compute_rms_div_%s_err" % kind,
                          ctypes.byref(r), df.ptr, G, None)
                want = getattr(orc, "compute_rms_div_%s_err" % kind)(fo, g)
                assert abs(r.value - want) <= 1e-12 * abs(want), (name, kind)
            continue
        if op == "synchronize_tang_e_norm_b":
            err = ctypes.c_double()
            _lib.call("vpic_hip_" + op, ctypes.byref(err), df.ptr, G, None)
            want = orc.synchronize_tang_e_norm_b(fo, g)
            assert abs(err.value - want) <= 1e-12 * max(abs(want), 1e-300)
        elif op in ("compute_div_b_err", "clean_div_b"):
            _lib.call("vpic_hip_" + op, df.ptr, G, None)
            getattr(orc, op)(fo, g)
        else:
            _lib.call("vpic_hip_" + op, df.ptr, G, ctypes.byref(mp), None)
            getattr(orc, op)(fo, g, mp)
        got = df.download()
        for c in T.FIELD_FLOATS:
            assert H.same_bits(got[c], fo[c]), (name, op, c)


def _hydro_err(got, want):
    worst = 0.0
    for c in T.HYDRO_FIELDS:
        den = float(np.max(np.abs(want[c])))
        if den:
            worst = max(worst, float(np.max(np.abs(got[c] - want[c]))) / den)
        else:
            assert not np.any(got[c]), c
    return worst


@pytest.mark.parametrize("sorted_", [False, True], ids=["flat", "tiled"])
def test_accumulate_hydro_p(dev, orc, sorted_):
    """accumulate_hydro_p + synchronize_hydro_array against the oracle (node
    sums to 1e-5 of each moment's largest value), flat and tiled, with
    particles appended after the sort."""
    for name, g in _field_bc_grids():
        p, fi = _inputs(orc, g, 9, 0.6, 81, e_amp=0.2, b_amp=0.3,
                        shuffle=True)
        part = None
        if sorted_:
            part = orc.sort_p(p, g)
            p = np.concatenate([p, H.thermal_particles(g, 1, 0.6, 82,
                                                       shuffle=True)[:53]])
        ho = np.zeros(g.nv, dtype=T.HYDRO)
        orc.accumulate_hydro_p(ho, p, -1.0, 1.0, fi, g)
        orc.synchronize_hydro(ho, g)
        dh = dev.DeviceArray((g.nv,), T.HYDRO)
        dp = dev.DeviceArray.from_host(p)
        dfi = dev.DeviceArray.from_host(fi)
        dpart = dev.DeviceArray.from_host(part) if part is not None else None
        G = ctypes.byref(g)
        say("accumulate_hydro_p,", name, "tiled" if sorted_ else "flat")
        _lib.call("vpic_hip_clear_hydro_array", dh.ptr, G, None)
        _lib.call("vpic_hip_accumulate_hydro_p", dh.ptr, dp.ptr, len(p), -1.0,
                  1.0, dfi.ptr, G, dpart.ptr if dpart else None, None)
        _lib.call("vpic_hip_synchronize_hydro_array", dh.ptr, G, None)
        assert _hydro_err(dh.download(), ho) <= 1e-5, name


def test_synchronize_rho_bitexact(dev, orc):
    for name, g in _field_bc_grids():
        f = H.random_full_fields(g, 5)
        fo = f.copy()
        orc.synchronize_rho(fo, g)
        df = dev.DeviceArray.from_host(f)
        say("synchronize_rho,", name)
        _lib.call("vpic_hip_synchronize_rho", df.ptr, ctypes.byref(g), None)
        got = df.download()
        for c in T.FIELD_FLOATS:
            assert H.same_bits(got[c], fo[c]), (name, c)


@pytest.mark.parametrize("mode", ["shuffled", "drifted", "one-hot-voxel"])
def test_sort_p_bitexact(dev, orc, mode):
    g = _box(6, 5, 4)
    if mode == "shuffled":
        p = H.thermal_particles(g, 5, 0.1, 9, shuffle=True)
    elif mode == "drifted":
        # sorted, then advanced a few steps so cells drift
        p, fi = _inputs(orc, g, 40, 0.3, 4)
        hg = host_grid(g)
        for _ in range(3):
            orc.advance_p(p, -1.0, 1.0, H.acc_array(g), fi, hg)
    else:
        # one voxel with more particles than one LDS segment holds
        p = H.thermal_particles(g, 3, 0.1, 9, shuffle=True)
        extra = H.thermal_particles(g, 1, 0.1, 10)[:1].repeat(3000)
        extra["dx"] = np.linspace(-1, 1, 3000, dtype=np.float32)
        p = np.concatenate([p, extra])
        p = p[np.random.default_rng(2).permutation(len(p))]
    po = p.copy()
    part_o = orc.sort_p(po, g)
    dp = dev.DeviceArray.from_host(p)
    dx = dev.particles(len(p))
    dpart = dev.DeviceArray((g.nv + 1,), np.int32)
    say("sort_p", mode, "np =", len(p))
    _lib.call("vpic_hip_sort_p", dp.ptr, dx.ptr, len(p), dpart.ptr,
              ctypes.byref(g), 1, None)
    assert H.same_bits(dp.download(), po)
    assert H.same_bits(dpart.download(), part_o)


def _stale_count():
    c = ctypes.c_int64()
    _lib.call("vpic_hip_sort_keys_stale", ctypes.byref(c))
    return c.value


@pytest.mark.parametrize("flat", [False, True], ids=["tiled", "flat"])
def test_push_keys_then_sort_p_keys(dev, orc, flat):
    """The keys the push leaves (vpic_hip_advance_p_keys, tiled and flat,
    with crossers) are the particles' voxel indices, and the sort from them
    is vpic_hip_sort_p's, bit for bit; keys that are not the particles' own
    (one changed, one outside the grid) give the same sort through the
    fallback, which is counted."""
    g = _box(7, 6, 5)
    p, fi = _inputs(orc, g, 9, 0.3, 44, shuffle=True)
    hg = host_grid(g)
    part = orc.sort_p(p, g)
    for _ in range(2):  # drift off the sort-time cells
        orc.advance_p(p, -1.0, 1.0, H.acc_array(g), fi, hg)
    po, ao = p.copy(), H.acc_array(g)
    orc.advance_p(po, -1.0, 1.0, ao, fi, hg)
    n = len(p)
    dp = dev.DeviceArray.from_host(p)
    dfi = dev.DeviceArray.from_host(fi)
    da = dev.DeviceArray.from_host(H.acc_array(g))
    dpm = dev.movers(n)
    dnm = dev.DeviceArray((4,), np.int32)
    dpart = dev.DeviceArray.from_host(part.astype(np.int32))
    dkeys = dev.DeviceArray((n,), np.int32)
    G = ctypes.byref(g)
    say("advance_p_keys", "flat" if flat else "tiled", "np =", n)
    _lib.call("vpic_hip_advance_p_keys", dp.ptr, n, -1.0, 1.0, dpm.ptr, n,
              dnm.ptr, da.ptr, dfi.ptr, G, None if flat else dpart.ptr,
              dkeys.ptr, None)
    pg = dp.download()
    assert H.same_bits(pg, po)
    assert H.acc_rel_err(da.download(), ao) <= ACC_TOL
    assert np.array_equal(dkeys.download(), po["i"])
    assert np.any(po["i"] != p["i"])  # crossers exercised
    ps = po.copy()
    part_o = orc.sort_p(ps, g)
    aux = dev.particles(n)
    for case in ("keys", "one changed", "out of range"):
        k = po["i"].copy()
        if case == "one changed":
            k[n // 2] = (k[n // 2] + 1) % g.nv
        elif case == "out of range":
            k[7] = g.nv + 5
        dk = dev.DeviceArray.from_host(k.astype(np.int32))
        dq = dev.DeviceArray.from_host(po)
        dpt = dev.DeviceArray((g.nv + 1,), np.int32)
        before = _stale_count()
        say("sort_p_keys", case)
        _lib.call("vpic_hip_sort_p_keys", dq.ptr, aux.ptr, n, dpt.ptr,
                  dk.ptr, G, 1, None)
        assert H.same_bits(dq.download(), ps), case
        assert H.same_bits(dpt.download(), part_o), case
        assert _stale_count() - before == (0 if case == "keys" else 1), case


@pytest.mark.parametrize("case", ["both sorted", "one unsorted"])
def test_advance_p_species_pair(dev, orc, case, tile_reset):
    """Electrons and ions in one vpic_hip_advance_p_species call (one tiled
    launch through the same windows when both have partitions, one call
    each otherwise): each species bit-exact, the shared accumulator
    against the oracle's two advance_p calls, the keys the particles'."""
    _lib.call("vpic_hip_push_tile", 4, 3, 5)
    g = _box(9, 7, 6)
    f = H.random_full_fields(g, 81, amp=0.3)
    fi = H.interp_array(g)
    orc.load_interpolator(fi, f, g)
    hg = host_grid(g)
    specs = [(-1.0, 1.0, 0.3, 11, 91), (1.0, 25.0, 0.05, 7, 92)]
    ps, parts = [], []
    for q, m, uth, ppc, seed in specs:
        p = H.thermal_particles(g, ppc, uth, seed, shuffle=True)
        parts.append(orc.sort_p(p, g))
        for _ in range(3):  # drift off the sort-time cells
            orc.advance_p(p, q, m, H.acc_array(g), fi, hg)
        ps.append(p)
    ao = H.acc_array(g)
    pos = [p.copy() for p in ps]
    for (q, m, _, _, _), po in zip(specs, pos):
        orc.advance_p(po, q, m, ao, fi, hg)
    dfi = dev.DeviceArray.from_host(fi)
    da = dev.DeviceArray.from_host(H.acc_array(g))
    keep, arr = [], (T.PushSpecies * 2)()
    for k, ((q, m, _, _, _), p, part) in enumerate(zip(specs, ps, parts)):
        dp = dev.DeviceArray.from_host(p)
        dpm = dev.movers(len(p))
        dnm = dev.DeviceArray((4,), np.int32)
        dkeys = dev.DeviceArray((len(p),), np.int32)
        dpart = (dev.DeviceArray.from_host(part.astype(np.int32))
                 if case == "both sorted" or k == 0 else None)
        keep.append((dp, dpm, dnm, dkeys, dpart))
        arr[k] = T.PushSpecies(dp.ptr, len(p), q, m, dpm.ptr, len(p),
                               dnm.ptr, dpart.ptr if dpart else None,
                               dkeys.ptr)
    say("advance_p_species", case)
    _lib.call("vpic_hip_advance_p_species", 2, arr, da.ptr, dfi.ptr,
              ctypes.byref(g), None)
    for k, (dp, dpm, dnm, dkeys, dpart) in enumerate(keep):
        assert H.same_bits(dp.download(), pos[k]), k
        assert np.array_equal(dkeys.download(), pos[k]["i"]), k
        assert tuple(dnm.download()[:3]) == (0, 0, 0), k
    assert H.acc_rel_err(da.download(), ao) <= ACC_TOL


def test_energy_p_and_rho(dev, orc):
    g = _box(6, 5, 4)
    p, fi = _inputs(orc, g, 7, 0.1, 12)
    eo = orc.energy_p(p, -1.0, 1.0, fi, g)
    dp = dev.DeviceArray.from_host(p)
    dfi = dev.DeviceArray.from_host(fi)
    en = ctypes.c_double()
    say("energy_p")
    _lib.call("vpic_hip_energy_p", dp.ptr, len(p), -1.0, 1.0, dfi.ptr,
              ctypes.byref(g), ctypes.byref(en), None)
    assert abs(en.value - eo) <= 1e-12 * abs(eo)
    f = np.zeros(g.nv, dtype=T.FIELD)
    fo = f.copy()
    orc.accumulate_rho_p(fo, p, -1.0, g)
    df = dev.DeviceArray.from_host(f)
    say("accumulate_rho_p")
    _lib.call("vpic_hip_accumulate_rho_p", df.ptr, dp.ptr, len(p), -1.0,
              ctypes.byref(g), None)
    got = df.download()["rhof"]
    assert H.acc_rel_err(got, fo["rhof"]) <= 1e-6


@pytest.mark.parametrize("uncenter", [0, 1], ids=["center", "uncenter"])
def test_center_p_bitexact(dev, orc, uncenter):
    g = _box(6, 5, 4)
    p, fi = _inputs(orc, g, 9, 0.3, 21)
    po = p.copy()
    orc.center_p(po, 1.0, 25.0, fi, g, bool(uncenter))
    dp = dev.DeviceArray.from_host(p)
    dfi = dev.DeviceArray.from_host(fi)
    say("center_p uncenter =", uncenter)
    _lib.call("vpic_hip_center_p", dp.ptr, len(p), 1.0, 25.0, dfi.ptr,
              ctypes.byref(g), uncenter, None)
    assert H.same_bits(dp.download(), po)


def test_loader_statistics(dev):
    g = _box(8, 8, 8)
    ppc = 64
    dp = dev.particles(ppc * 512)
    say("load_uniform_plasma")
    _lib.call("vpic_hip_load_uniform_plasma", dp.ptr, ctypes.byref(g), ppc,
              0.1, 0.1, 0.1, 0.0, 0.0, 0.05, 2.0, 1234, None)
    p = dp.download()
    assert np.all(np.abs(p["dx"]) <= 1) and np.all(np.abs(p["dz"]) <= 1)
    cells = np.bincount(p["i"], minlength=g.nv)
    x, y, z = H.voxel_coords(g, 512)
    vox = x + (g.nx + 2) * (y + (g.ny + 2) * z)
    assert np.all(cells[vox] == ppc)
    assert abs(np.std(p["ux"]) - 0.1) < 0.003
    assert abs(np.mean(p["uz"]) - 0.05) < 0.003
    assert np.all(p["w"] == 2.0)
    # reproducible
    dq = dev.particles(ppc * 512)
    _lib.call("vpic_hip_load_uniform_plasma", dq.ptr, ctypes.byref(g), ppc,
              0.1, 0.1, 0.1, 0.0, 0.0, 0.05, 2.0, 1234, None)
    assert H.same_bits(dq.download(), p)


# ---------------------------------------------------------------------------
# The partition-driven (tiled, LDS-accumulating) push
# ---------------------------------------------------------------------------
TILES = [(7, 8, 7), (8, 8, 8), (8, 8, 4), (1, 1, 1), (16, 4, 4), (3, 2, 5)]


@pytest.fixture
def tile_reset():
    was = (ctypes.c_int32 * 4)()
    _lib.call("vpic_hip_push_get_tile", was)
    yield
    _lib.call("vpic_hip_push_tile_window", *was)
    _lib.call("vpic_hip_push_group", 16)


GROUPS = [1, 2, 4, 8, 16, 32, 64]


@pytest.mark.parametrize("gs", GROUPS, ids=["gs%d" % gs for gs in GROUPS])
def test_tiled_push_groups(dev, orc, gs, tile_reset):
    """Every lane-group size of the tiled push, on drifted particles and on
    fast ones whose crossers fill the wavefront queues many times over."""
    _lib.call("vpic_hip_push_group", gs)
    g = _box(9, 8, 5)
    for (q, m, uth, ppc, steps) in ((-1.0, 1.0, 0.3, 13, 3),
                                    (-1.0, 1.0, 1.5, 40, 0)):
        p, fi = _inputs(orc, g, ppc, uth, 37, shuffle=True)
        hg = host_grid(g)
        part = orc.sort_p(p, g)
        for _ in range(steps):
            orc.advance_p(p, q, m, H.acc_array(g), fi, hg)
        po, ao = p.copy(), H.acc_array(g)
        orc.advance_p(po, q, m, ao, fi, hg)
        say("tiled advance_p group", gs, "uth", uth, "ppc", ppc)
        pg, ag, nmg, igg, _ = _gpu_advance_p(dev, p, q, m, fi, g,
                                             partition=part)
        assert H.same_bits(pg, po)
        assert (nmg, igg) == (0, 0)
        assert H.acc_rel_err(ag, ao) <= ACC_TOL


@pytest.mark.parametrize("tile", TILES, ids=["x".join(map(str, t)) for t in TILES])
def test_tiled_push_parity(dev, orc, tile, tile_reset):
    _lib.call("vpic_hip_push_tile", *tile)
    g = _box(7, 5, 6)
    for (q, m, uth, steps) in ((-1.0, 1.0, 0.1, 0), (-1.0, 1.0, 0.3, 4),
                               (1.0, 25.0, 0.02, 2)):
        p, fi = _inputs(orc, g, 11, uth, 31, shuffle=True)
        hg = host_grid(g)
        part = orc.sort_p(p, g)
        # let the particles drift off their sort-time cells (stale partition)
        for _ in range(steps):
            orc.advance_p(p, q, m, H.acc_array(g), fi, hg)
        po, ao = p.copy(), H.acc_array(g)
        orc.advance_p(po, q, m, ao, fi, hg)
        say("tiled advance_p tile", tile, "uth", uth, "drift steps", steps)
        pg, ag, nmg, igg, _ = _gpu_advance_p(dev, p, q, m, fi, g,
                                             partition=part)
        assert H.same_bits(pg, po)
        assert (nmg, igg) == (0, 0)
        assert H.acc_rel_err(ag, ao) <= ACC_TOL


# tiles with a two-cell window halo (tx, ty, tz, halo): tiles narrower than
# the halo put a voxel in the windows of up to five tiles per axis
WINDOWS = [(8, 8, 4, 2), (6, 6, 6, 2), (1, 1, 1, 2), (3, 2, 5, 2),
           (2, 3, 2, 2)]


@pytest.mark.parametrize("win", WINDOWS,
                         ids=["x".join(map(str, w[:3])) + "h%d" % w[3]
                              for w in WINDOWS])
def test_tiled_push_halo2_parity(dev, orc, win, tile_reset):
    """The two-cell halo: particles drifted up to several cells since the
    sort (their deposits in the halo of their tile's window, in the window
    of a neighbouring tile's halo, or outside every window)."""
    _lib.call("vpic_hip_push_tile_window", *win)
    g = _box(9, 7, 6)
    for (q, m, uth, steps) in ((-1.0, 1.0, 0.1, 0), (-1.0, 1.0, 0.3, 6),
                               (-1.0, 1.0, 0.6, 5), (1.0, 25.0, 0.02, 2)):
        p, fi = _inputs(orc, g, 11, uth, 33, shuffle=True)
        hg = host_grid(g)
        part = orc.sort_p(p, g)
        for _ in range(steps):
            orc.advance_p(p, q, m, H.acc_array(g), fi, hg)
        po, ao = p.copy(), H.acc_array(g)
        orc.advance_p(po, q, m, ao, fi, hg)
        say("tiled advance_p window", win, "uth", uth, "drift steps", steps)
        pg, ag, nmg, igg, _ = _gpu_advance_p(dev, p, q, m, fi, g,
                                             partition=part)
        assert H.same_bits(pg, po)
        assert (nmg, igg) == (0, 0)
        assert H.acc_rel_err(ag, ao) <= ACC_TOL


# grids one cell thick along an axis (2D runs): the tile is clamped to the
# grid and its window has no halo along that axis
THIN = [((9, 7, 1), (8, 8, 8, 1)), ((9, 1, 6), (4, 8, 4, 2)),
        ((1, 6, 5), (8, 3, 2, 1)), ((10, 9, 1), (32, 16, 1, 1))]


@pytest.mark.parametrize("n,win", THIN,
                         ids=["%dx%dx%d" % t[0] for t in THIN])
def test_tiled_push_thin_grid(dev, orc, n, win, tile_reset):
    """A 2D grid (one cell along x, y or z): every deposit stays in the
    plane, with particles drifted off their sort-time cells."""
    _lib.call("vpic_hip_push_tile_window", *win)
    g = _box(*n)
    for (q, m, uth, steps) in ((-1.0, 1.0, 0.3, 3), (-1.0, 1.0, 0.8, 4)):
        p, fi = _inputs(orc, g, 13, uth, 35, shuffle=True)
        hg = host_grid(g)
        part = orc.sort_p(p, g)
        for _ in range(steps):
            orc.advance_p(p, q, m, H.acc_array(g), fi, hg)
        po, ao = p.copy(), H.acc_array(g)
        orc.advance_p(po, q, m, ao, fi, hg)
        say("tiled advance_p thin grid", n, "window", win, "uth", uth)
        pg, ag, nmg, igg, _ = _gpu_advance_p(dev, p, q, m, fi, g,
                                             partition=part)
        assert H.same_bits(pg, po)
        assert (nmg, igg) == (0, 0)
        assert H.acc_rel_err(ag, ao) <= ACC_TOL


def test_tiled_push_multistep_device_sort(dev, orc):
    """sort_p and several tiled pushes entirely on the device (the partition
    goes stale as the particles drift), each step against the oracle."""
    g = _box(8, 8, 4)
    f = H.random_full_fields(g, 51, amp=0.3)
    fi = H.interp_array(g)
    orc.load_interpolator(fi, f, g)
    p = H.thermal_particles(g, 6, 0.3, 61, shuffle=True)
    hg = host_grid(g)
    po = p.copy()
    orc.sort_p(po, g)
    dp = dev.DeviceArray.from_host(p)
    aux = dev.particles(len(p))
    dpart = dev.DeviceArray((g.nv + 1,), np.int32)
    dfi = dev.DeviceArray.from_host(fi)
    da = dev.DeviceArray.from_host(H.acc_array(g))
    dpm = dev.movers(len(p))
    G = ctypes.byref(g)
    say("device sort + 4 tiled pushes")
    _lib.call("vpic_hip_sort_p", dp.ptr, aux.ptr, len(p), dpart.ptr, G, 1,
              None)
    assert H.same_bits(dp.download(), po)
    nm = ctypes.c_int32()
    for step in range(4):
        ao = H.acc_array(g)
        orc.advance_p(po, -1.0, 1.0, ao, fi, hg)
        _lib.call("vpic_hip_clear_accumulator_array", da.ptr, G, None)
        _lib.call("vpic_hip_advance_p", dp.ptr, len(p), -1.0, 1.0, dpm.ptr,
                  len(p), ctypes.byref(nm), None, da.ptr, dfi.ptr, G,
                  dpart.ptr, None)
        say("  step", step)
        assert H.same_bits(dp.download(), po), step
        assert H.acc_rel_err(da.download(), ao) <= ACC_TOL, step


BIN_MODES = [0, 1, 2]


@pytest.mark.parametrize("mode", BIN_MODES, ids=["bins%d" % m for m in BIN_MODES])
def test_tiled_push_tile_lists(dev, orc, mode, tile_reset):
    """Fast particles on small tiles drift across several tiles between
    sorts: every push (partition rows or tile lists, whichever the mode
    picks) against the oracle, with a re-sort in the middle, a flat push, and
    a change of the particle count."""
    _lib.call("vpic_hip_push_bin_mode", mode)
    _lib.call("vpic_hip_push_tile", 2, 3, 2)
    try:
        g = _box(9, 8, 6)
        f = H.random_full_fields(g, 71, amp=0.3)
        fi = H.interp_array(g)
        orc.load_interpolator(fi, f, g)
        p = H.thermal_particles(g, 7, 0.5, 73, shuffle=True)
        hg = host_grid(g)
        n0 = len(p)
        po = p.copy()
        dp = dev.DeviceArray.from_host(p)
        aux = dev.particles(n0)
        dpart = dev.DeviceArray((g.nv + 1,), np.int32)
        dfi = dev.DeviceArray.from_host(fi)
        da = dev.DeviceArray.from_host(H.acc_array(g))
        dpm = dev.movers(n0)
        G = ctypes.byref(g)
        nm = ctypes.c_int32()
        np_cur = n0
        # step kinds: s = sort then tiled push, t = tiled push, f = flat push,
        # c = tiled push with the last 37 particles dropped
        plan = "sttttsttfttctt"
        for step, kind in enumerate(plan):
            if kind == "s":
                orc.sort_p(po[:np_cur], g)
                _lib.call("vpic_hip_sort_p", dp.ptr, aux.ptr, np_cur,
                          dpart.ptr, G, 1, None)
            if kind == "c":
                np_cur -= 37
            part = None if kind == "f" else dpart.ptr
            ao = H.acc_array(g)
            sub = po[:np_cur]
            orc.advance_p(sub, -1.0, 1.0, ao, fi, hg)
            po[:np_cur] = sub
            _lib.call("vpic_hip_clear_accumulator_array", da.ptr, G, None)
            _lib.call("vpic_hip_advance_p", dp.ptr, np_cur, -1.0, 1.0,
                      dpm.ptr, n0, ctypes.byref(nm), None, da.ptr, dfi.ptr, G,
                      part, None)
            say("  bin mode", mode, "step", step, kind, "np", np_cur)
            assert nm.value == 0
            assert H.same_bits(dp.download()[:np_cur], po[:np_cur]), step
            assert H.acc_rel_err(da.download(), ao) <= ACC_TOL, step
    finally:
        _lib.call("vpic_hip_push_bin_mode", 0)


def test_tiled_push_np_changed(dev, orc):
    """Particles appended after the sort (pushed by the flat tail) and
    particles removed (partition clamped to np)."""
    g = _box(6, 5, 4)
    p, fi = _inputs(orc, g, 9, 0.2, 41, shuffle=True)
    part = orc.sort_p(p, g)
    hg = host_grid(g)
    extra = H.thermal_particles(g, 1, 0.2, 42, shuffle=True)[:97]
    for sub in (np.concatenate([p, extra]), p[:len(p) - 123].copy()):
        po, ao = sub.copy(), H.acc_array(g)
        orc.advance_p(po, -1.0, 1.0, ao, fi, hg)
        say("tiled advance_p with np", len(sub), "partition total", part[g.nv])
        pg, ag, nmg, _, _ = _gpu_advance_p(dev, sub, -1.0, 1.0, fi, g,
                                           partition=part)
        assert H.same_bits(pg, po)
        assert nmg == 0
        assert H.acc_rel_err(ag, ao) <= ACC_TOL


def test_tiled_push_walls_movers(dev, orc):
    g = _box(4, 5, 6)
    g.pbc[0] = g.pbc[3] = T.BC_REFLECT_PARTICLES
    g.pbc[2] = g.pbc[5] = T.BC_ABSORB_PARTICLES
    p, fi = _inputs(orc, g, 7, 0.9, 5, e_amp=0.2, b_amp=0.2, shuffle=True)
    part = orc.sort_p(p, g)
    hg = host_grid(g)
    po, ao = p.copy(), H.acc_array(g)
    nmo, _, pmo = orc.advance_p(po, -1.0, 1.0, ao, fi, hg, max_nm=len(p))
    say("tiled advance_p walls/movers")
    pg, ag, nmg, igg, pmg = _gpu_advance_p(dev, p, -1.0, 1.0, fi, g,
                                           max_nm=len(p), partition=part)
    assert (nmg, igg) == (nmo, 0) and nmo > 0
    assert H.same_bits(pmg, pmo)
    assert H.same_bits(pg, po)
    assert H.acc_rel_err(ag, ao) <= ACC_TOL


def test_tile_shape_rejected():
    with pytest.raises(_lib.VpicError, match="rejected"):
        _lib.call("vpic_hip_push_tile", 64, 64, 64)
    # 8x8x8 with a two-cell halo: 12^3 window cells of 96 B exceed the LDS
    with pytest.raises(_lib.VpicError, match="rejected"):
        _lib.call("vpic_hip_push_tile_window", 8, 8, 8, 2)
    with pytest.raises(_lib.VpicError, match="not 1 or 2"):
        _lib.call("vpic_hip_push_tile_window", 4, 4, 4, 3)
