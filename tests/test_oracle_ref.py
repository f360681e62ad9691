"""Pins the oracle (oracle/oracle.c) to the reference's own scalar kernels
compiled from the lanl/vpic sources (oracle/_ref), bit for bit.

What each reference function is pinned by:
  advance_p_pipeline_scalar + move_p    -> test_advance_p_* (bit-exact,
                                           particles, accumulators, movers)
  move_p (single mover)                 -> test_move_p_single
  load_interpolator_pipeline_scalar     -> test_load_interpolator
  unload_accumulator_pipeline_scalar    -> test_unload_accumulator
  advance_b_pipeline_scalar (bulk)      -> test_advance_b (+ translation
                                           invariance for the face leftovers)
  vacuum_advance_e_pipeline_scalar      -> test_vacuum_advance_e (same)
  subsort_pipeline_scalar               -> test_sort_p
  energy_p_pipeline_scalar              -> test_energy_p
  vacuum_energy_f_pipeline_scalar       -> test_energy_f
"""
import numpy as np
import pytest

from oracle.oracle import host_grid
from vpic_amd import types as T

import helpers as H

pytestmark = pytest.mark.ref


def _box(nx, ny, nz, cell=1.0, frac=0.95):
    L = (nx * cell, ny * cell, nz * cell)
    dt = frac * T.courant_length(L[0], L[1], L[2], nx, ny, nz)
    return T.grid_periodic(0, 0, 0, L[0], L[1], L[2], nx, ny, nz, dt)


def _push_inputs(g, ppc, uth, seed, e_amp=0.05, b_amp=0.05):
    from oracle.oracle import Oracle
    f = H.random_fields(g, seed, e_amp, b_amp)
    fi = H.interp_array(g)
    Oracle.get().load_interpolator(fi, f, g)
    p = H.thermal_particles(g, ppc, uth, seed + 1)
    return p, fi


@pytest.mark.parametrize("q,m,uth,ppc", [(-1.0, 1.0, 0.1, 8),
                                          (1.0, 25.0, 0.02, 8),
                                          (-1.0, 1.0, 0.5, 13)])
def test_advance_p_bitexact(orc, ref, q, m, uth, ppc):
    g = _box(8, 8, 8)
    p, fi = _push_inputs(g, ppc, uth, 11)
    hg = host_grid(g)
    pa, pb = p.copy(), p.copy()
    aa, ab = H.acc_array(g), H.acc_array(g)
    nma, iga, pma = orc.advance_p(pa, q, m, aa, fi, hg)
    nmb, igb, pmb = ref.advance_p(pb, q, m, ab, fi, hg)
    assert H.same_bits(pa, pb)
    assert H.same_bits(aa, ab)
    assert (nma, iga) == (nmb, igb) == (0, 0)
    # crossers were exercised
    assert np.any(pa["i"] != p["i"])


def test_advance_p_walls_and_movers(orc, ref):
    """outbndj-style 4x5x6 box, fast particles, reflecting x walls and
    absorbing z walls: exercises reflection, returned movers, the mover
    segment split and its overflow rollback (advance_p_pipeline.cc:221-237)."""
    g = _box(4, 5, 6)
    g.pbc[0] = g.pbc[3] = T.BC_REFLECT_PARTICLES
    g.pbc[2] = g.pbc[5] = T.BC_ABSORB_PARTICLES
    p, fi = _push_inputs(g, 7, 0.9, 5, e_amp=0.2, b_amp=0.2)
    hg = host_grid(g)
    for max_nm in (10000, 40, 3):
        pa, pb = p.copy(), p.copy()
        aa, ab = H.acc_array(g), H.acc_array(g)
        ra = orc.advance_p(pa, -1.0, 1.0, aa, fi, hg, max_nm=max_nm)
        rb = ref.advance_p(pb, -1.0, 1.0, ab, fi, hg, max_nm=max_nm)
        assert H.same_bits(pa, pb)
        assert H.same_bits(aa, ab)
        assert ra[0] == rb[0] and ra[1] == rb[1]
        assert H.same_bits(ra[2], rb[2])
        if max_nm == 10000:
            assert ra[0] > 0 and ra[1] == 0
        if max_nm == 3:
            assert ra[1] > 0


def test_advance_p_threaded_particles_match(orc, ref):
    """Particle results do not depend on the pipeline count; only the
    deposit summation order does."""
    g = _box(8, 8, 8)
    p, fi = _push_inputs(g, 9, 0.1, 3)
    hg = host_grid(g)
    pa, pb = p.copy(), p.copy()
    aa, ab = H.acc_array(g), H.acc_array(g)
    orc.advance_p(pa, -1.0, 1.0, aa, fi, hg)
    ref.advance_p(pb, -1.0, 1.0, ab, fi, hg, n_pipeline=4)
    assert H.same_bits(pa, pb)
    assert H.acc_rel_err(aa, ab) < 1e-6


def test_move_p_single(orc, ref):
    g = _box(4, 5, 6)
    g.pbc[1] = g.pbc[4] = T.BC_REFLECT_PARTICLES
    hg = host_grid(g)
    rng = np.random.default_rng(7)
    p = H.thermal_particles(g, 2, 0.1, 8)
    for trial in range(200):
        k = int(rng.integers(len(p)))
        mv = np.zeros(1, dtype=T.MOVER)
        mv["dispx"], mv["dispy"], mv["dispz"] = rng.uniform(-3, 3, 3)
        mv["i"] = k
        pa, pb = p.copy(), p.copy()
        ma, mb = mv.copy(), mv.copy()
        aa, ab = H.acc_array(g), H.acc_array(g)
        ra = orc.move_p(pa, ma, aa, hg, -1.0)
        rb = ref.move_p(pb, mb, ab, hg, -1.0)
        assert ra == rb
        assert H.same_bits(pa, pb) and H.same_bits(ma, mb)
        assert H.same_bits(aa, ab)


def test_load_interpolator(orc, ref):
    g = _box(7, 5, 6)
    f = H.random_full_fields(g, 2)
    a, b = H.interp_array(g), H.interp_array(g)
    orc.load_interpolator(a, f, g)
    ref.load_interpolator(b, f, g)
    assert H.same_bits(a, b)


def test_unload_accumulator(orc, ref):
    g = _box(7, 5, 6)
    rng = np.random.default_rng(4)
    acc = rng.standard_normal((g.nv, 12)).astype(np.float32)
    # the reference assumes zero ghost accumulators
    x, y, z = np.meshgrid(np.arange(g.nx + 2), np.arange(g.ny + 2),
                          np.arange(g.nz + 2), indexing="ij")
    ghost = ((x == 0) | (x == g.nx + 1) | (y == 0) | (y == g.ny + 1) |
             (z == 0) | (z == g.nz + 1))
    acc[g.voxel(x[ghost], y[ghost], z[ghost])] = 0
    fa = H.random_full_fields(g, 3)
    fb = fa.copy()
    orc.unload_accumulator(fa, acc, g)
    ref.unload_accumulator(fb, acc, g)
    assert H.same_bits(fa, fb)


def _interior_mask(g, lo):
    v = np.arange(g.nv)
    x = v % (g.nx + 2)
    y = (v // (g.nx + 2)) % (g.ny + 2)
    z = v // ((g.nx + 2) * (g.ny + 2))
    return ((x >= lo) & (x <= g.nx) & (y >= lo) & (y <= g.ny) & (z >= lo) &
            (z <= g.nz))


def test_advance_b(orc, ref):
    g = _box(7, 5, 6)
    f = H.random_full_fields(g, 5)
    fa, fb = f.copy(), f.copy()
    orc.advance_b(fa, g, 0.5)
    ref.advance_b_bulk(fb, g, 0.5)
    m = _interior_mask(g, 1)
    for c in ("cbx", "cby", "cbz"):
        assert H.same_bits(fa[c][m], fb[c][m])


def _roll_field(g, f, sx, sy, sz):
    """Periodic translation of the interior+face data by (sx,sy,sz) cells:
    interior voxel (x,y,z) <- (x-sx,y-sy,z-sz) wrapped on 1..n."""
    a = f.reshape(g.nz + 2, g.ny + 2, g.nx + 2)
    out = f.copy().reshape(a.shape)
    inner = a[1:g.nz + 1, 1:g.ny + 1, 1:g.nx + 1]
    out[1:g.nz + 1, 1:g.ny + 1, 1:g.nx + 1] = np.roll(
        inner, shift=(sz, sy, sx), axis=(0, 1, 2))
    return out.reshape(-1)


def _periodize(g, f):
    """Make the +faces (x=nx+1 etc.) copies of the x=1 faces, as a
    periodic field is after a step."""
    a = f.reshape(g.nz + 2, g.ny + 2, g.nx + 2)
    a[:, :, g.nx + 1] = a[:, :, 1]
    a[:, g.ny + 1, :] = a[:, 1, :]
    a[g.nz + 1, :, :] = a[1, :, :]
    a[:, :, 0] = a[:, :, g.nx]
    a[:, 0, :] = a[:, g.ny, :]
    a[0, :, :] = a[g.nz, :, :]
    return f


def test_fdtd_translation_invariance(orc):
    """On a periodic box the face/leftover updates must equal the bulk
    update of the translated field; with the bulk pinned to the reference
    this pins every voxel."""
    g = _box(6, 5, 4)
    mp = T.vacuum_params()
    f = _periodize(g, H.random_full_fields(g, 6))
    for op in ("b", "e"):
        fa = f.copy()
        fr = _periodize(g, _roll_field(g, f, 2, 3, 1))
        if op == "b":
            orc.advance_b(fa, g, 0.5)
            orc.advance_b(fr, g, 0.5)
            comps = ("cbx", "cby", "cbz")
        else:
            orc.vacuum_advance_e(fa, g, mp)
            orc.vacuum_advance_e(fr, g, mp)
            comps = ("ex", "ey", "ez", "tcax", "tcay", "tcaz")
        ra = _roll_field(g, fa, 2, 3, 1)
        m = _interior_mask(g, 1)
        for c in comps:
            assert H.same_bits(ra[c][m], fr[c][m]), (op, c)
        # the components living on the x = nx+1 face stay copies of x = 1
        a = fa.reshape(g.nz + 2, g.ny + 2, g.nx + 2)
        face = ("cbx",) if op == "b" else ("ey", "ez", "tcay", "tcaz")
        for c in face:
            assert H.same_bits(a[c][1:g.nz + 1, 1:g.ny + 1, g.nx + 1],
                               a[c][1:g.nz + 1, 1:g.ny + 1, 1])


def test_vacuum_advance_e(orc, ref):
    g = _box(7, 5, 6)
    mp = T.vacuum_params(damp=0.01)
    f = H.random_full_fields(g, 8)
    fa, fb = f.copy(), f.copy()
    orc.vacuum_advance_e(fa, g, mp)
    ref.vacuum_advance_e_bulk(fb, g, mp)
    m = _interior_mask(g, 2)
    for c in ("ex", "ey", "ez", "tcax", "tcay", "tcaz"):
        assert H.same_bits(fa[c][m], fb[c][m]), c


def test_sort_p(orc, ref):
    g = _box(6, 5, 4)
    p = H.thermal_particles(g, 5, 0.1, 9, shuffle=True)
    pa, pb = p.copy(), p.copy()
    qa = orc.sort_p(pa, g)
    qb = ref.sort_p(pb, g)
    assert H.same_bits(pa, pb)
    assert H.same_bits(qa[:g.nv], qb[:g.nv])
    # stable: equals numpy's stable sort by voxel
    order = np.argsort(p["i"], kind="stable")
    assert H.same_bits(pa, p[order])


def test_energy_p(orc, ref):
    g = _box(6, 5, 4)
    p, fi = _push_inputs(g, 7, 0.1, 12)
    ea = orc.energy_p(p, -1.0, 1.0, fi, g)
    eb = ref.energy_p(p, -1.0, 1.0, fi, g)
    assert ea == eb


@pytest.mark.parametrize("uncenter", [False, True], ids=["center", "uncenter"])
def test_center_p_bitexact(orc, ref, uncenter):
    g = _box(6, 5, 4)
    p, fi = _push_inputs(g, 7, 0.3, 14, e_amp=0.3, b_amp=0.4)
    pa, pb = p.copy(), p.copy()
    orc.center_p(pa, -1.0, 1.0, fi, g, uncenter)
    ref.center_p(pb, -1.0, 1.0, fi, g, uncenter)
    assert H.same_bits(pa, pb)
    assert not H.same_bits(pa, p)


def test_center_p_closed_forms(orc):
    """Zero fields: both are the identity.  Uniform E, no B: center_p adds
    exactly qdt_2mc*E and uncenter_p subtracts it.  uncenter_p undoes
    center_p to rounding."""
    g = _box(4, 4, 4)
    p = H.thermal_particles(g, 5, 0.2, 3)
    fi = H.interp_array(g)
    for unc in (False, True):
        q = p.copy()
        orc.center_p(q, -1.0, 1.0, fi, g, unc)
        assert H.same_bits(q, p)
    fi["ex"], fi["ey"], fi["ez"] = 0.5, -0.25, 0.125
    qdt_2mc = T.push_coefficients(-1.0, 1.0, g)[0]
    q = p.copy()
    orc.center_p(q, -1.0, 1.0, fi, g, False)
    for c, e in (("ux", 0.5), ("uy", -0.25), ("uz", 0.125)):
        assert np.array_equal(q[c], p[c] + np.float32(qdt_2mc * np.float32(e)))
    fi2 = H.interp_array(g)
    f = H.random_fields(g, 5, 0.2, 0.3)
    orc.load_interpolator(fi2, f, g)
    q = p.copy()
    orc.center_p(q, -1.0, 1.0, fi2, g, False)
    orc.center_p(q, -1.0, 1.0, fi2, g, True)
    for c in ("ux", "uy", "uz"):
        assert np.allclose(q[c], p[c], rtol=0, atol=4e-7 * (1 + np.abs(p[c])))


def test_energy_f(orc, ref):
    g = _box(6, 5, 4)
    f = H.random_full_fields(g, 13)
    mp = T.vacuum_params()
    assert np.array_equal(orc.energy_f(f, g, mp), ref.energy_f(f, g, mp))


# -- divergence cleaning / initialization kernels of the vacuum table -------

def _mat(damp=0.0):
    mp = T.vacuum_params(damp)
    return mp


def test_div_e_err_and_rhob_bulk(orc, ref):
    """Every interior node (2..n) against the reference's own pipelines."""
    g = _box(7, 5, 6)
    mp = _mat()
    f = H.random_full_fields(g, 21)
    m = _interior_mask(g, 2)
    fa, fb = f.copy(), f.copy()
    orc.compute_div_e_err(fa, g, mp)
    ref.div_e_err_bulk(fb, g, mp)
    assert H.same_bits(fa["div_e_err"][m], fb["div_e_err"][m])
    fa, fb = f.copy(), f.copy()
    orc.compute_rhob(fa, g, mp)
    ref.rhob_bulk(fb, g, mp)
    assert H.same_bits(fa["rhob"][m], fb["rhob"][m])


def test_clean_div_e_curl_b_bulk(orc, ref):
    g = _box(6, 7, 5)
    mp = _mat()
    f = H.random_full_fields(g, 22)
    fa, fb = f.copy(), f.copy()
    orc.clean_div_e(fa, g, mp)
    ref.clean_div_e_bulk(fb, g, mp)
    m = _interior_mask(g, 1)
    for c in ("ex", "ey", "ez"):
        assert H.same_bits(fa[c][m], fb[c][m]), c
    fa, fb = f.copy(), f.copy()
    orc.compute_curl_b(fa, g, mp)
    ref.curl_b_bulk(fb, g, mp)
    m = _interior_mask(g, 2)
    for c in ("tcax", "tcay", "tcaz"):
        assert H.same_bits(fa[c][m], fb[c][m]), c


def test_div_b_err_and_clean_div_b(orc, ref):
    g = _box(5, 6, 7)
    f = H.random_full_fields(g, 23)
    fa, fb = f.copy(), f.copy()
    orc.compute_div_b_err(fa, g)
    ref.div_b_err(fb, g)
    m = _interior_mask(g, 1)
    assert H.same_bits(fa["div_b_err"][m], fb["div_b_err"][m])
    fa, fb = f.copy(), f.copy()
    orc.clean_div_b(fa, g)
    ref.clean_div_b_bulk(fb, g)
    m = _interior_mask(g, 2)
    for c in ("cbx", "cby", "cbz"):
        assert H.same_bits(fa[c][m], fb[c][m]), c


def test_rms_div_errors(orc, ref):
    g = _box(6, 5, 7)
    f = H.random_full_fields(g, 24)
    ro, rr = orc.compute_rms_div_b_err(f, g), ref.rms_div_b_err(f, g)
    assert abs(ro - rr) <= 1e-12 * abs(rr)
    # with div_e_err zero on the boundary nodes only the pipelined interior
    # sum remains
    a = f.reshape(g.nz + 2, g.ny + 2, g.nx + 2)
    for sl in ((slice(None), slice(None), 1), (slice(None), slice(None), g.nx + 1),
               (slice(None), 1, slice(None)), (slice(None), g.ny + 1, slice(None)),
               (1, slice(None), slice(None)), (g.nz + 1, slice(None), slice(None))):
        a["div_e_err"][sl] = 0
    bulk = ref.rms_div_e_err_bulk(f, g)
    n = g.nx * g.ny * g.nz
    want = float(g.eps0) * np.sqrt(bulk * float(g.dV) /
                                   float(np.float32(n) * g.dV))
    assert abs(orc.compute_rms_div_e_err(f, g) - want) <= 1e-12 * want


def test_rms_closed_forms(orc):
    """A uniform error c gives eps0*|c|: the node weights (1, 1/2, 1/4, 1/8)
    sum to nx*ny*nz, as the cell count does for div_b_err."""
    g = _box(5, 4, 6)
    g.eps0 = np.float32(2.0)
    f = np.zeros(g.nv, dtype=T.FIELD)
    f["div_e_err"] = np.float32(0.75)
    f["div_b_err"] = np.float32(-1.25)
    assert abs(orc.compute_rms_div_e_err(f, g) - 1.5) <= 1e-12
    assert abs(orc.compute_rms_div_b_err(f, g) - 2.5) <= 1e-12


def test_cleaning_translation_invariance(orc):
    """On a periodic box the face/leftover parts of each kernel equal the
    bulk part on the translated field (the bulk is pinned to the reference
    above), so every node, edge and face is pinned."""
    g = _box(6, 5, 4)
    mp = _mat()
    f = _periodize(g, H.random_full_fields(g, 25))
    ops = {
        "div_e_err": (lambda x: orc.compute_div_e_err(x, g, mp), ("div_e_err",)),
        "rhob": (lambda x: orc.compute_rhob(x, g, mp), ("rhob",)),
        "clean_div_e": (lambda x: orc.clean_div_e(x, g, mp), ("ex", "ey", "ez")),
        "curl_b": (lambda x: orc.compute_curl_b(x, g, mp),
                   ("tcax", "tcay", "tcaz")),
        "div_b_err": (lambda x: orc.compute_div_b_err(x, g), ("div_b_err",)),
        "clean_div_b": (lambda x: orc.clean_div_b(x, g), ("cbx", "cby", "cbz")),
    }
    m = _interior_mask(g, 1)
    for name, (op, comps) in ops.items():
        fa = f.copy()
        fr = _periodize(g, _roll_field(g, f, 2, 3, 1))
        op(fa)
        op(fr)
        ra = _roll_field(g, fa, 2, 3, 1)
        for c in comps:
            assert H.same_bits(ra[c][m], fr[c][m]), (name, c)


def test_discrete_divergence(orc):
    """div_e_err is the backward difference of E minus rho/eps0 on the
    periodic node lattice (numpy roll as the independent statement)."""
    g = _box(6, 5, 4)
    g.eps0 = np.float32(1.0)
    mp = _mat()
    f = _periodize(g, H.random_full_fields(g, 26))
    fa = f.copy()
    orc.compute_div_e_err(fa, g, mp)
    a = f.reshape(g.nz + 2, g.ny + 2, g.nx + 2)
    inn = (slice(1, g.nz + 1), slice(1, g.ny + 1), slice(1, g.nx + 1))
    ex, ey, ez = (a[c][inn].astype(np.float64) for c in ("ex", "ey", "ez"))
    rho = (a["rhof"][inn].astype(np.float64) + a["rhob"][inn])
    d = ((ex - np.roll(ex, 1, axis=2)) * g.rdx + (ey - np.roll(ey, 1, axis=1)) *
         g.rdy + (ez - np.roll(ez, 1, axis=0)) * g.rdz - rho)
    got = fa.reshape(a.shape)["div_e_err"][inn]
    assert np.max(np.abs(got - d)) <= 1e-5 * np.max(np.abs(d))


def test_synchronize_tang_e_norm_b_closed_forms(orc):
    g = _box(5, 4, 3)
    f = _periodize(g, H.random_full_fields(g, 27))
    fa = f.copy()
    assert orc.synchronize_tang_e_norm_b(fa, g) == 0.0
    assert H.same_bits(fa, f)
    # one face value off by d: err = 2 d^2 (seen from both faces), both
    # faces take the FP64 average
    a = fa.reshape(g.nz + 2, g.ny + 2, g.nx + 2)
    d = np.float32(0.5)
    lo = a["cbx"][2, 2, 1]
    a["cbx"][2, 2, g.nx + 1] = lo + d
    err = orc.synchronize_tang_e_norm_b(fa, g)
    assert abs(err - 2 * float(d) ** 2) <= 1e-12
    want = np.float32(0.5 * (float(lo) + float(lo + d)))
    assert a["cbx"][2, 2, 1] == want and a["cbx"][2, 2, g.nx + 1] == want


# -- hydro moments -----------------------------------------------------------

@pytest.mark.parametrize("q,m,uth", [(-1.0, 1.0, 0.2), (2.0, 25.0, 1.5)])
def test_accumulate_hydro_p_bitexact(orc, ref, q, m, uth):
    g = _box(6, 5, 4)
    p, fi = _push_inputs(g, 9, uth, 31, e_amp=0.2, b_amp=0.3)
    h0 = np.zeros(g.nv, dtype=T.HYDRO)
    h0[:] = np.random.default_rng(2).standard_normal((g.nv, 16)).astype(
        np.float32).view(T.HYDRO).reshape(-1)
    ha, hb = h0.copy(), h0.copy()
    orc.accumulate_hydro_p(ha, p, q, m, fi, g)
    ref.accumulate_hydro_p(hb, p, q, m, fi, g)
    for c in T.HYDRO_FIELDS:
        assert H.same_bits(ha[c], hb[c]), c


def test_hydro_closed_forms(orc):
    """A particle at rest at a cell centre with zero fields puts w r8V q
    (rho) and w r8V m c (only through t; px = 0) on each of its 8 nodes;
    sums over the nodes give the particle's totals."""
    g = _box(4, 4, 4)
    fi = np.zeros(g.nv, dtype=T.INTERPOLATOR)
    p = np.zeros(1, dtype=T.PARTICLE)
    p["i"] = g.voxel(2, 2, 2)
    p["w"] = 3.0
    p["ux"] = 0.5
    h = np.zeros(g.nv, dtype=T.HYDRO)
    orc.accumulate_hydro_p(h, p, -2.0, 4.0, fi, g)
    r8V = float(g.r8V)
    gam = np.sqrt(1 + 0.25)
    nodes = [g.voxel(2 + a, 2 + b, 2 + c) for c in (0, 1) for b in (0, 1)
             for a in (0, 1)]
    assert np.allclose(h["rho"][nodes], -2.0 * 3.0 * r8V, rtol=1e-6)
    assert np.allclose(h["px"][nodes], 4.0 * 3.0 * r8V * 0.5, rtol=1e-6)
    assert np.allclose(h["jx"][nodes], -2.0 * 3.0 * r8V * 0.5 / gam, rtol=1e-6)
    assert np.allclose(h["ke"][nodes], 4.0 * 3.0 * r8V * (gam - 1), rtol=1e-5)
    assert np.allclose(h["txx"][nodes], 4.0 * 3.0 * r8V * 0.25 / gam, rtol=1e-6)
    other = np.ones(g.nv, bool)
    other[nodes] = False
    assert not np.any(h["rho"][other])


def test_synchronize_hydro_closed_forms(orc):
    """Periodic faces: both node planes take the sum; local faces double."""
    g = _box(4, 3, 5)
    h = np.zeros(g.nv, dtype=T.HYDRO)
    rng = np.random.default_rng(4)
    for c in T.HYDRO_FIELDS:
        h[c] = rng.integers(-8, 8, g.nv).astype(np.float32)
    a = h.reshape(g.nz + 2, g.ny + 2, g.nx + 2)
    want = a.copy()
    s = want["rho"][2, 2, 1] + want["rho"][2, 2, g.nx + 1]
    orc.synchronize_hydro(h, g)
    assert a["rho"][2, 2, 1] == s and a["rho"][2, 2, g.nx + 1] == s
    g2 = g.copy()
    g2.fbc[0] = g2.fbc[3] = T.BC_PEC_FIELDS
    h2 = want.copy().reshape(-1)
    orc.synchronize_hydro(h2, g2)
    b = h2.reshape(a.shape)
    assert b["jx"][2, 2, 1] == 2 * want["jx"][2, 2, 1]
    assert b["jx"][2, 2, g.nx + 1] == 2 * want["jx"][2, 2, g.nx + 1]
