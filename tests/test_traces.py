"""Energy and charge traces over many steps, against the FP64 oracle.

The oracle's step functions are compiled twice from one source
(oracle/oracle_step.inc): over float they are the FP32 restatement, pinned
bit for bit against the reference's scalar kernels (tests/test_oracle_ref.py);
over double they are the FP64 oracle.  A run of each from the same initial
state (oracle/trace.py: OracleRun) differs only in its arithmetic.

What pins the FP64 build beyond sharing the pinned source:
  * one step from the same state agrees with the FP32 build to FP32
    rounding (positions, momenta, cell indices, currents);
  * charge is conserved to FP64 rounding: after every step the rms of
    div E - rho/eps0 (compute_rms_div_e_err, the reference's charge
    diagnostic, advance.cc:146-160) stays at 1e-15 of the charge scale,
    on a periodic box and on a box with PEC/reflecting walls -- a dropped,
    doubled or misplaced current term would give ~1e-3.

The traces compared step by step (dump_energies, dump.cc:38-77: the six
field energies and each species' kinetic energy, FP64 sums; and the rms
div E error) and their bounds, all measured first on the CPU (FP32 oracle
against FP64 oracle) and then stated here:
  * thermal plasma (the 32^3, 2 x 64 ppc, 100-step box and reduced boxes):
    electric and magnetic energy sums within 1e-3 relative of the FP64
    trace, kinetic energy within 2e-5, total energy within 2e-5, rms div E
    error below 2e-6 of the mean charge density (FP32 rounding of the node
    sums, flat over the run).  Measured, oracle FP32 against FP64: 1.6e-4,
    2.5e-4, 3.3e-6, 3.0e-6 and 6.1e-7 on the 32^3 box (1.5e-5, 2.7e-5,
    2.4e-6, 2.3e-6 and 5.9e-7 on the 8^3 box); the field energies are the
    plasma's noise, whose FP32 and FP64 realizations drift apart as the
    particle orbits do;
  * Harris sheet (walls, nonuniform density, heavy crossing): the same
    bounds (measured 2.0e-4, 1.6e-6, 1.8e-6, 5.1e-7, 2.1e-7 on a 64 x 32
    sheet over 200 steps).
The GPU traces (FP32 push, FP64 device reductions) must meet the same
bounds against the FP64 oracle, and their own rms div E error the same
ceiling.

The reference's own multi-step fixture, test/unit/energy_comparison/
energies_gold.3d_test (a 16^3 box, 2 x 16 ppc, mi/me 1836, 50 steps;
3d_test.cc), was made with the reference's SFMT stream seeded by rank,
which cannot be reproduced, so it is compared the way its own test does
(compare_energies.h: sums over the masks e = 0x0E, b = 0x70,
"particle" = 0xC0, relative error |A-B|/min(A,B)) where that is a property
of the physics rather than of one random realization: the electric-energy
sum within 30% and the 0xC0 sum (bz plus the first species' kinetic
energy, as the mask selects) within 1% at every step, as there; the
magnetic-energy sum, noise of one realization whose per-step 3% tolerance
holds only for the same random numbers, within 15% as a mean over the 50
steps (six of our realizations: +5% +-3.5%).  Lines where the reference
switches to its absolute branch (min(A,B) < 10 FLT_EPSILON, which demands
equality to DBL_EPSILON) are skipped: they too need the same random
numbers.
"""
import ctypes
import os
import sys

import numpy as np
import pytest

from oracle.trace import OracleRun
from vpic_amd import types as T

import backends as B
import helpers as H

GOLD_3D = os.path.join(os.path.dirname(__file__), "golden",
                       "energies_gold.3d_test")

# trace bounds (see the module docstring)
THERMAL = {"E": 1e-3, "B": 1e-3, "KE": 2e-5, "total": 2e-5, "div": 2e-6}
HARRIS = THERMAL


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


# --------------------------------------------------------------------------
# workloads
# --------------------------------------------------------------------------

def thermal_box(n, ppc, seed=1):
    """The plain-method plasma: n^3 periodic cells (dx = d_e = 1), electrons
    uth 0.1c and ions (mi/me = 25) uth 0.02c, ppc each, positions uniform
    in each cell, weights n0 dV / ppc, dt = 0.95 Courant, fields zero."""
    dt = 0.95 * T.courant_length(n, n, n, n, n, n)
    g = T.grid_periodic(0, 0, 0, n, n, n, n, n, n, dt)
    w = 1.0 / ppc
    pe = H.thermal_particles(g, ppc, 0.1, seed, w=w)
    pi = H.thermal_particles(g, ppc, 0.02, seed + 1, w=w)
    return g, [("electron", -1.0, 1.0, pe), ("ion", 1.0, 25.0, pi)], None


def box_3d_test(seed=1):
    """test/unit/energy_comparison/3d_test.cc:60-250: a periodic
    2.0944 x 1 x 1 box of 16^3 cells from (0, -0.5, -0.5), dt = 0.99
    Courant, 16 x 4096 electron-ion pairs loaded at common uniform
    positions, bi-Maxwellian vth = 0.25/sqrt(2) across and 0.05/sqrt(2)
    along x, mi/me = 1836, weights n0 V / Ne.  The species are listed ion
    first (define_species prepends to the species list), as the gold file's
    columns are."""
    lx, ly, lz, n, nppc = 2.09439510239320, 1.0, 1.0, 16, 16
    dt = 0.99 * T.courant_length(lx, ly, lz, n, n, n)
    g = T.grid_periodic(0, -0.5 * ly, -0.5 * lz, lx, 0.5 * ly, 0.5 * lz,
                        n, n, n, dt)
    ne = nppc * n * n * n
    we = lx * ly * lz / ne
    vth, vthx = 0.25 / np.sqrt(2.0), 0.05 / np.sqrt(2.0)
    rng = np.random.default_rng(seed)
    x = rng.uniform(0, lx, ne)
    y = rng.uniform(-0.5 * ly, 0.5 * ly, ne)
    z = rng.uniform(-0.5 * lz, 0.5 * lz, ne)
    lo, hi = (0.0, -0.5 * ly, -0.5 * lz), (lx, 0.5 * ly, 0.5 * lz)
    pe = B.inject(g, x, y, z, rng.normal(0, vthx, ne), rng.normal(0, vth, ne),
                  rng.normal(0, vth, ne), we, lo, hi)
    pi = B.inject(g, x, y, z, rng.normal(0, vthx, ne), rng.normal(0, vth, ne),
                  rng.normal(0, vth, ne), we, lo, hi)
    return g, [("ion", 1.0, 1836.0, pi), ("electron", -1.0, 1.0, pe)], None


def harris_sheet(nx, nz, lx, lz, ppc, seed=1):
    """A 2D Harris current sheet in the x-z plane (ny = 1), after
    sample/harris:84-164 with the sheet normal along x: PEC field walls and
    reflecting particle walls on the x faces, periodic z.  Natural units
    (d_e = c = m_e = |e| = eps0 = 1, peak density n0 = 1), mi/me = 25,
    half-thickness L = 0.5 d_i = 2.5 d_e, Ti/Te = 5, wpe/wce = 3 (so
    b0 = 1/3 and pressure balance gives kTe = 1/108), background density
    0.2 n0 with no drift, no guide field; B_z = b0 tanh(x/L) plus a 10%
    flux perturbation A_y = 0.1 b0 d_i cos(2 pi z/Lz) cos(pi x/Lx), B taken
    as the discrete curl of A_y on the Yee grid (div B = 0 to rounding).
    Particles (equal weights, ppc per species on average) are split
    between the sech^2 sheet (x = L atanh(u), drifting along y with the
    deck's vdre, vdri = -Ti/Te vdre) and the background in proportion to
    their physical numbers; electron and ion of a pair share a position."""
    mi, L, ti_te, wpe_wce, nb = 25.0, 2.5, 5.0, 3.0, 0.2
    b0 = 1.0 / wpe_wce
    kte = 1.0 / (2 * wpe_wce ** 2 * (1 + ti_te))
    kti = kte * ti_te
    uthe, uthi = np.sqrt(kte), np.sqrt(kti / mi)
    vdre = b0 / (L * (1 + ti_te))
    vdri = -ti_te * vdre
    dx, dz = lx / nx, lz / nz
    dt = 0.99 * T.courant_length(lx, dx, lz, nx, 1, nz)
    g = T.grid_periodic(-0.5 * lx, 0, 0, 0.5 * lx, dx, lz, nx, 1, nz, dt)
    for k in (0, 3):
        g.fbc[k] = T.BC_PEC_FIELDS
        g.pbc[k] = T.BC_REFLECT_PARTICLES
    # fields: A_y at the nodes, B = curl A (cbx on x-faces, cbz on z-faces)
    xn = -0.5 * lx + dx * np.arange(nx + 3) - dx  # node x of index 0..nx+2
    zn = dz * np.arange(nz + 3) - dz
    X, Z = np.meshgrid(xn, zn)  # [z, x]
    A = b0 * L * np.log(np.cosh(X / L)) + 0.1 * b0 * 5.0 * \
        np.cos(2 * np.pi * Z / lz) * np.cos(np.pi * X / lx)
    f = np.zeros(g.nv, dtype=T.FIELD)
    cbx = np.zeros((nz + 2, 3, nx + 2))
    cbz = np.zeros((nz + 2, 3, nx + 2))
    cbx[:, 1, :] = -(A[1:, :-1] - A[:-1, :-1]) / dz
    cbz[:, 1, :] = (A[:-1, 1:] - A[:-1, :-1]) / dx
    f["cbx"] = cbx.reshape(-1).astype(np.float32)
    f["cbz"] = cbz.reshape(-1).astype(np.float32)
    # particles
    rng = np.random.default_rng(seed)
    n_sheet_phys = 2 * L * np.tanh(0.5 * lx / L)
    n_bg_phys = nb * lx
    ntot = ppc * nx * nz
    ns = int(round(ntot * n_sheet_phys / (n_sheet_phys + n_bg_phys)))
    nbg = ntot - ns
    w = (n_sheet_phys + n_bg_phys) * dx * lz / ntot  # per unit y = dx
    xs = []
    while sum(len(v) for v in xs) < ns:
        v = L * np.arctanh(rng.uniform(-1, 1, 2 * ns))
        xs.append(v[(v > -0.5 * lx) & (v < 0.5 * lx)])
    xs = np.concatenate(xs)[:ns]
    x = np.concatenate([xs, rng.uniform(-0.5 * lx, 0.5 * lx, nbg)])
    y = rng.uniform(0, dx, ntot)
    z = rng.uniform(0, lz, ntot)
    drift = np.concatenate([np.ones(ns), np.zeros(nbg)])
    lo, hi = (-0.5 * lx, 0.0, 0.0), (0.5 * lx, dx, lz)
    out = []
    for name, q, m, uth, vd in (("ion", 1.0, mi, uthi, vdri),
                                ("electron", -1.0, 1.0, uthe, vdre)):
        ux, uy, uz = (rng.normal(0, uth, ntot) for _ in range(3))
        gd = 1 / np.sqrt(1 - vd * vd)
        ud = vd * gd
        # boost along y to the drift frame (sample/harris:263-266)
        uy = np.where(drift > 0, gd * uy +
                      np.sqrt(ux * ux + uy * uy + uz * uz + 1) * ud, uy)
        out.append((name, q, m, B.inject(g, x, y, z, ux, uy, uz, w, lo, hi)))
    return g, out, f


# --------------------------------------------------------------------------
# trace helpers
# --------------------------------------------------------------------------

def oracle_trace(g, species, fields, steps, fp64):
    r = OracleRun(g, species, fp64=fp64, fields=fields)
    r.initialize()
    rows = []
    for _ in range(steps):
        r.advance()
        ef, ep = r.energies()
        rows.append(np.r_[ef, ep, r.div_e_rms()])
    r.close()
    return np.array(rows)


def compare_traces(a, want, bound, nsp, label):
    """a, want: rows of [6 field energies, nsp kinetic energies, rms div E
    error]; want is the FP64 oracle's trace."""
    E, Ew = a[:, 0:3].sum(1), want[:, 0:3].sum(1)
    Bf, Bw = a[:, 3:6].sum(1), want[:, 3:6].sum(1)
    K, Kw = a[:, 6:6 + nsp].sum(1), want[:, 6:6 + nsp].sum(1)
    tot, totw = E + Bf + K, Ew + Bw + Kw
    errs = {"E": np.max(np.abs(E - Ew) / np.abs(Ew)),
            "B": np.max(np.abs(Bf - Bw) / np.abs(Bw)),
            "KE": np.max(np.abs(K - Kw) / np.abs(Kw)),
            "total": np.max(np.abs(tot - totw) / np.abs(totw))}
    say(label, " ".join("%s %.2e" % kv for kv in errs.items()),
        "div E rms max %.2e (FP64 %.2e)" % (a[:, -1].max(), want[:, -1].max()))
    for k, v in errs.items():
        assert v <= bound[k], (label, k, v, bound[k])
    return errs


def compare_gold_3d(rows, gold):
    """compare_energies (compare_energies.h) over the reference's masks on
    the lines that do not need its random numbers; returns the errors."""
    thr = 10 * np.finfo(np.float32).eps

    def err(a, b):
        with np.errstate(divide="ignore", invalid="ignore"):
            return np.abs(a - b) / np.minimum(a, b)
    out = {}
    E, Eg = rows[:, 1:4].sum(1), gold[:, 1:4].sum(1)
    m = np.minimum(np.abs(E), np.abs(Eg)) >= thr
    out["E"] = float(err(E, Eg)[m].max())
    assert out["E"] < 0.3, out
    P, Pg = rows[:, 6] + rows[:, 7], gold[:, 6] + gold[:, 7]
    out["particle"] = float(err(P, Pg).max())
    assert out["particle"] < 0.01, out
    Bs, Bg = rows[1:, 4:7].sum(1), gold[1:, 4:7].sum(1)
    out["B_mean"] = float(abs(Bs.mean() / Bg.mean() - 1))
    assert out["B_mean"] < 0.15, out
    say("3d_test vs gold:", out)
    return out


# --------------------------------------------------------------------------
# CPU: pins of the FP64 oracle, and the bounds measured on the oracle
# --------------------------------------------------------------------------

def test_fp64_oracle_one_step_matches_fp32():
    """One step of the same state (random E and B, thermal particles, walls
    on x) in the FP32 and FP64 builds agrees to FP32 rounding."""
    g = T.grid_periodic(0, 0, 0, 6, 5, 4, 6, 5, 4,
                        0.95 * T.courant_length(6, 5, 4, 6, 5, 4))
    for k in (0, 3):
        g.fbc[k] = T.BC_PEC_FIELDS
        g.pbc[k] = T.BC_REFLECT_PARTICLES
    f = H.random_fields(g, 3, 0.05, 0.05)
    p = H.thermal_particles(g, 16, 0.2, 4, w=1.0 / 16)
    runs = [OracleRun(g, [("e", -1.0, 1.0, p)], fp64=fp, fields=f,
                      threads=1) for fp in (False, True)]
    for r in runs:
        r.load_interpolator()
        r.advance()
    a, b = runs[0].species[0][3], runs[1].species[0][3]
    assert np.array_equal(a["i"], b["i"])
    for c in ("dx", "dy", "dz", "ux", "uy", "uz"):
        d = np.abs(a[c].astype(np.float64) - b[c])
        assert d.max() <= 4e-6 * max(1.0, float(np.abs(b[c]).max())), c
    for c in ("ex", "ey", "ez", "cbx", "cby", "cbz", "jfx", "jfy", "jfz"):
        x, y = runs[0].f[c].astype(np.float64), runs[1].f[c]
        assert np.abs(x - y).max() <= 1e-5 * np.abs(y).max(), c


@pytest.mark.parametrize("walls", [False, True], ids=["periodic", "walls"])
def test_fp64_oracle_conserves_charge(walls):
    """The FP64 run conserves charge to FP64 rounding at every step; the
    FP32 run to FP32 rounding."""
    g, sp, f = thermal_box(6, 16, seed=5)
    if walls:
        for k in (0, 3):
            g.fbc[k] = T.BC_PEC_FIELDS
            g.pbc[k] = T.BC_REFLECT_PARTICLES
    for fp64, ceiling in ((True, 1e-13), (False, 2e-6)):
        tr = oracle_trace(g, sp, f, 12, fp64)
        assert tr[:, -1].max() <= ceiling, (fp64, tr[:, -1].max())


def test_trace_bounds_thermal_box_cpu():
    """The bounds the GPU trace is held to, measured on the oracle: the
    FP32 build against the FP64 build over 100 steps of an 8^3 box."""
    g, sp, f = thermal_box(8, 64)
    a = oracle_trace(g, sp, f, 100, False)
    want = oracle_trace(g, sp, f, 100, True)
    compare_traces(a, want, THERMAL, 2, "oracle FP32 thermal 8^3")
    assert a[:, -1].max() <= THERMAL["div"]


def test_trace_bounds_harris_cpu():
    """The same on a small Harris sheet (64 x 1 x 32 cells, walls on x)
    over 200 steps."""
    g, sp, f = harris_sheet(64, 32, 16.0, 8.0, 32)
    a = oracle_trace(g, sp, f, 200, False)
    want = oracle_trace(g, sp, f, 200, True)
    compare_traces(a, want, HARRIS, 2, "oracle FP32 Harris 64x32")
    assert a[:, -1].max() <= HARRIS["div"]


def test_3d_test_against_gold_oracle():
    """The reference's energies_gold.3d_test against the FP32 oracle's
    run of the same deck (our random numbers)."""
    g, sp, f = box_3d_test()
    r = OracleRun(g, sp, fields=f)
    r.initialize()
    rows = []
    for s in range(51):
        if s:
            r.advance()
        ef, ep = r.energies()
        rows.append(np.r_[s, ef, ep])
    r.close()
    compare_gold_3d(np.array(rows), np.loadtxt(GOLD_3D))


# --------------------------------------------------------------------------
# GPU: the device path's traces against the FP64 oracle
# --------------------------------------------------------------------------

def gpu_run(g, species, fields, steps, sort_interval=20, every=1,
            progress=None):
    """The device step loop (vpic_amd.sim.Simulation) from the same
    initial state; returns rows as oracle_trace does (every `every`
    steps)."""
    import torch
    from vpic_amd import _lib
    from vpic_amd import sim as S
    _lib.call("vpic_hip_init", 0)
    s = S.Simulation(g)
    if fields is not None:
        s.fields.upload(fields)
    for name, q, m, p in species:
        S.host_species(s, name, q, m, p, sort_interval=sort_interval)
    diag = s.initialize()
    rows = []
    for k in range(steps):
        s.advance()
        if (k + 1) % every == 0:
            ef = s.energy_f()
            ep = [s.energy_p(sp) for sp in s.species]
            rows.append(np.r_[ef, ep, s.div_e_rms()])
        if progress and (k + 1) % progress == 0:
            say("step", k + 1, "E+B %.5e" % (rows[-1][:6].sum()),
                "KE %.8e" % sum(rows[-1][6:-1]))
    torch.cuda.synchronize()
    for sp in s.species:
        nm, lost, bad = s.check_movers(sp)
        assert (nm, lost, bad) == (0, 0, 0), (sp.name, nm, lost, bad)
    return np.array(rows), diag, s


@pytest.mark.gpu
def test_thermal_box_trace_gpu():
    """The plain-method plasma at its CPU-runnable size: 32^3 cells,
    2 x 64 ppc (4,194,304 particles), 100 steps, sorted every 20 (the
    tiled push), against the FP64 oracle."""
    g, sp, f = thermal_box(32, 64)
    say("32^3 x 2 x 64 ppc: GPU run")
    got, diag, _ = gpu_run(g, sp, f, 100)
    say("FP64 oracle run")
    want = oracle_trace(g, sp, f, 100, True)
    compare_traces(got, want, THERMAL, 2, "GPU thermal 32^3")
    assert got[:, -1].max() <= THERMAL["div"]


@pytest.mark.gpu
def test_3d_test_trace_gpu():
    """The reference's 3d_test deck on the device: against the gold file as
    in the CPU test, and against the FP64 oracle's run of the same state."""
    g, sp, f = box_3d_test()
    import torch
    from vpic_amd import _lib
    from vpic_amd import sim as S
    _lib.call("vpic_hip_init", 0)
    s = S.Simulation(g)
    for name, q, m, p in sp:
        S.host_species(s, name, q, m, p)
    s.initialize()
    rows = []
    for k in range(51):
        if k:
            s.advance()
        rows.append(np.r_[k, s.energy_f(), [s.energy_p(x) for x in s.species],
                          s.div_e_rms()])
    torch.cuda.synchronize()
    rows = np.array(rows)
    compare_gold_3d(rows[:, :-1], np.loadtxt(GOLD_3D))
    want = oracle_trace(g, sp, f, 50, True)
    compare_traces(rows[1:, 1:], want, THERMAL, 2, "GPU 3d_test")


@pytest.mark.gpu
def test_harris_trace_gpu():
    """A reduced Harris sheet (128 x 1 x 64 cells, 64 ppc per species,
    walls on x) for 400 steps against the FP64 oracle."""
    g, sp, f = harris_sheet(128, 64, 16.0, 8.0, 64)
    say("Harris 128x1x64: GPU run")
    got, diag, _ = gpu_run(g, sp, f, 400)
    assert diag["div_b_rms"] <= 1e-6 * (1.0 / 3.0)
    say("FP64 oracle run")
    want = oracle_trace(g, sp, f, 400, True)
    compare_traces(got, want, HARRIS, 2, "GPU Harris 128x64")
    assert got[:, -1].max() <= HARRIS["div"]


@pytest.mark.gpu
@pytest.mark.timeout(1200)
def test_harris_full_size_gpu():
    """The Harris sheet at full size: 1024 x 1 x 512 cells (128 x 64 d_e),
    200 ppc per species on average (104,857,600 ions and as many
    electrons), 2000 steps, sorted every 20, through size-independent
    bounds: total energy (FP64 device sums) changes by at most 1% over the
    run, the rms div E error stays at FP32 rounding (<= 2e-6 of n0) at every
    sampled step, no particle leaves the box or is lost, and B starts
    divergence free."""
    say("Harris 1024x1x512, 2 x 200 ppc: loading")
    g, sp, f = harris_sheet(1024, 512, 128.0, 64.0, 200, seed=11)
    assert sum(len(p) for _, _, _, p in sp) == 209715200
    got, diag, s = gpu_run(g, sp, f, 2000, every=20, progress=200)
    del sp
    say("init diagnostics", diag)
    assert diag["div_b_rms"] <= 1e-6
    tot = got[:, :6].sum(1) + got[:, 6:-1].sum(1)
    drift = float(np.max(np.abs(tot - tot[0])) / tot[0])
    say("total energy max change %.3e, div E rms max %.3e"
        % (drift, got[:, -1].max()))
    assert drift <= 1e-2
    assert got[:, -1].max() <= HARRIS["div"]
    for x in s.species:
        assert x.np == 104857600


@pytest.mark.gpu
def test_unequal_species_sorted_alternately_gpu():
    """Two species of different sizes (20 and 8 ppc) sorted at different
    intervals (every 2 and every 3 steps), so that each sort runs on the
    species' own buffers whatever the other did: 12 steps stay whole (every
    particle kept, the partitions cover each species) and follow the FP64
    oracle's trace."""
    n = 8
    dt = 0.95 * T.courant_length(n, n, n, n, n, n)
    g = T.grid_periodic(0, 0, 0, n, n, n, n, n, n, dt)
    pe = H.thermal_particles(g, 20, 0.1, 21, w=1.0 / 20)
    pi = H.thermal_particles(g, 8, 0.02, 22, w=1.0 / 8)
    sp = [("electron", -1.0, 1.0, pe), ("ion", 1.0, 25.0, pi)]
    import torch
    from vpic_amd import _lib
    from vpic_amd import sim as S
    _lib.call("vpic_hip_init", 0)
    s = S.Simulation(g)
    S.host_species(s, "electron", -1.0, 1.0, pe, sort_interval=2)
    S.host_species(s, "ion", 1.0, 25.0, pi, sort_interval=3)
    s.initialize()
    rows = []
    for _ in range(12):
        s.advance()
        rows.append(np.r_[s.energy_f(), [s.energy_p(x) for x in s.species],
                          s.div_e_rms()])
    torch.cuda.synchronize()
    for x, p in zip(s.species, (pe, pi)):
        assert x.np == len(p) and x.p.shape[0] == len(p)
        assert x.aux.shape[0] == len(p)
        part = x.partition.download()
        assert part[g.nv] == x.np
        assert np.all(np.diff(part.astype(np.int64)) >= 0)
        got = x.p.download()[:x.np]
        assert np.array_equal(np.sort(got["w"]), np.sort(p["w"]))
    want = oracle_trace(g, sp, None, 12, True)
    compare_traces(np.array(rows), want, THERMAL, 2, "GPU unequal species")
