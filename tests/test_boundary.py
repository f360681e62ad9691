"""boundary_p for one domain: the mover loop (absorb -> rhob, custom
boundary hits, removal by backfill) and the injection loop.

CPU tests pin the oracle: accumulate_rhob bit-exact against the reference's
own function (rho_p.cc, compiled from source), closed forms of the
absorbed charge, a hand-traced backfill, and the invariants of removal.
GPU tests compare the HIP path with the oracle on movers that the pinned
advance_p produces against absorbing and custom walls: particles, hits and
counts bit-exact, rhob to 1e-5 (the order of the absorbed charges'
additions differs), injection bit-exact (one lane, the reference's order).
"""
import ctypes
import sys

import numpy as np
import pytest

from oracle.oracle import host_grid
from vpic_amd import types as T

import helpers as H


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _box(nx, ny, nz, cell=1.0, frac=0.95):
    L = (nx * cell, ny * cell, nz * cell)
    dt = frac * T.courant_length(L[0], L[1], L[2], nx, ny, nz)
    return T.grid_periodic(0, 0, 0, L[0], L[1], L[2], nx, ny, nz, dt)


def _pushed(orc, g, ppc, uth, seed, q=-1.0, m=1.0):
    """Particles after one oracle push against g's walls, and its movers."""
    f = H.random_fields(g, seed, 0.2, 0.2)
    fi = H.interp_array(g)
    orc.load_interpolator(fi, f, g)
    p = H.thermal_particles(g, ppc, uth, seed + 1, shuffle=True)
    acc = H.acc_array(g)
    hg = host_grid(g)
    nm, ign, pm = orc.advance_p(p, q, m, acc, fi, hg, max_nm=len(p))
    assert ign == 0
    return p, pm[:nm].copy(), hg


# -- oracle pins (CPU) --------------------------------------------------------

def test_accumulate_rhob_bitexact(orc, ref):
    """Random particles over every voxel, faces included (the weights are
    doubled there), against the reference's accumulate_rhob."""
    g = _box(5, 4, 6)
    hg = host_grid(g)
    p = H.thermal_particles(g, 3, 0.1, 17, shuffle=True)
    fa = H.random_full_fields(g, 3)
    fb = fa.copy()
    orc.accumulate_rhob(fa, p, -1.5, hg)
    ref.accumulate_rhob(fb, p, -1.5, hg)
    assert H.same_bits(fa, fb)


def test_accumulate_rhob_closed_forms(orc):
    """A charge at the centre of an interior voxel puts q w r8V on each of
    its 8 nodes; in the corner voxel (1,1,1) the three low faces double
    the low nodes' weights (node (1,1,1) gets 8x)."""
    g = _box(5, 5, 5)
    hg = host_grid(g)
    r8V = float(np.float32(g.r8V))
    for (x, y, z) in ((3, 3, 3), (1, 1, 1)):
        p = np.zeros(1, dtype=T.PARTICLE)
        p["i"] = g.voxel(x, y, z)
        p["w"] = 2.0
        f = np.zeros(g.nv, dtype=T.FIELD)
        orc.accumulate_rhob(f, p, 0.75, hg)
        for dz in (0, 1):
            for dy in (0, 1):
                for dx in (0, 1):
                    mult = 1
                    if (x, y, z) == (1, 1, 1):
                        mult = 2 ** ((dx == 0) + (dy == 0) + (dz == 0))
                    v = g.voxel(x + dx, y + dy, z + dz)
                    want = np.float32(0.75 * r8V) * np.float32(2.0) * mult
                    assert f["rhob"][v] == want, ((x, y, z), dx, dy, dz)
        assert np.count_nonzero(f["rhob"]) == 8


def test_boundary_p_backfill_known_answer(orc):
    """15 particles, movers at 2 5 6 7 8 11 13 (all absorbed): removing them
    last first, each hole taking the array's last particle, leaves
    0 1 12 3 4 9 10 14 (traced by hand: 13<-14, 11<-14, 8<-12, 7<-14,
    6<-10, 5<-9, 2<-12)."""
    g = _box(4, 4, 4)
    g.pbc[0] = g.pbc[3] = T.BC_ABSORB_PARTICLES
    hg = host_grid(g)
    p = np.zeros(15, dtype=T.PARTICLE)
    p["w"] = np.arange(15, dtype=np.float32)
    p["i"] = g.voxel(2, 2, 2)
    movers = [2, 5, 6, 7, 8, 11, 13]
    for k in movers:
        p["i"][k] = 8 * g.voxel(1, 2, 2) + 0  # on the -x face: absorbing
        p["dx"][k] = -1.0
    pm = np.zeros(len(movers), dtype=T.MOVER)
    pm["i"] = movers
    f = np.zeros(g.nv, dtype=T.FIELD)
    n, hits, counts = orc.boundary_p(p, len(p), pm, f, -1.0, hg)
    assert n == 8
    assert list(p["w"][:n].astype(int)) == [0, 1, 12, 3, 4, 9, 10, 14]
    assert list(counts) == [0, 0]
    assert np.all(hits["code"] == T.BC_ABSORB_PARTICLES)
    assert np.all(hits["p"]["i"] == g.voxel(1, 2, 2))
    # absorbed charge (w = the particle's index): each particle sits on the
    # -x face (dx = -1, dy = dz = 0), so its 4 low-x nodes get 2 q w r8V,
    # doubled on the domain's low x face: 16 q w r8V in all
    qw = -1.0 * sum(movers)
    assert np.isclose(np.sum(f["rhob"], dtype=np.float64),
                      16 * qw * float(np.float32(g.r8V)), rtol=1e-6)


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_boundary_p_invariants(orc, seed):
    """Random movers: the survivors are the original particles minus the
    movers' (as a multiset), np drops by nm, and custom / unknown codes are
    reported, not absorbed."""
    rng = np.random.default_rng(seed)
    g = _box(4, 5, 3)
    g.pbc[0] = T.BC_ABSORB_PARTICLES
    g.pbc[3] = -3          # a custom particle boundary
    g.pbc[4] = -7          # another one
    hg = host_grid(g)
    N = int(rng.integers(1, 400))
    p = H.thermal_particles(g, 7, 0.1, seed)[:N].copy()
    p["w"] = np.arange(N, dtype=np.float32)
    k = int(rng.integers(0, N + 1))
    movers = np.sort(rng.choice(N, size=k, replace=False))
    faces = rng.choice([0, 3, 4], size=k)
    for j, (i, fc) in enumerate(zip(movers, faces)):
        x, y = (1 if fc == 0 else 4), (5 if fc == 4 else 2)
        p["i"][i] = 8 * g.voxel(x, y, 2) + int(fc)
    pm = np.zeros(k, dtype=T.MOVER)
    pm["i"] = movers
    pm["dispx"] = rng.uniform(-1, 1, k)
    before = p.copy()
    f = np.zeros(g.nv, dtype=T.FIELD)
    n, hits, counts = orc.boundary_p(p, N, pm, f, 1.0, hg)
    assert n == N - k
    keep = np.setdiff1d(np.arange(N), movers)
    assert sorted(p["w"][:n].astype(int)) == list(keep)
    assert counts[0] == np.count_nonzero(faces != 0)
    assert counts[1] == 0
    want = np.where(faces == 0, T.BC_ABSORB_PARTICLES,
                    np.where(faces == 3, -3, -7))
    assert list(hits["code"]) == list(want)
    assert H.same_bits(hits["m"], pm)
    for j, i in enumerate(movers):
        assert hits["p"][j]["w"] == before["w"][i]
        assert hits["p"][j]["i"] == before["i"][i] >> 3
    total = -np.sum(f["rhob"], dtype=np.float64)
    assert (total != 0) == bool(np.any(faces == 0))


def test_inject_p_closed_forms(orc):
    """Injectors last first; one with no displacement is appended as it is
    and deposits nothing; one that leaves through an absorbing wall keeps a
    mover with its index; a full array drops and counts."""
    g = _box(4, 4, 4)
    g.pbc[0] = g.pbc[3] = T.BC_ABSORB_PARTICLES
    hg = host_grid(g)
    pi = np.zeros(3, dtype=T.INJECTOR)
    pi["i"] = [g.voxel(2, 2, 2), g.voxel(1, 3, 3), g.voxel(3, 1, 2)]
    pi["w"] = [1.0, 2.0, 3.0]
    pi["dispx"][1] = -3.0      # leaves through -x (absorbing)
    p = np.zeros(10, dtype=T.PARTICLE)
    pm = np.zeros(10, dtype=T.MOVER)
    acc = H.acc_array(g)
    n, nm, counts = orc.inject_p(p, 4, pm, 0, pi, acc, -1.0, hg)
    assert (n, nm) == (7, 1) and list(counts) == [0, 0]
    assert list(p["w"][4:7]) == [3.0, 2.0, 1.0]        # last first
    assert pm["i"][0] == 5 and p["i"][5] == 8 * g.voxel(1, 3, 3) + 0
    assert p["dx"][5] == -1.0
    assert H.same_bits(p[4:5]["i"], pi[2:3]["i"])
    # the full-array branch
    p2 = np.zeros(5, dtype=T.PARTICLE)
    n2, nm2, c2 = orc.inject_p(p2, 4, pm, 0, pi, H.acc_array(g), -1.0, hg)
    assert (n2, list(c2)) == (5, [2, 0])


def test_inject_p_matches_move_p(orc, ref):
    """Each injection is an append plus the pinned move_p: replaying it with
    the reference's move_p gives the same particles and currents."""
    g = _box(5, 4, 4)
    g.pbc[1] = g.pbc[4] = T.BC_ABSORB_PARTICLES
    hg = host_grid(g)
    rng = np.random.default_rng(9)
    n = 40
    pi = np.zeros(n, dtype=T.INJECTOR)
    src = H.thermal_particles(g, 1, 0.1, 4, shuffle=True)[:n]
    for c in ("dx", "dy", "dz", "i", "ux", "uy", "uz", "w"):
        pi[c] = src[c]
    for c in ("dispx", "dispy", "dispz"):
        pi[c] = rng.uniform(-2.5, 2.5, n).astype(np.float32)
    p = np.zeros(64, dtype=T.PARTICLE)
    pm = np.zeros(64, dtype=T.MOVER)
    acc = H.acc_array(g)
    np_, nm, counts = orc.inject_p(p, 0, pm, 0, pi, acc, 0.5, hg)
    pr = np.zeros(64, dtype=T.PARTICLE)
    accr = H.acc_array(g)
    nmr = 0
    for j, k in enumerate(range(n - 1, -1, -1)):
        for c in ("dx", "dy", "dz", "i", "ux", "uy", "uz", "w"):
            pr[c][j] = pi[c][k]
        mv = np.zeros(1, dtype=T.MOVER)
        mv["dispx"], mv["dispy"], mv["dispz"] = (pi["dispx"][k],
                                                 pi["dispy"][k],
                                                 pi["dispz"][k])
        mv["i"] = j
        nmr += ref.move_p(pr, mv, accr, hg, 0.5)
    assert np_ == n and nm == nmr and nm > 0
    assert H.same_bits(p[:n], pr[:n])
    assert H.same_bits(acc, accr)


# -- HIP path (GPU) ---------------------------------------------------------------

@pytest.fixture(scope="module")
def dev():
    import torch
    from vpic_amd import _lib
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    _lib.call("vpic_hip_init", 0)
    from vpic_amd import device
    return device


def _gpu_boundary(dev, p, pm, f, q, g):
    from vpic_amd import _lib
    dp = dev.DeviceArray.from_host(p)
    dpm = dev.DeviceArray.from_host(pm) if len(pm) else None
    df = dev.DeviceArray.from_host(f)
    dh = dev.DeviceArray((max(len(pm), 1),), T.BOUNDARY_HIT)
    npc = ctypes.c_int32(len(p))
    counts = (ctypes.c_int32 * 2)()
    _lib.call("vpic_hip_boundary_p", dp.ptr, ctypes.byref(npc),
              dpm.ptr if dpm else None, len(pm), df.ptr, q, ctypes.byref(g),
              dh.ptr, counts, None)
    return (dp.download(), npc.value, dh.download()[:len(pm)], df.download(),
            list(counts))


@pytest.mark.gpu
@pytest.mark.parametrize("walls", ["absorb", "custom"])
def test_boundary_p_gpu(dev, orc, walls):
    """Movers from a push against absorbing z walls (and custom x walls):
    the same particles in the same slots, hits and counts as the oracle;
    rhob to 1e-5."""
    g = _box(6, 5, 7)
    g.pbc[2] = g.pbc[5] = T.BC_ABSORB_PARTICLES
    if walls == "custom":
        g.pbc[0] = g.pbc[3] = -3
    p, pm, hg = _pushed(orc, g, 9, 0.7, 31)
    assert len(pm) > 20
    f = H.random_full_fields(g, 8)
    po, fo = p.copy(), f.copy()
    no, ho, co = orc.boundary_p(po, len(po), pm, fo, -1.0, hg)
    say("boundary_p", walls, "np", len(p), "nm", len(pm))
    pg, ng, hgot, fg, cg = _gpu_boundary(dev, p, pm, f, -1.0, g)
    assert ng == no and cg == list(co)
    assert H.same_bits(pg[:ng], po[:no])
    assert H.same_bits(hgot, ho)
    den = float(np.max(np.abs(fo["rhob"])))
    assert float(np.max(np.abs(fg["rhob"] - fo["rhob"]))) <= 1e-5 * den
    for c in T.FIELD_FLOATS:
        if c != "rhob":
            assert H.same_bits(fg[c], fo[c]), c


@pytest.mark.gpu
def test_boundary_p_gpu_tail_chains(dev, orc):
    """Movers crowded at the end of the array (long backfill chains) and
    the edge cases: no movers, every particle a mover."""
    g = _box(4, 4, 4)
    g.pbc[0] = g.pbc[3] = T.BC_ABSORB_PARTICLES
    hg = host_grid(g)
    rng = np.random.default_rng(5)
    for N, movers in ((300, list(range(150, 300, 1))),
                      (300, sorted(set(rng.integers(200, 300, 70)) |
                                   {3, 17, 40})),
                      (64, []), (64, list(range(64)))):
        p = H.thermal_particles(g, 5, 0.1, 6)[:N].copy()
        p["w"] = np.arange(N, dtype=np.float32)
        for i in movers:
            p["i"][i] = 8 * g.voxel(4, 2, 3) + 3
            p["dx"][i] = 1.0
        pm = np.zeros(len(movers), dtype=T.MOVER)
        pm["i"] = movers
        f = np.zeros(g.nv, dtype=T.FIELD)
        po, fo = p.copy(), f.copy()
        no, ho, co = orc.boundary_p(po, N, pm, fo, 1.0, hg)
        say("boundary_p chains N", N, "nm", len(movers))
        pg, ng, hgot, fg, cg = _gpu_boundary(dev, p, pm, f, 1.0, g)
        assert ng == no
        assert H.same_bits(pg[:ng], po[:no])


@pytest.mark.gpu
def test_boundary_p_gpu_rejects_bad_movers(dev, orc):
    """Movers out of order, or whose particle is not on a face, are an
    error and nothing is changed."""
    from vpic_amd import _lib
    g = _box(4, 4, 4)
    p = H.thermal_particles(g, 2, 0.1, 6)
    pm = np.zeros(2, dtype=T.MOVER)
    pm["i"] = [7, 3]
    p["i"][3] = 8 * g.voxel(1, 1, 1)
    p["i"][7] = 8 * g.voxel(1, 1, 1)
    dp = dev.DeviceArray.from_host(p)
    dpm = dev.DeviceArray.from_host(pm)
    df = dev.DeviceArray((g.nv,), T.FIELD)
    npc = ctypes.c_int32(len(p))
    counts = (ctypes.c_int32 * 2)()
    say("boundary_p bad movers")
    rc = _lib.load().vpic_hip_boundary_p(dp.ptr, ctypes.byref(npc), dpm.ptr, 2,
                                        df.ptr, 1.0, ctypes.byref(g), None,
                                        counts, None)
    assert rc < 0 and npc.value == len(p)
    assert H.same_bits(dp.download(), p)


@pytest.mark.gpu
def test_inject_p_gpu(dev, orc):
    """Injection against absorbing y walls: particles, movers and currents
    bit-exact (one lane, the reference's order), and the full-array drop."""
    from vpic_amd import _lib
    g = _box(5, 4, 4)
    g.pbc[1] = g.pbc[4] = T.BC_ABSORB_PARTICLES
    hg = host_grid(g)
    rng = np.random.default_rng(19)
    n = 200
    pi = np.zeros(n, dtype=T.INJECTOR)
    src = H.thermal_particles(g, 3, 0.1, 4, shuffle=True)[:n]
    for c in ("dx", "dy", "dz", "i", "ux", "uy", "uz", "w"):
        pi[c] = src[c]
    for c in ("dispx", "dispy", "dispz"):
        pi[c] = rng.uniform(-2.5, 2.5, n).astype(np.float32)
    for cap in (512, 150):
        p = H.thermal_particles(g, 1, 0.1, 2)[:cap // 4].copy()
        n0 = len(p)
        p = np.concatenate([p, np.zeros(cap - n0, dtype=T.PARTICLE)])
        pm = np.zeros(300, dtype=T.MOVER)
        acc = H.acc_array(g)
        flat = acc.view(np.float32).reshape(-1)
        flat[:] = rng.uniform(-1, 1, flat.size)
        po, pmo, acco = p.copy(), pm.copy(), acc.copy()
        no, nmo, co = orc.inject_p(po, n0, pmo, 0, pi, acco, -1.0, hg)
        dp = dev.DeviceArray.from_host(p)
        dpm = dev.DeviceArray.from_host(pm)
        dpi = dev.DeviceArray.from_host(pi)
        da = dev.DeviceArray.from_host(acc)
        npc, nmc = ctypes.c_int32(n0), ctypes.c_int32(0)
        counts = (ctypes.c_int32 * 2)()
        say("inject_p cap", cap)
        _lib.call("vpic_hip_inject_p", dp.ptr, ctypes.byref(npc), cap, dpm.ptr,
                  ctypes.byref(nmc), 300, dpi.ptr, n, da.ptr, -1.0,
                  ctypes.byref(g), counts, None)
        assert (npc.value, nmc.value, list(counts)) == (no, nmo, list(co))
        assert H.same_bits(dp.download()[:no], po[:no])
        assert H.same_bits(dpm.download()[:nmo], pmo[:nmo])
        assert H.same_bits(da.download(), acco)


@pytest.mark.gpu
def test_simulation_absorbing_walls_gpu(orc):
    """The step driver with absorbing z walls: each step's push and
    boundary_p leave the same particles, in the same slots, as the oracle's
    advance_p + boundary_p; rhob gets the absorbed charge."""
    import torch
    from vpic_amd import _lib
    from vpic_amd import sim as S
    assert torch.cuda.is_available()
    _lib.call("vpic_hip_init", 0)
    g = _box(6, 5, 4)
    g.pbc[2] = g.pbc[5] = T.BC_ABSORB_PARTICLES
    s = S.Simulation(g)
    f = H.random_fields(g, 21, 0.05, 0.05)
    s.fields.upload(f)
    s.load_interpolator_array()
    p = H.thermal_particles(g, 6, 0.4, 22, shuffle=True)
    sp = S.Species("electron", -1.0, 1.0, len(p), g)
    sp.p.upload(p)
    s.add_species(sp)
    hg = host_grid(g)
    po = p.copy()
    absorbed = 0
    for step in range(4):
        fi = s.interp.download().copy()
        say("absorbing walls, step", step, "np", sp.np)
        s.advance()
        nmo, ign, pm = orc.advance_p(po, -1.0, 1.0, H.acc_array(g), fi, hg)
        assert ign == 0
        fo = np.zeros(g.nv, dtype=T.FIELD)
        n, hits, counts = orc.boundary_p(po, len(po), pm, fo, -1.0, hg)
        po = po[:n].copy()
        absorbed += nmo
        assert sp.np == n and sp.nm == 0
        assert H.same_bits(sp.p.download()[:n], po)
    assert absorbed > 0
    assert np.any(s.fields.download()["rhob"] != 0)
