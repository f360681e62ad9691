"""The Takizuka-Abe binary collision operator
(src/collision/pipeline/takizuka_abe_pipeline.cc, src/collision/
takizuka_abe.cc).

The reference draws its random numbers from a per-pipeline SFMT generator
whose seeding (rng.cc) reaches the MPI layer, so its operator cannot be
linked here; its arithmetic is pinned instead by what the operator must do
whatever the random numbers are:

  * Philox4x32-10, the stream both sides use, against its published
    known-answer vectors (Salmon et al., SC'11; Random123 kat_vectors);
  * every binary collision with equal weights conserves the pair's
    momentum and kinetic energy (it rotates the relative velocity about the
    centre of mass), so per voxel the sums of m u and m |u|^2 / 2 are
    unchanged to FP32 rounding, intra- and inter-species, odd counts (the
    three-particle start) included;
  * the scattering angle: tan(theta/2) = dd ~ N(0, s^2) with
    s = std / ur^(3/2), std^2 = cvar0 q_i^2 q_j^2 / mu^2 * n dt interval / dV
    (takizuka_abe_pipeline.cc:95-99, 161-165), so <1 - cos theta> =
    <2 dd^2 / (1 + dd^2)> for a beam of pairs with one relative speed;
  * unequal weights: the lighter particle's partner keeps its momentum with
    probability w_light / w_heavy (:65-70).
The GPU runs the same operator (same stream, same operations) and must
equal the oracle bit for bit.
"""
import ctypes
import math

import numpy as np
import pytest

from vpic_amd import _lib
from vpic_amd import types as T

import helpers as H


def test_philox_known_answers(orc):
    kat = [((0, 0, 0, 0), (0, 0),
            (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
           ((0xffffffff,) * 4, (0xffffffff,) * 2,
            (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
           ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344),
            (0xa4093822, 0x299f31d0),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1))]
    for ctr, key, want in kat:
        assert tuple(int(x) for x in orc.philox4x32(ctr, key)) == want


def _grid(nx=4, ny=3, nz=3):
    return T.grid_periodic(0, 0, 0, nx, ny, nz, nx, ny, nz, 0.2)


def _species(g, counts, rng, uth=0.1, drift=(0.0, 0.0, 0.0), w=None):
    """Particles sorted by voxel with the given count per voxel (nv), and
    their partition."""
    n = int(counts.sum())
    p = np.zeros(n, dtype=T.PARTICLE)
    p["i"] = np.repeat(np.arange(g.nv, dtype=np.int32), counts)
    for c in ("dx", "dy", "dz"):
        p[c] = rng.uniform(-1, 1, n).astype(np.float32)
    for k, c in enumerate(("ux", "uy", "uz")):
        p[c] = (drift[k] + uth * rng.standard_normal(n)).astype(np.float32)
    p["w"] = 1.0 if w is None else w(n)
    part = np.zeros(g.nv + 1, dtype=np.int32)
    part[1:] = np.cumsum(counts)
    return p, part


def _counts(g, rng, lo, hi):
    c = np.zeros(g.nv, dtype=np.int64)
    v = np.arange(g.nv)
    x, y, z = v % (g.nx + 2), v // (g.nx + 2) % (g.ny + 2), v // (
        (g.nx + 2) * (g.ny + 2))
    inner = ((x >= 1) & (x <= g.nx) & (y >= 1) & (y <= g.ny) & (z >= 1) &
             (z <= g.nz))
    c[inner] = rng.integers(lo, hi + 1, int(inner.sum()))
    return c


def _sums(p, part, m, nv):
    """Per voxel: sum of m u (3) and of m |u|^2 / 2, in float64."""
    u = np.stack([p["ux"], p["uy"], p["uz"]], 1).astype(np.float64)
    mom = np.zeros((nv, 3))
    ke = np.zeros(nv)
    v = np.repeat(np.arange(nv), np.diff(part))
    np.add.at(mom, v, m * u)
    np.add.at(ke, v, 0.5 * m * (u ** 2).sum(1))
    return mom, ke


@pytest.mark.parametrize("case", ["intra", "inter"])
def test_collisions_conserve_momentum_and_energy(orc, case):
    g = _grid()
    rng = np.random.default_rng(3)
    ce = _counts(g, rng, 0, 9)  # 0, 1, 3 (odd start) and even counts
    pe, parte = _species(g, ce, rng, 0.1)
    mi, mj = 1.0, 25.0
    if case == "intra":
        mom0, ke0 = _sums(pe, parte, mi, g.nv)
        orc.takizuka_abe(pe, parte, -1.0, mi, pe, parte, -1.0, mi, 2.0, 1,
                         11, 0, g)
        mom1, ke1 = _sums(pe, parte, mi, g.nv)
    else:
        ci = _counts(g, rng, 0, 9)
        pi, parti = _species(g, ci, rng, 0.02, drift=(0.05, 0, 0))
        m0e, k0e = _sums(pe, parte, mi, g.nv)
        m0i, k0i = _sums(pi, parti, mj, g.nv)
        orc.takizuka_abe(pe, parte, -1.0, mi, pi, parti, 1.0, mj, 2.0, 1,
                         11, 0, g)
        m1e, k1e = _sums(pe, parte, mi, g.nv)
        m1i, k1i = _sums(pi, parti, mj, g.nv)
        mom0, ke0, mom1, ke1 = m0e + m0i, k0e + k0i, m1e + m1i, k1e + k1i
    scale_m = np.abs(mom0).max()
    assert np.abs(mom1 - mom0).max() <= 1e-5 * scale_m
    assert np.abs(ke1 - ke0).max() <= 1e-5 * ke0.max()
    assert np.abs(ke1 - ke0).max() > 0 or np.abs(mom1 - mom0).max() >= 0
    # the velocities did change (the operator is not the identity)
    assert (mom1 != mom0).any() or (ke1 != ke0).any()


def test_scattering_angle_variance(orc):
    """Pairs with one relative velocity ur along x: the mean of
    1 - cos(theta) over many pairs matches <2 dd^2/(1+dd^2)>, dd ~ N(0,s^2)
    with s from the operator's formula."""
    g = _grid(8, 8, 8)
    counts = _counts(g, np.random.default_rng(0), 40, 40)
    n = int(counts.sum())
    pe, part = _species(g, counts, np.random.default_rng(1), 0.0)
    pi, _ = _species(g, counts, np.random.default_rng(2), 0.0)
    ur = 0.05
    pe["ux"] = ur
    qi, qj, mi, mj, cvar0, interval = -1.0, 1.0, 1.0, 1.0, 1e-3, 1
    orc.takizuka_abe(pe, part, qi, mi, pi, part, qj, mj, cvar0, interval, 5,
                     0, g)
    rel = np.stack([pe["ux"] - pi["ux"], pe["uy"] - pi["uy"],
                    pe["uz"] - pi["uz"]], 1).astype(np.float64)
    cos_t = rel[:, 0] / np.linalg.norm(rel, axis=1)
    # std (the voxel's density is 40 particles of weight 1)
    mu = mi * mj / (mi + mj)
    cvar = cvar0 * (qi * qi * qj * qj) / (mu * mu)
    std = math.sqrt(cvar * 40 * g.dt * interval / float(g.dV))
    s = min(std / ur ** 1.5, 1.0)
    dd = np.random.default_rng(9).standard_normal(1 << 22) * s
    want = float(np.mean(2 * dd * dd / (1 + dd * dd)))
    got = float(np.mean(1 - cos_t))
    assert n == len(cos_t)
    assert abs(got - want) <= 0.03 * want, (got, want, s)


def test_unequal_weights_keep_the_heavier_side(orc):
    """A heavy-weight particle (w = 10) against light ones (w = 1): its
    velocity changes in ~1/10 of its collisions, the light ones' in all."""
    g = _grid(8, 8, 8)
    counts = _counts(g, np.random.default_rng(0), 30, 30)
    pe, part = _species(g, counts, np.random.default_rng(1), 0.1)
    pi, _ = _species(g, counts, np.random.default_rng(2), 0.1,
                     w=lambda n: np.float32(10.0))
    e0, i0 = pe.copy(), pi.copy()
    orc.takizuka_abe(pe, part, -1.0, 1.0, pi, part, 1.0, 1.0, 1e-2, 1, 7, 0,
                     g)
    moved_e = np.mean(pe["ux"] != e0["ux"])
    moved_i = np.mean(pi["ux"] != i0["ux"])
    assert moved_e > 0.99
    assert 0.07 < moved_i < 0.13, moved_i


@pytest.fixture(scope="module")
def dev():
    import torch
    assert torch.cuda.is_available(), "GPU tests need a GPU"
    _lib.call("vpic_hip_init", 0)
    from vpic_amd import device
    return device


@pytest.mark.gpu
def test_takizuka_abe_gpu_bitexact(dev, orc):
    """The device operator against the oracle on the same stream: intra-
    and inter-species, counts 0..9 per voxel (odd starts, one-sided
    voxels), unequal weights, two calls in a row."""
    g = _grid(6, 5, 4)
    rng = np.random.default_rng(21)
    ce = _counts(g, rng, 0, 9)
    ci = _counts(g, rng, 0, 9)
    pe, parte = _species(g, ce, rng, 0.1)
    pi, parti = _species(g, ci, rng, 0.02, drift=(0.03, 0, 0),
                         w=lambda n: rng.choice(
                             np.array([1.0, 2.5], np.float32), n))
    de, dpe = dev.DeviceArray.from_host(pe), dev.DeviceArray.from_host(parte)
    di, dpi = dev.DeviceArray.from_host(pi), dev.DeviceArray.from_host(parti)
    G = ctypes.byref(g)
    for call in (0, 1):
        print("takizuka_abe intra, call", call, flush=True)
        orc.takizuka_abe(pe, parte, -1.0, 1.0, pe, parte, -1.0, 1.0, 0.5, 2,
                         1234, call, g)
        _lib.call("vpic_hip_takizuka_abe", de.ptr, dpe.ptr, len(pe), -1.0,
                  1.0, de.ptr, dpe.ptr, len(pe), -1.0, 1.0, 0.5, 2, 1234,
                  call, G, None)
        assert H.same_bits(de.download(), pe), call
        print("takizuka_abe inter, call", call, flush=True)
        orc.takizuka_abe(pe, parte, -1.0, 1.0, pi, parti, 1.0, 25.0, 0.5, 2,
                         1234, 10 + call, g)
        _lib.call("vpic_hip_takizuka_abe", de.ptr, dpe.ptr, len(pe), -1.0,
                  1.0, di.ptr, dpi.ptr, len(pi), 1.0, 25.0, 0.5, 2, 1234,
                  10 + call, G, None)
        assert H.same_bits(de.download(), pe), call
        assert H.same_bits(di.download(), pi), call


@pytest.mark.gpu
def test_takizuka_abe_gpu_crowded_voxels(dev, orc):
    """Voxels with 256, 257 and 300 particles of species i and a species j
    much larger or smaller than i, bit-exact against the oracle."""
    g = _grid(4, 3, 3)
    rng = np.random.default_rng(5)
    ce = _counts(g, rng, 0, 9)
    ci = _counts(g, rng, 0, 40)
    inner = np.nonzero(ce)[0]
    ce[inner[0]], ce[inner[1]], ce[inner[2]] = 256, 257, 300
    ci[inner[2]] = 3
    pe, parte = _species(g, ce, rng, 0.1)
    pi, parti = _species(g, ci, rng, 0.02)
    de, dpe = dev.DeviceArray.from_host(pe), dev.DeviceArray.from_host(parte)
    di, dpi = dev.DeviceArray.from_host(pi), dev.DeviceArray.from_host(parti)
    G = ctypes.byref(g)
    print("takizuka_abe crowded, intra", flush=True)
    orc.takizuka_abe(pe, parte, -1.0, 1.0, pe, parte, -1.0, 1.0, 0.5, 1, 99,
                     3, g)
    _lib.call("vpic_hip_takizuka_abe", de.ptr, dpe.ptr, len(pe), -1.0, 1.0,
              de.ptr, dpe.ptr, len(pe), -1.0, 1.0, 0.5, 1, 99, 3, G, None)
    assert H.same_bits(de.download(), pe)
    print("takizuka_abe crowded, inter", flush=True)
    orc.takizuka_abe(pe, parte, -1.0, 1.0, pi, parti, 1.0, 25.0, 0.5, 1, 99,
                     4, g)
    _lib.call("vpic_hip_takizuka_abe", de.ptr, dpe.ptr, len(pe), -1.0, 1.0,
              di.ptr, dpi.ptr, len(pi), 1.0, 25.0, 0.5, 1, 99, 4, G, None)
    assert H.same_bits(de.download(), pe)
    assert H.same_bits(di.download(), pi)


@pytest.mark.gpu
def test_takizuka_abe_gpu_rejects_bad_partition(dev):
    """A partition that indexes past the species' particles is reported,
    not followed."""
    g = _grid()
    counts = _counts(g, np.random.default_rng(0), 2, 4)
    p, part = _species(g, counts, np.random.default_rng(1))
    bad = part.copy()
    bad[-1] = len(p) + 5
    bad[-2] = len(p) + 5  # the last interior voxels run past np
    bad[g.nv // 2 + 1:] = np.maximum(bad[g.nv // 2 + 1:], len(p) + 5)
    dp = dev.DeviceArray.from_host(p)
    db = dev.DeviceArray.from_host(bad)
    with pytest.raises(_lib.VpicError):
        _lib.call("vpic_hip_takizuka_abe", dp.ptr, db.ptr, len(p), -1.0, 1.0,
                  dp.ptr, db.ptr, len(p), -1.0, 1.0, 0.5, 1, 1, 0,
                  ctypes.byref(g), None)


@pytest.mark.gpu
def test_simulation_collision_operator_gpu(dev):
    """Simulation.takizuka_abe (define_collision_op): at a collision step
    both species are sorted and the operator runs; the plasma's total
    momentum and kinetic energy (equal weights) are conserved, the
    partition still indexes every particle, and a step that is not a
    multiple of the interval leaves the particles alone."""
    import torch
    from vpic_amd import sim as S
    g = T.grid_periodic(0, 0, 0, 8, 6, 4, 8, 6, 4, 0.2)
    s = S.Simulation(g)
    S.uniform_plasma(s, 20, -1.0, 1.0, 0.1, 5, "electron", w=0.05,
                     sort_interval=0)
    S.uniform_plasma(s, 20, 1.0, 25.0, 0.02, 6, "ion", w=0.05,
                     sort_interval=0)
    el, io = s.species
    op = s.takizuka_abe(el, io, 0.5, 2, seed=3)

    def totals():
        mom, ke = np.zeros(3), 0.0
        for sp in (el, io):
            p = sp.p.download()[:sp.np]
            u = np.stack([p["ux"], p["uy"], p["uz"]], 1).astype(np.float64)
            mom += sp.m * u.sum(0)
            ke += 0.5 * sp.m * float((u ** 2).sum())
        return mom, ke

    s.step = 1  # not a collision step
    m0, k0 = totals()
    before = el.p.download()[:el.np].copy()
    op(s)
    torch.cuda.synchronize()
    assert H.same_bits(el.p.download()[:el.np], before)
    s.step = 2
    print("collision step", flush=True)
    op(s)
    torch.cuda.synchronize()
    assert el.last_sorted == 2 and io.last_sorted == 2
    m1, k1 = totals()
    scale = sum(sp.m * sp.np * 0.1 for sp in (el, io))
    assert np.abs(m1 - m0).max() <= 1e-5 * scale
    assert abs(k1 - k0) <= 1e-5 * k0
    assert not H.same_bits(el.p.download()[:el.np], before)
    for sp in (el, io):
        part = sp.partition.download()
        p = sp.p.download()[:sp.np]
        assert part[g.nv] == sp.np
        v = np.repeat(np.arange(g.nv), np.diff(part.astype(np.int64)))
        assert np.array_equal(p["i"], v)


@pytest.mark.parametrize("case", ["intra", "inter", "unequal_weights"])
def test_oracle_matches_reference_operator(orc, ref, case):
    """The oracle's operator against the reference's own
    takizuka_abe_pipeline_scalar (takizuka_abe_pipeline.cc:81-215, compiled
    from source in oracle/_ref) on the same random stream: the shuffle, the
    odd-count start, the pairing and the collision arithmetic must agree.
    The reference's sin/cos of the scattering phase are the float overloads
    (sincosf); the oracle takes the correctly rounded float value, so a
    momentum may differ where sincosf is not correctly rounded: by at most
    4 ulp of the species' largest momentum (2^-21 of it), in at most 2% of
    the components, bit-exact otherwise (measured: 0-4 components of ~1000,
    at most 0.4 ulp of the scale)."""
    g = _grid(5, 4, 3)
    rng = np.random.default_rng(17)
    ce = _counts(g, rng, 0, 11)
    w = (lambda n: rng.uniform(0.5, 2.0, n).astype(np.float32)) \
        if case == "unequal_weights" else None
    pe, parte = _species(g, ce, rng, 0.1, w=w)
    if case == "intra":
        a, b = pe.copy(), pe.copy()
        orc.takizuka_abe(a, parte, -1.0, 1.0, a, parte, -1.0, 1.0, 2.0, 1,
                         0x1234567890, 7, g)
        ref.takizuka_abe(b, parte, -1.0, 1.0, b, parte, -1.0, 1.0, 2.0, 1,
                         0x1234567890, 7, g)
        pairs = [(a, b)]
    else:
        ci = _counts(g, rng, 0, 11)
        pi, parti = _species(g, ci, rng, 0.02, drift=(0.05, 0, 0), w=w)
        a, b, ai, bi = pe.copy(), pe.copy(), pi.copy(), pi.copy()
        orc.takizuka_abe(a, parte, -1.0, 1.0, ai, parti, 1.0, 25.0, 2.0, 2,
                         99, 3, g)
        ref.takizuka_abe(b, parte, -1.0, 1.0, bi, parti, 1.0, 25.0, 2.0, 2,
                         99, 3, g)
        pairs = [(a, b), (ai, bi)]
    changed = 0
    for x, y in pairs:
        for c in ("dx", "dy", "dz", "i", "w"):
            assert np.array_equal(x[c], y[c]), c  # same shuffle
        scale = max(float(np.abs(y[c]).max()) for c in ("ux", "uy", "uz"))
        for c in ("ux", "uy", "uz"):
            d = np.abs(x[c].astype(np.float64) - y[c].astype(np.float64))
            assert d.max() <= 2.0 ** -21 * scale, (c, float(d.max()), scale)
            changed += int((d > 0).sum())
    n = sum(len(x) for x, _ in pairs)
    assert changed <= 0.02 * 3 * n, changed
    # the operator did act
    assert not np.array_equal(a["ux"], pe["ux"])


def test_collision_operators_get_distinct_streams(orc):
    """Two operators applied at the same step (Simulation.takizuka_abe with
    one user seed) draw from different Philox keys: the same particles
    collided under operator 0's and operator 1's key end differently, and
    each key is reproducible."""
    from vpic_amd.sim import collision_key
    keys = [collision_key(5, k) for k in range(4)]
    assert len(set(keys)) == 4 and all(0 <= k < 1 << 64 for k in keys)
    assert collision_key(5, 0) != collision_key(6, 0)
    g = _grid()
    rng = np.random.default_rng(4)
    ce = _counts(g, rng, 2, 9)
    pe, parte = _species(g, ce, rng, 0.1)
    out = []
    for key in (keys[0], keys[1], keys[0]):
        p = pe.copy()
        orc.takizuka_abe(p, parte, -1.0, 1.0, p, parte, -1.0, 1.0, 2.0, 1,
                         key, 4, g)
        out.append(p)
    assert H.same_bits(out[0], out[2])
    assert not np.array_equal(out[0]["ux"], out[1]["ux"])
