"""The benchmark plasma at its full size (256x128x128 cells, 2 x 100 ppc,
838,860,800 particles), checked through properties that do not depend on
the size:

  * charge continuity at every node: rho(t+dt) - rho(t) + dt div J = 0 for
    the step's currents (accumulate_rho_p + synchronize_rho before and
    after, jf from unload_accumulator_array).  Every particle's current
    must be counted exactly once, whichever path it took (the tile's LDS
    window, a neighbour's halo, an out-of-window sub-bucket, a crosser's
    move_p, a float atomic), so a dropped or doubled deposit of a single
    particle (q w = 0.01 here) breaks it by orders of magnitude;
  * bit-exact particles on a random sample: a million particles of each
    species, pushed again by the oracle from their state before the step
    with the same interpolators, equal the GPU's results bit for bit;
  * no particle leaves or is lost (periodic box: no movers), and the
    sort-time partition still indexes every particle.
Taken mid sort interval (13 steps after a sort), when the drift paths are
busy.
"""
import ctypes
import importlib
import sys

import numpy as np
import pytest

from oracle.oracle import host_grid
from vpic_amd import types as T

import helpers as H

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def _rho(s, _lib):
    G = ctypes.byref(s.g)
    _lib.call("vpic_hip_clear_rhof", s.fields.ptr, G, None)
    for sp in s.species:
        _lib.call("vpic_hip_accumulate_rho_p", s.fields.ptr, sp.p.ptr, sp.np,
                  sp.q, G, None)
    _lib.call("vpic_hip_synchronize_rho", s.fields.ptr, G, None)
    return s.fields.download()["rhof"].astype(np.float64)


def _take(sp, idx):
    """Particles idx (a torch index on the device) of species sp."""
    import torch
    t = sp.p.t[:sp.np * 32].view(torch.int32).view(sp.np, 8)
    return t.index_select(0, idx).cpu().numpy().view(T.PARTICLE).reshape(-1)


def test_full_size_plasma_step(orc):
    import torch
    from vpic_amd import _lib
    assert torch.cuda.is_available()
    _lib.call("vpic_hip_init", 0)
    bench = importlib.import_module("bench")
    say("building the 256x128x128 plasma, 2 x 100 ppc")
    s = bench.build_sim(256, 128, 128, 100)
    g = s.g
    np_total = sum(sp.np for sp in s.species)
    assert np_total == 838860800
    for _ in range(13):
        s.advance()
    torch.cuda.synchronize()
    say("step", s.step, "(13 after the sort)")
    rng = np.random.default_rng(7)
    samples = []
    for sp in s.species:
        idx = np.sort(rng.choice(sp.np, size=1 << 20, replace=False))
        tidx = torch.from_numpy(idx).to("cuda")
        samples.append((sp, tidx, _take(sp, tidx)))
    fi = s.interp.download().copy()
    rho0 = _rho(s, _lib)
    s.advance()
    torch.cuda.synchronize()
    nm, lost, bad = s.check_movers()
    assert (nm, lost, bad) == (0, 0, 0)
    f = s.fields.download()
    rho1 = _rho(s, _lib)

    # continuity at every node of the periodic box (cells are unit cubes)
    nx, ny, nz = g.nx, g.ny, g.nz
    shp = (nz + 2, ny + 2, nx + 2)
    r0, r1 = rho0.reshape(shp), rho1.reshape(shp)
    jx = f["jfx"].astype(np.float64).reshape(shp)
    jy = f["jfy"].astype(np.float64).reshape(shp)
    jz = f["jfz"].astype(np.float64).reshape(shp)
    nodes = (slice(1, nz + 1), slice(1, ny + 1), slice(1, nx + 1))
    div = (jx[nodes] - np.roll(jx[1:nz + 1, 1:ny + 1, 1:nx + 1], 1, axis=2) +
           jy[nodes] - np.roll(jy[1:nz + 1, 1:ny + 1, 1:nx + 1], 1, axis=1) +
           jz[nodes] - np.roll(jz[1:nz + 1, 1:ny + 1, 1:nx + 1], 1, axis=0))
    resid = (r1[nodes] - r0[nodes]) + float(g.dt) * div
    scale = float(np.max(np.abs(r0[nodes])))
    worst = float(np.max(np.abs(resid)))
    say("continuity: max |d rho + dt div J| = %.3g of max |rho| %.3g" %
        (worst, scale))
    assert worst <= 1e-4 * scale

    # the sampled particles, pushed again by the oracle
    hg = host_grid(g)
    for sp, tidx, before in samples:
        after = _take(sp, tidx)
        mine = before.copy()
        acc = H.acc_array(g)
        nmo, ign, _ = orc.advance_p(mine, sp.q, sp.m, acc, fi, hg)
        assert nmo == 0 and ign == 0
        assert H.same_bits(after, mine), sp.name
        say(sp.name, "1,048,576 sampled particles bit-exact")

    # the partition of the last sort still covers every particle once
    for sp in s.species:
        part = sp.partition.download()
        assert part[0] >= 0 and part[g.nv] == sp.np
        assert np.all(np.diff(part.astype(np.int64)) >= 0)
