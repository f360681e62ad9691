"""Checkpoint and restart (vpic_amd.checkpoint; the reference's checkpt.h:
101-110 checkpt_objects / restore_objects).

Restart parity: a run checkpointed after some steps and continued from the
file in a new Simulation ends bit for bit where the uninterrupted run ends.
Once every species has been sorted the step is deterministic on the device
(the tiled push sums currents in 64-bit fixed point, whose integer sums do
not depend on the order of the additions; deposits outside a tile's window
go to sub-buckets summed the same way, which on a small grid hold every
particle of a tile, so no float atomic is taken; and the merge adds the
windows in a fixed order), so the two continuations must agree in every bit
of the particles, fields, interpolators and accumulators.
"""
import os
import sys

import pytest

import helpers as H
from vpic_amd import types as T

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, file=sys.stderr, flush=True)


def _state(s):
    out = {"fields": s.fields.download().copy(),
           "interp": s.interp.download().copy(),
           "acc": s.acc.download().copy()}
    for k, sp in enumerate(s.species):
        out["p%d" % k] = sp.p.download()[:sp.np].copy()
        out["part%d" % k] = sp.partition.download().copy()
    return out


def _build(n=8, walls=False):
    from vpic_amd import _lib
    from vpic_amd import sim as S
    _lib.call("vpic_hip_init", 0)
    dt = 0.95 * T.courant_length(n, n, n, n, n, n)
    g = T.grid_periodic(0, 0, 0, n, n, n, n, n, n, dt)
    if walls:
        g.pbc[0] = g.pbc[3] = T.BC_REFLECT_PARTICLES
    s = S.Simulation(g)
    pe = H.thermal_particles(g, 12, 0.2, 31, w=1.0 / 12)
    pi = H.thermal_particles(g, 6, 0.03, 32, w=1.0 / 6)
    S.host_species(s, "electron", -1.0, 1.0, pe, sort_interval=4)
    S.host_species(s, "ion", 1.0, 25.0, pi, sort_interval=5)
    s.initialize()
    return s


@pytest.mark.parametrize("walls", [False, True], ids=["periodic", "walls"])
def test_restart_parity(tmp_path, walls):
    from vpic_amd import sim as S
    s = _build(walls=walls)
    for _ in range(6):
        s.advance()
    path = os.path.join(str(tmp_path), "run.ckpt")
    s.checkpoint(path)
    say("checkpoint at step", s.step, "size", os.path.getsize(path))
    for _ in range(7):
        s.advance()
    want = _state(s)
    r = S.Simulation.restore(path)
    r2 = S.Simulation.restore(path)
    assert r.step == 6
    assert [sp.name for sp in r.species] == ["electron", "ion"]
    assert [sp.sort_interval for sp in r.species] == [4, 5]
    for _ in range(7):
        r.advance()
        r2.advance()
    got, got2 = _state(r), _state(r2)
    assert r.step == s.step == 13
    for key in want:
        assert H.same_bits(got[key], got2[key]), key
        assert H.same_bits(got[key], want[key]), key


def test_checkpoint_rejects_other_files(tmp_path):
    from vpic_amd import _lib
    from vpic_amd import sim as S
    path = os.path.join(str(tmp_path), "not.ckpt")
    with open(path, "wb") as fh:
        fh.write(b"something else entirely\n")
    with pytest.raises(_lib.VpicError, match="not a checkpoint"):
        S.Simulation.restore(path)
