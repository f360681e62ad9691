"""The batched ensemble (vpic_hip_ensemble_*, vpic_amd/ensemble.py): many
independent periodic boxes in one persistent, graph-captured step.

Each box must be exactly a single-grid run of its own: after every step,
for every box, from the device's state before the step,
  * the particles equal the oracle's advance_p of that box alone bit for
    bit (same interpolators in);
  * the box's accumulators equal the oracle's within 1e-5 (only the
    summation order of the deposits differs);
  * the fields equal the oracle's unload / synchronize_jf / advance_b /
    vacuum_advance_e / advance_b of the box from the device's accumulators,
    and the interpolators its load_interpolator_array, bit for bit.
Run with the default tile of the tiled push (one tile per small box), with
the flat push, and with 2x2x1 tiles (several tiles per box, so deposits
cross tile windows and go through the out-of-window records and the
merge).  Steps alternate between graph replays (several in a row, with host
work between) and phase-by-phase launches; every other step starts with the
batched sort_p of every box (the oracle sorts its copy too; the batched sort
keeps no order within a voxel, so those steps compare the particles as a
set, each pushed bit for bit, and check the voxel order).  The tuner keeps one of its
candidate shapes and leaves the state as it found it.
"""
import numpy as np
import pytest

from vpic_amd import types as T

import helpers as H


def test_beam_config_from_seed():
    from vpic_amd.ensemble import SIDES, beam_config, beam_grid
    a = beam_config(256, 7)
    assert a == beam_config(256, 7) and a != beam_config(256, 8)
    assert all(c["nx"] in SIDES and c["ny"] in SIDES for c in a)
    assert all(16 <= c["ppc"] <= 256 for c in a)
    assert all(0.05 <= c["drift"] <= 0.3 for c in a)
    g = beam_grid(64, 32)
    assert (g.nx, g.ny, g.nz) == (64, 32, 1)
    assert abs(float(g.dt) - 0.95 * 0.5 / np.sqrt(2.0)) < 1e-6


def _small_boxes(seed):
    from vpic_amd.ensemble import beam_grid
    rng = np.random.default_rng(seed)
    out = []
    for b in range(5):
        nx, ny = (int(v) for v in rng.choice((3, 4, 6, 8), 2))
        ppc = int(rng.integers(4, 13))
        g = beam_grid(nx, ny)
        p = H.thermal_particles(g, ppc, 0.05, 100 + b, w=0.125 / ppc)
        half = len(p) // 2
        d = float(rng.uniform(0.05, 0.3))
        p["ux"][:half] += np.float32(d)
        p["ux"][half:] -= np.float32(d)
        f = H.random_fields(g, 200 + b, 0.02, 0.02)
        out.append((g, p, f))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("tile", [None, (0, 0, 0), (2, 2, 1), (3, 2, 1, 3)],
                         ids=["tile-default", "flat", "tile-2x2x1",
                              "tile-3x2x1-halo3"])
def test_ensemble_boxes_match_single_box_oracle(orc, tile):
    import torch
    from oracle.oracle import host_grid
    from vpic_amd import _lib
    from vpic_amd.ensemble import Ensemble
    assert torch.cuda.is_available()
    _lib.call("vpic_hip_init", 0)
    boxes = _small_boxes(3)
    e = Ensemble([g for g, _, _ in boxes], [len(p) for _, p, _ in boxes])
    for b, (g, p, f) in enumerate(boxes):
        e.set_particles(b, p)
        e.set_fields(b, f)
    e.prepare()
    if tile is not None:
        e.tile(*tile)
    e.sort_interval(2)  # sorts before steps 0, 2 and 4
    mp = T.vacuum_params()
    for step, graph in enumerate((True, True, True, False, True)):
        before = [(e.particles(b), e.fields(b), e.interpolators(b))
                  for b in range(len(boxes))]
        e.step(1, graph=graph)
        sorted_step = step % 2 == 0
        for b, (g, _, _) in enumerate(boxes):
            p0, f0, fi0 = before[b]
            hg = host_grid(g)
            p = p0.copy()
            if sorted_step:
                orc.sort_p(p, g)
            pre_voxels = p["i"].copy()
            acc = H.acc_array(g)
            nm, ign, _ = orc.advance_p(p, -1.0, 1.0, acc, fi0, hg)
            assert nm == 0 and ign == 0
            got_p = e.particles(b)
            if sorted_step:
                # the batched sort keeps no order within a voxel: the same
                # particles, each pushed bit for bit, in voxel order
                where = {bytes(r): k for k, r in
                         enumerate(p.view(np.uint8).reshape(len(p), 32))}
                assert len(where) == len(p)
                idx = [where.get(bytes(r), -1) for r in
                       got_p.view(np.uint8).reshape(len(got_p), 32)]
                assert min(idx) >= 0 and len(set(idx)) == len(p), (step, b)
                assert np.all(np.diff(pre_voxels[idx]) >= 0), (step, b)
            else:
                assert H.same_bits(got_p, p), (step, b)
            got_a = e.accumulators(b)
            assert H.acc_rel_err(got_a, acc) <= 1e-5, (step, b)
            f = f0.copy()
            orc.clear_jf(f, g)
            orc.unload_accumulator(f, got_a, g)
            orc.synchronize_jf(f, g)
            orc.advance_b(f, g, 0.5)
            orc.vacuum_advance_e(f, g, mp)
            orc.advance_b(f, g, 0.5)
            got_f = e.fields(b)
            for c in ("ex", "ey", "ez", "cbx", "cby", "cbz", "jfx", "jfy",
                      "jfz"):
                assert H.same_bits(got_f[c], f[c]), (step, b, c)
            fi = H.interp_array(g)
            orc.load_interpolator(fi, f, g)
            # interior voxels (the reference loads 1..n)
            got_fi = e.interpolators(b)
            x, y, z = H.voxel_coords(g, g.nx * g.ny * g.nz)
            v = (x + (g.nx + 2) * (y + (g.ny + 2) * z)).astype(np.int64)
            assert H.same_bits(got_fi[v], fi[v]), (step, b)
        print("step", step, "graph" if graph else "eager", "ok", flush=True)
        # host work between replays
        _ = [e.particles(b)[:4] for b in range(len(boxes))]
    e.close()


@pytest.mark.gpu
def test_ensemble_tuner_keeps_a_candidate_and_the_state():
    import torch
    from vpic_amd import _lib
    from vpic_amd.ensemble import beam_config, beam_ensemble
    assert torch.cuda.is_available()
    _lib.call("vpic_hip_init", 0)
    cfg = beam_config(12, 5, sides=(16, 32), ppc=(8, 32))
    e = beam_ensemble(cfg, 5)
    p0 = [e.particles(b) for b in range(len(cfg))]
    shape, ms = e.tune(2)
    print("tuned shape", shape, "%.3f ms per step" % ms, flush=True)
    assert ms > 0
    for b in range(len(cfg)):
        assert H.same_bits(e.particles(b), p0[b])
    e.step(3)
    e.close()
