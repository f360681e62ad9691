"""Checkpoint and restart of a Simulation's hot-path state.

The reference checkpoints its object graph (checkpt.h:101-110:
checkpt_objects / restore_objects over the registered species, grids,
field, interpolator and accumulator arrays).  Here the state of the path
lives in HBM, so a checkpoint is one file holding what a restarted run
needs to continue bit for bit:

  the step, the domain's corner, the grid (vpic_grid_t without its
  pointer), the field parameters, and per species its couplings, sizes,
  sort interval and last sort step; then, as raw sections, the field,
  interpolator and accumulator arrays and every species' particles and
  sort-time partition.

Layout: the 16-byte magic b"VPICHIP-CKPT-01\\n", an 8-byte little-endian
header length, the header as UTF-8 JSON (which lists every section's
offset and size), then the sections, each starting at a multiple of 4096
bytes.  Arrays are streamed between HBM and the file in pieces, so a
checkpoint of the 27-GB benchmark plasma needs no host copy of it whole.

Not saved: collision operators (closures; define them again after a
restore), the hydro array (recomputed at each hydro dump), the movers of
the last push (every push rewrites them), and the sort keys the push
leaves before a sort (the restarted sort reads the voxel indices from the
particles instead: the same order).
"""
import ctypes
import json

import numpy as np
import torch

from . import _lib, device
from . import types as T

MAGIC = b"VPICHIP-CKPT-01\n"
ALIGN = 4096
CHUNK = 1 << 28  # bytes streamed per piece


def _struct_dict(st):
    return {name: (list(getattr(st, name)) if hasattr(getattr(st, name),
                                                      "__len__")
                   else getattr(st, name))
            for name, _ in st._fields_}


def _struct_from(cls, d):
    st = cls()
    for name, ctype in cls._fields_:
        v = d[name]
        if isinstance(v, list):
            arr = getattr(st, name)
            for k, x in enumerate(v):
                arr[k] = x
        else:
            setattr(st, name, v)
    return st


def _sections(sim):
    """(name, device array, bytes to save) of every array section."""
    out = [("fields", sim.fields, sim.fields.nbytes),
           ("interpolators", sim.interp, sim.interp.nbytes),
           ("accumulators", sim.acc, sim.acc.nbytes)]
    for k, sp in enumerate(sim.species):
        out.append(("species%d.particles" % k, sp.p,
                    sp.np * T.PARTICLE.itemsize))
        out.append(("species%d.partition" % k, sp.partition,
                    sp.partition.nbytes))
    return out


def save(sim, path):
    """Write sim's state to `path` (see the module docstring)."""
    if sim.g.neighbor:
        raise _lib.VpicError("checkpoint: grids with an explicit neighbor "
                             "table are not saved")
    torch.cuda.synchronize()
    g = _struct_dict(sim.g)
    g.pop("neighbor")
    secs = _sections(sim)
    header = {
        "format": 1,
        "step": sim.step,
        "origin": list(sim.origin),
        "grid": g,
        "field_params": _struct_dict(sim.mp),
        "species": [{"name": sp.name, "q": sp.q, "m": sp.m, "np": sp.np,
                     "max_np": sp.max_np, "max_nm": sp.max_nm,
                     "sort_interval": sp.sort_interval,
                     "last_sorted": sp.last_sorted}
                    for sp in sim.species],
        "sections": [],
    }
    # offsets once the header's own size is known (it holds them)
    for _ in range(2):
        blob = json.dumps(header).encode()
        off = _round(len(MAGIC) + 8 + len(blob))
        header["sections"] = []
        for name, _, nbytes in secs:
            header["sections"].append({"name": name, "offset": off,
                                       "nbytes": nbytes})
            off = _round(off + nbytes)
    blob = json.dumps(header).encode()
    assert _round(len(MAGIC) + 8 + len(blob)) == \
        header["sections"][0]["offset"]
    with open(path, "wb") as fh:
        fh.write(MAGIC)
        fh.write(np.uint64(len(blob)).tobytes())
        fh.write(blob)
        for (name, arr, nbytes), sec in zip(secs, header["sections"]):
            fh.seek(sec["offset"])
            for a in range(0, nbytes, CHUNK):
                b = min(a + CHUNK, nbytes)
                fh.write(arr.t[a:b].cpu().numpy().tobytes())
        fh.truncate(_round(fh.tell()))


def load(path):
    """A Simulation continued from the checkpoint at `path`."""
    from . import sim as S
    with open(path, "rb") as fh:
        if fh.read(len(MAGIC)) != MAGIC:
            raise _lib.VpicError("checkpoint: %s is not a checkpoint of "
                                 "this format" % path)
        n = int(np.frombuffer(fh.read(8), dtype=np.uint64)[0])
        header = json.loads(fh.read(n).decode())
        gd = dict(header["grid"])
        gd["neighbor"] = None
        g = _struct_from(T.Grid, gd)
        mp = _struct_from(T.FieldParams, header["field_params"])
        sim = S.Simulation(g, field_params=mp,
                           origin=tuple(header["origin"]))
        sim.step = int(header["step"])
        for d in header["species"]:
            sp = S.Species(d["name"], d["q"], d["m"], d["max_np"], sim.g,
                           max_nm=d["max_nm"],
                           sort_interval=d["sort_interval"])
            sp.np = int(d["np"])
            sp.last_sorted = int(d["last_sorted"])
            sim.add_species(sp)
        secs = {s["name"]: s for s in header["sections"]}
        for name, arr, nbytes in _sections(sim):
            sec = secs[name]
            if sec["nbytes"] != nbytes:
                raise _lib.VpicError("checkpoint: section %s holds %d "
                                     "bytes, %d expected" % (
                                         name, sec["nbytes"], nbytes))
            fh.seek(sec["offset"])
            for a in range(0, nbytes, CHUNK):
                b = min(a + CHUNK, nbytes)
                raw = np.frombuffer(fh.read(b - a), dtype=np.uint8)
                if raw.size != b - a:
                    raise _lib.VpicError("checkpoint: %s is truncated" % path)
                arr.t[a:b].copy_(torch.from_numpy(raw.copy()))
    torch.cuda.synchronize()
    return sim


def _round(n):
    return (n + ALIGN - 1) // ALIGN * ALIGN
