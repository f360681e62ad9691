"""The push entry point that allocates nothing (vpic_hip_advance_p_noalloc)
against the allocating stream-ordered push it shares its kernels with
(vpic_hip_advance_p_async without a partition), through the C-ABI.

Two calls in a row on the same GPU (the second with a smaller input, so the
fixed scratch holds the first call's counters and lists when the second
starts): particles, movers and {nm, n_ignored, n_bad} bit-exact, currents
within 1e-5 relative (only the order of the float atomics differs).  The
inputs have reflecting and absorbing walls, so movers are returned, and the
second call caps max_nm below the mover count (the rollback of the rest).
"""
import ctypes
import sys

import numpy as np
import pytest

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


def _walled_box(n):
    nx, ny, nz = n
    dt = 0.95 * T.courant_length(nx, ny, nz, nx, ny, nz)
    g = T.grid_periodic(0, 0, 0, nx, ny, nz, nx, ny, nz, dt)
    g.pbc[0] = g.pbc[3] = T.BC_REFLECT_PARTICLES
    g.pbc[2] = g.pbc[5] = T.BC_ABSORB_PARTICLES
    return g


def _inputs(orc, g, ppc, uth, seed):
    f = H.random_fields(g, seed, 0.2, 0.2)
    fi = H.interp_array(g)
    orc.load_interpolator(fi, f, g)
    return H.thermal_particles(g, ppc, uth, seed + 1), fi


def _push(dev, name, p, fi, g, max_nm):
    dp = dev.DeviceArray.from_host(p)
    dfi = dev.DeviceArray.from_host(fi)
    da = dev.DeviceArray.from_host(H.acc_array(g))
    dpm = dev.movers(max(max_nm, 1))
    dnm = dev.DeviceArray.from_host(np.full(3, -7, dtype=np.int32))
    args = [dp.ptr, len(p), -1.0, 1.0, dpm.ptr, max_nm, dnm.ptr, da.ptr,
            dfi.ptr, ctypes.byref(g)]
    if name == "vpic_hip_advance_p_async":
        args.append(None)  # no partition: the flat push
    _lib.call(name, *args, None)
    out = dnm.download()
    nm = int(out[0])
    return dp.download(), da.download(), out, dpm.download()[:nm].copy()


@pytest.mark.parametrize("shape,ppc,uth,seed,cap", [
    ((10, 12, 14), 24, 0.6, 11, None),   # 40,320 particles, all movers kept
    ((6, 5, 7), 9, 0.9, 12, 3),          # 1,890 particles, max_nm below nm
])
def test_noalloc_matches_allocating_push_twice(dev, orc, shape, ppc, uth,
                                               seed, cap):
    calls = []
    for k, (sh, pp, s) in enumerate(((shape, ppc, seed),
                                     ((4, 5, 6), 5, seed + 100))):
        g = _walled_box(sh)
        p, fi = _inputs(orc, g, pp, uth, s)
        calls.append((g, p, fi))
    assert len(calls[1][1]) < len(calls[0][1])
    for k, (g, p, fi) in enumerate(calls):
        max_nm = len(p) if (cap is None or k == 0) else cap
        say("advance_p_noalloc call %d: %d particles, grid %dx%dx%d, "
            "max_nm %d" % (k + 1, len(p), g.nx, g.ny, g.nz, max_nm))
        pa, aa, na, ma = _push(dev, "vpic_hip_advance_p_async", p, fi, g,
                               max_nm)
        pz, az, nz, mz = _push(dev, "vpic_hip_advance_p_noalloc", p, fi, g,
                               max_nm)
        assert na[0] > 0, "the walled box returns movers"
        assert list(nz) == list(na)
        assert H.same_bits(mz, ma)
        assert H.same_bits(pz, pa)
        assert H.acc_rel_err(az, aa) <= ACC_TOL


def test_noalloc_rejects_more_movers_than_its_scratch(dev, orc):
    g = _walled_box((4, 4, 4))
    p, fi = _inputs(orc, g, 2, 0.1, 3)
    say("advance_p_noalloc with max_nm beyond its scratch")
    with pytest.raises(_lib.VpicError, match="max_nm"):
        _push(dev, "vpic_hip_advance_p_noalloc", p, fi, g, 65537)
