"""Host-side step driver over the C-ABI, in the call order of
vpic_simulation::advance() (src/vpic/advance.cc:15-207) restricted to the
hot path: sort_p, clear_accumulator_array, advance_p per species,
reduce_accumulator_array, clear_jf / unload_accumulator_array /
synchronize_jf, advance_b(0.5), advance_e(1), advance_b(0.5),
load_interpolator_array.  Collisions, emitters and boundary_p's
inter-domain exchange are not part of the path.  The dump_* methods write
the reference's output formats (src/vpic/dump.cc) from the device arrays.

All arrays live in HBM (vpic_amd.device); every call is stream-ordered on
one HIP stream, and the only per-step host synchronisation is the sort's.
"""
import ctypes

import numpy as np
import torch

from . import _lib, device
from . import types as T


class Species:
    """species_t (src/species_advance/species_advance_aos.h:54-94) with
    HBM-resident particle, mover and partition arrays."""

    def __init__(self, name, q, m, np_, grid, max_nm=None, sort_interval=0):
        self.name = name
        self.q = float(q)
        self.m = float(m)
        self.np = int(np_)
        self.max_np = int(np_)
        self.p = device.particles(np_)
        # vpic.h:507-511: default max_nm = 2*np/25, at least 16*(MAX_PIPELINE+1)
        if max_nm is None:
            max_nm = max(2 * np_ // 25, 16 * 273)
        self.max_nm = int(max_nm)
        self.pm = device.movers(self.max_nm)
        self.nm = 0
        self.sort_interval = int(sort_interval)
        self.last_sorted = -1
        self.partition = device.DeviceArray((grid.nv + 1,), np.int32)
        self.dev_nm = device.DeviceArray((4,), np.int32)  # last push's movers
        # sort_p's out-of-place buffer (sort_p.cc:91 takes a fresh max_np
        # buffer per sort); allocated by the first sort, then p and aux
        # swap roles at every sort, so both always hold max_np records
        self.aux = None
        # voxel keys written by the push of the step before a sort
        # (allocated by the first such push) and that step
        self.keys = None
        self.keys_step = -1


class Simulation:
    """The per-step hot path on one GPU."""

    def __init__(self, grid, field_params=None, origin=(0.0, 0.0, 0.0)):
        self.g = grid
        self.origin = tuple(float(v) for v in origin)  # grid_t.x0, y0, z0
        self.hydro = None  # allocated by the first dump_hydro
        self.collisions = []  # define_collision_op, applied each step
        self.mp = field_params or T.vacuum_params()
        self.fields = device.fields(grid)
        self.fields.zero()
        self.interp = device.interpolators(grid)
        self.interp.zero()
        self.acc = device.accumulators(grid)
        self.acc.zero()
        self.species = []
        self.step = 0
        self.stream = None  # the default stream
        self.nm_log = []

    def add_species(self, sp):
        self.species.append(sp)
        return sp

    # -- the reference-named operations ------------------------------------
    def _G(self):
        return ctypes.byref(self.g)

    def sort_p(self, sp):
        """sort_p (sort_p.cc:154-165): stable by voxel, into the species'
        own aux buffer of max_np records, then the two buffers swap roles
        (the species' particle array stays max_np records long)."""
        if sp.aux is None:
            sp.aux = device.particles(max(sp.max_np, 1))
        if sp.np > sp.max_np or sp.aux.shape[0] < sp.np or \
                sp.p.shape[0] < sp.np:
            raise _lib.VpicError(
                "sort_p: species %s has %d particles, its buffers hold %d "
                "and %d" % (sp.name, sp.np, sp.p.shape[0], sp.aux.shape[0]))
        if sp.keys is not None and sp.keys_step == self.step - 1:
            # the keys the last push wrote (checked against the particles
            # by the sort, which falls back to reading them on a mismatch)
            _lib.call("vpic_hip_sort_p_keys", sp.p.ptr, sp.aux.ptr, sp.np,
                      sp.partition.ptr, sp.keys.ptr, self._G(), 0,
                      self.stream)
        else:
            _lib.call("vpic_hip_sort_p", sp.p.ptr, sp.aux.ptr, sp.np,
                      sp.partition.ptr, self._G(), 0, self.stream)
        sp.keys_step = -1
        sp.p, sp.aux = sp.aux, sp.p
        sp.last_sorted = self.step

    def takizuka_abe(self, spi, spj, cvar0, interval, seed=0):
        """takizuka_abe(name, spi, spj, rp, cvar0, interval)
        (src/collision/takizuka_abe.cc:50-75) added to the step's collision
        operators (define_collision_op): every interval steps, both species
        sorted at this step first (apply_takizuka_abe, takizuka_abe.cc:6-12),
        then vpic_hip_takizuka_abe with the step as the stream's counter.
        The reference's operators draw in turn from one advancing generator
        pool; here each operator gets its own Philox key, collision_key(seed,
        its index among the operators), so two operators applied at the
        same step never share draws."""
        if interval < 1:
            raise _lib.VpicError("takizuka_abe: interval must be >= 1")
        key = collision_key(seed, len(self.collisions))

        def apply(sim):
            if sim.step % interval:
                return
            for sp in (spi, spj):
                if sp.last_sorted != sim.step:
                    sim.sort_p(sp)
            _lib.call("vpic_hip_takizuka_abe", spi.p.ptr, spi.partition.ptr,
                      spi.np, spi.q, spi.m, spj.p.ptr, spj.partition.ptr,
                      spj.np, spj.q, spj.m, float(cvar0), int(interval),
                      key, int(sim.step), sim._G(), sim.stream)
        self.collisions.append(apply)
        return apply

    def _push_args(self, sp):
        """vpic_push_species_t of one species for this step: after a sort
        its partition drives the tiled (LDS) push, and when the next step
        starts with its sort_p (advance.cc:25-29) the push leaves it the
        voxel keys."""
        part = sp.partition.ptr if sp.last_sorted >= 0 else None
        keys = None
        if sp.sort_interval > 0 and (self.step + 1) % sp.sort_interval == 0:
            if sp.keys is None or sp.keys.shape[0] < sp.max_np:
                sp.keys = device.DeviceArray((max(sp.max_np, 1),), np.int32)
            keys = sp.keys.ptr
            sp.keys_step = self.step
        else:
            sp.keys_step = -1
        return T.PushSpecies(sp.p.ptr, sp.np, sp.q, sp.m, sp.pm.ptr,
                             sp.max_nm, sp.dev_nm.ptr, part, keys)

    def advance_p(self, sp):
        a = self._push_args(sp)
        _lib.call("vpic_hip_advance_p_keys", a.p, a.np, a.q, a.m, a.pm,
                  a.max_nm, a.dev_nm, self.acc.ptr, self.interp.ptr,
                  self._G(), a.partition, a.keys, self.stream)

    def advance_species(self):
        """advance_p of every species into the accumulator (advance.cc:
        56-62): species with a partition go two to a call, which pushes
        both through the same tiles in one launch."""
        pending = []
        for sp in self.species:
            if sp.last_sorted < 0 or sp.np == 0:
                self.advance_p(sp)
                continue
            pending.append(sp)
            if len(pending) == 2:
                arr = (T.PushSpecies * 2)(*[self._push_args(x)
                                            for x in pending])
                _lib.call("vpic_hip_advance_p_species", 2, arr,
                          self.acc.ptr, self.interp.ptr, self._G(),
                          self.stream)
                pending = []
        for sp in pending:
            self.advance_p(sp)

    def load_interpolator_array(self):
        _lib.call("vpic_hip_load_interpolator_array", self.interp.ptr,
                  self.fields.ptr, self._G(), self.stream)

    def clear_accumulator_array(self):
        _lib.call("vpic_hip_clear_accumulator_array", self.acc.ptr,
                  self._G(), self.stream)

    def reduce_accumulator_array(self):
        _lib.call("vpic_hip_reduce_accumulator_array", self.acc.ptr,
                  self._G(), self.stream)

    def unload_accumulator_array(self):
        # clear_jf + unload_accumulator_array (advance.cc:107-109) in one
        # pass over the field records
        _lib.call("vpic_hip_clear_jf_unload_accumulator_array",
                  self.fields.ptr, self.acc.ptr, self._G(), self.stream)
        _lib.call("vpic_hip_synchronize_jf", self.fields.ptr, self._G(),
                  self.stream)

    def advance_b(self, frac):
        _lib.call("vpic_hip_advance_b", self.fields.ptr, self._G(), frac,
                  self.stream)

    def advance_e(self, frac=1.0):
        _lib.call("vpic_hip_vacuum_advance_e", self.fields.ptr, self._G(),
                  ctypes.byref(self.mp), frac, self.stream)

    def energy_f(self):
        en = (ctypes.c_double * 6)()
        _lib.call("vpic_hip_energy_f", en, self.fields.ptr, self._G(),
                  ctypes.byref(self.mp), self.stream)
        return np.array(en[:])

    def energy_p(self, sp):
        en = ctypes.c_double()
        _lib.call("vpic_hip_energy_p", sp.p.ptr, sp.np, sp.q, sp.m,
                  self.interp.ptr, self._G(), ctypes.byref(en), self.stream)
        return en.value

    def rho(self):
        """rhof of every species: clear_rhof, accumulate_rho_p,
        synchronize_rho (initialize.cc:33-36, advance.cc:147-150)."""
        _lib.call("vpic_hip_clear_rhof", self.fields.ptr, self._G(),
                  self.stream)
        for sp in self.species:
            _lib.call("vpic_hip_accumulate_rho_p", self.fields.ptr, sp.p.ptr,
                      sp.np, sp.q, self._G(), self.stream)
        _lib.call("vpic_hip_synchronize_rho", self.fields.ptr, self._G(),
                  self.stream)

    def div_e_rms(self):
        """The charge-conservation diagnostic of advance.cc:146-160: rho,
        compute_div_e_err, compute_rms_div_e_err (FP64 reduction on the
        device)."""
        self.rho()
        _lib.call("vpic_hip_compute_div_e_err", self.fields.ptr, self._G(),
                  ctypes.byref(self.mp), self.stream)
        err = ctypes.c_double()
        _lib.call("vpic_hip_compute_rms_div_e_err", ctypes.byref(err),
                  self.fields.ptr, self._G(), self.stream)
        return err.value

    def initialize(self):
        """vpic_simulation::initialize after the user's set-up
        (initialize.cc:16-60): synchronize_tang_e_norm_b, one div B
        cleaning pass, compute_curl_b, the bound charge rhob from rho and
        the initial E, one div E cleaning pass, synchronize_tang_e_norm_b
        again, load_interpolator_array and uncenter_p of every species.
        Returns the diagnostics the reference logs."""
        G, S_ = self._G(), self.stream
        out = {}
        err = ctypes.c_double()
        _lib.call("vpic_hip_synchronize_tang_e_norm_b", ctypes.byref(err),
                  self.fields.ptr, G, S_)
        out["sync_error"] = err.value
        _lib.call("vpic_hip_compute_div_b_err", self.fields.ptr, G, S_)
        _lib.call("vpic_hip_compute_rms_div_b_err", ctypes.byref(err),
                  self.fields.ptr, G, S_)
        out["div_b_rms"] = err.value
        _lib.call("vpic_hip_clean_div_b", self.fields.ptr, G, S_)
        _lib.call("vpic_hip_compute_curl_b", self.fields.ptr, G,
                  ctypes.byref(self.mp), S_)
        self.rho()
        _lib.call("vpic_hip_compute_rhob", self.fields.ptr, G,
                  ctypes.byref(self.mp), S_)
        _lib.call("vpic_hip_compute_div_e_err", self.fields.ptr, G,
                  ctypes.byref(self.mp), S_)
        _lib.call("vpic_hip_compute_rms_div_e_err", ctypes.byref(err),
                  self.fields.ptr, G, S_)
        out["div_e_rms"] = err.value
        _lib.call("vpic_hip_clean_div_e", self.fields.ptr, G,
                  ctypes.byref(self.mp), S_)
        _lib.call("vpic_hip_synchronize_tang_e_norm_b", ctypes.byref(err),
                  self.fields.ptr, G, S_)
        out["sync_error_after"] = err.value
        self.load_interpolator_array()
        for sp in self.species:
            _lib.call("vpic_hip_center_p", sp.p.ptr, sp.np, sp.q, sp.m,
                      self.interp.ptr, G, 1, S_)
        return out

    def check_movers(self, sp=None):
        """(nm, movers lost to max_nm, bad voxels) of the last advance_p of
        sp (default: the last species; host read, synchronises)."""
        sp = sp if sp is not None else self.species[-1]
        torch.cuda.synchronize()
        v = sp.dev_nm.download()
        return int(v[0]), int(v[1]), int(v[2])

    def boundary_p(self):
        """boundary_p (boundary_p.cc:42-750) for this one domain: every
        species' movers (absorbing faces: charge to rhob, particle removed).
        Custom particle boundaries need host handlers, which this driver
        does not have: a hit on one is an error."""
        for sp in self.species:
            nm = self.check_movers(sp)[0]
            sp.nm = nm
            if nm == 0:
                continue
            sp.keys_step = -1  # particles are removed and moved below
            npc = ctypes.c_int32(sp.np)
            counts = (ctypes.c_int32 * 2)()
            _lib.call("vpic_hip_boundary_p", sp.p.ptr, ctypes.byref(npc),
                      sp.pm.ptr, nm, self.fields.ptr, sp.q, self._G(), None,
                      counts, self.stream)
            if counts[0] or counts[1]:
                raise _lib.VpicError(
                    "boundary_p: %d particles hit custom particle boundaries "
                    "and %d unknown ones (species %s); this driver handles "
                    "absorbing walls only" % (counts[0], counts[1], sp.name))
            sp.np = npc.value
            sp.nm = 0

    # -- output (src/vpic/dump.cc) -----------------------------------------
    def _info(self):
        return T.DumpInfo(self.step, self.origin[0], self.origin[1],
                          self.origin[2], 0, 1)

    def _fname(self, fbase, ftag):
        # dump.cc:192-193: fbase.step.rank with a tag, else fbase.rank
        return ("%s.%d.%d" % (fbase, self.step, 0) if ftag
                else "%s.%d" % (fbase, 0))

    def dump_energies(self, fname, append=1):
        """dump_energies (dump.cc:38-77): energy_f and energy_p of every
        species, one text line per call."""
        en_f = (ctypes.c_double * 6)(*self.energy_f())
        nsp = len(self.species)
        en_p = (ctypes.c_double * max(nsp, 1))(
            *[self.energy_p(sp) for sp in self.species])
        names = (ctypes.c_char_p * max(nsp, 1))(
            *[sp.name.encode() for sp in self.species])
        _lib.call("vpic_hip_write_energies", fname.encode(), int(append),
                  ctypes.byref(self._info()), self.g.dt, en_f, en_p, names,
                  nsp)

    def dump_fields(self, fbase, ftag=1):
        """dump_fields (dump.cc:181-219)."""
        f = self.fields.download()
        _lib.call("vpic_hip_write_fields", self._fname(fbase, ftag).encode(),
                  f.ctypes.data, self._G(), ctypes.byref(self._info()))

    def dump_hydro(self, sp, fbase, ftag=1):
        """dump_hydro (dump.cc:221-270): clear_hydro_array,
        accumulate_hydro_p, synchronize_hydro_array, then the file."""
        h = self.hydro_moments(sp)
        _lib.call("vpic_hip_write_hydro", self._fname(fbase, ftag).encode(),
                  h.ctypes.data, self._G(), ctypes.byref(self._info()),
                  sp_id(self, sp), np.float32(sp.q / sp.m))

    def hydro_moments(self, sp):
        """clear_hydro_array, accumulate_hydro_p, synchronize_hydro_array
        of species sp; returns the host copy of the hydro array."""
        if self.hydro is None:
            self.hydro = device.DeviceArray((self.g.nv,), T.HYDRO)
        part = sp.partition.ptr if sp.last_sorted >= 0 else None
        _lib.call("vpic_hip_clear_hydro_array", self.hydro.ptr, self._G(),
                  self.stream)
        _lib.call("vpic_hip_accumulate_hydro_p", self.hydro.ptr, sp.p.ptr,
                  sp.np, sp.q, sp.m, self.interp.ptr, self._G(), part,
                  self.stream)
        _lib.call("vpic_hip_synchronize_hydro_array", self.hydro.ptr,
                  self._G(), self.stream)
        return self.hydro.download()

    def global_header(self, base, field_params, species_params):
        """global_header (dump.cc:405-532): "<base>.vpc" for the strided
        dumps, field_params and one DumpParams per dumped species."""
        n = 1 + len(species_params)
        arr = (T.DumpParams * n)(field_params, *species_params)
        lo, hi = self.origin, [self.origin[0] + self.g.nx * self.g.dx,
                               self.origin[1] + self.g.ny * self.g.dy,
                               self.origin[2] + self.g.nz * self.g.dz]
        gi = T.GlobalInfo(lo[0], hi[0], lo[1], hi[1], lo[2], hi[2], 1, 1, 1,
                          0)
        _lib.call("vpic_hip_global_header", base.encode(), self._G(),
                  ctypes.byref(gi), arr, n)

    def field_dump(self, params, step=-1):
        """field_dump (dump.cc:533-706) of the field array."""
        f = self.fields.download()
        _lib.call("vpic_hip_field_dump", f.ctypes.data, self._G(),
                  ctypes.byref(self._info()), ctypes.byref(params), step)

    def hydro_dump(self, sp, params, step=-1):
        """hydro_dump (dump.cc:708-880): the species' moments, then the
        file."""
        h = self.hydro_moments(sp)
        _lib.call("vpic_hip_hydro_dump", h.ctypes.data, self._G(),
                  ctypes.byref(self._info()), sp_id(self, sp),
                  np.float32(sp.q / sp.m), ctypes.byref(params), step)

    def dump_particles(self, sp, fbase, ftag=1, chunk=1 << 24):
        """dump_particles (dump.cc:273-335): the particles time-centred
        with center_p (on a device copy, chunk by chunk; the species is
        unchanged)."""
        fname = self._fname(fbase, ftag)
        _lib.call("vpic_hip_write_particles", fname.encode(), None, 0, sp.np,
                  self._G(), ctypes.byref(self._info()), sp_id(self, sp),
                  np.float32(sp.q / sp.m))
        buf = device.particles(min(max(sp.np, 1), chunk))
        with open(fname, "ab") as fh:
            for s0 in range(0, sp.np, chunk):
                n = min(chunk, sp.np - s0)
                nb = n * T.PARTICLE.itemsize
                buf.t[:nb].copy_(sp.p.t[s0 * T.PARTICLE.itemsize:
                                        s0 * T.PARTICLE.itemsize + nb])
                _lib.call("vpic_hip_center_p", buf.ptr, n, sp.q, sp.m,
                          self.interp.ptr, self._G(), 0, self.stream)
                torch.cuda.synchronize()
                fh.write(buf.t[:nb].cpu().numpy().tobytes())

    def dump_grid(self, fbase):
        """dump_grid (dump.cc:138-179) of this one-rank domain: bc[27] (this
        rank, or the face's field boundary code), range {0, nv}, and the
        neighbor table."""
        bc = np.zeros(27, dtype=np.int32)
        for k, (x, y, z) in enumerate(((-1, 0, 0), (0, -1, 0), (0, 0, -1),
                                       (1, 0, 0), (0, 1, 0), (0, 0, 1))):
            if self.g.fbc[k] != T.BC_PERIODIC:
                bc[13 + x + 3 * y + 9 * z] = self.g.fbc[k]
        rng = np.array([0, self.g.nv], dtype=np.int64)
        nb = np.zeros(6 * self.g.nv, dtype=np.int64)
        _lib.call("vpic_hip_grid_neighbor_table", self._G(), nb.ctypes.data)
        _lib.call("vpic_hip_write_grid", ("%s.%d" % (fbase, 0)).encode(),
                  self._G(), ctypes.byref(self._info()), bc.ctypes.data,
                  rng.ctypes.data, nb.ctypes.data)

    # -- checkpoint / restart (checkpt.h:101-110) ----------------------------
    def checkpoint(self, path):
        """Write the state a restarted run continues from
        (vpic_amd.checkpoint)."""
        from . import checkpoint
        checkpoint.save(self, path)

    @staticmethod
    def restore(path):
        """The Simulation saved at `path` by checkpoint()."""
        from . import checkpoint
        return checkpoint.load(path)

    # -- one time step (advance.cc:15-207) ----------------------------------
    def advance(self, timer=None):
        t = timer or (lambda name: _NullCtx())
        with t("sort"):
            for sp in self.species:
                if sp.sort_interval > 0 and self.step % sp.sort_interval == 0:
                    self.sort_p(sp)
        # advance.cc:39-46: the collision operators, after the sorts
        for op in self.collisions:
            op(self)
        with t("push"):
            self.clear_accumulator_array()
            self.advance_species()
            # advance.cc:70-77: particles that hit a particle boundary
            if any(self.g.pbc[k] != T.BC_PERIODIC for k in range(6)):
                self.boundary_p()
            self.reduce_accumulator_array()
        with t("field"):
            self.unload_accumulator_array()
            self.advance_b(0.5)
            self.advance_e(1.0)
            self.advance_b(0.5)
        with t("interp"):
            self.load_interpolator_array()
        self.step += 1


def collision_key(seed, index):
    """Philox key of collision operator `index` for a user seed: the
    splitmix64 finaliser of seed * 2^16 + index, so different operators
    (and different seeds) get unrelated keys."""
    m = (1 << 64) - 1
    z = ((int(seed) << 16) + int(index) + 0x9E3779B97F4A7C15) & m
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & m
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & m
    return z ^ (z >> 31)


def sp_id(sim, sp):
    """species_t.id: the species' position in the simulation's list."""
    return sim.species.index(sp)


class _NullCtx:
    def __enter__(self):
        return self

    def __exit__(self, *a):
        return False


class PhaseTimer:
    """torch events (on the default stream, where the C-ABI calls run)
    around each phase; elapsed times are collected at the end."""

    def __init__(self):
        self.ev = {}

    def __call__(self, name):
        timer = self

        class _Ctx:
            def __enter__(self_):
                self_.a = torch.cuda.Event(enable_timing=True)
                self_.a.record()
                return self_

            def __exit__(self_, *exc):
                b = torch.cuda.Event(enable_timing=True)
                b.record()
                timer.ev.setdefault(name, []).append((self_.a, b))
                return False

        return _Ctx()

    def totals_ms(self):
        torch.cuda.synchronize()
        return {k: sum(a.elapsed_time(b) for a, b in v)
                for k, v in self.ev.items()}


def host_species(sim, name, q, m, p, sort_interval=0, max_np=None):
    """A species whose particles (a host particle array) are uploaded."""
    n = len(p)
    sp = Species(name, q, m, max(n, max_np or 0), sim.g,
                 sort_interval=sort_interval)
    sp.np = n
    if n:
        src = torch.from_numpy(np.ascontiguousarray(p).view(np.uint8))
        sp.p.t[:src.numel()].copy_(src)
    return sim.add_species(sp)


def uniform_plasma(sim, ppc, q, m, uth, seed, name, drift=(0.0, 0.0, 0.0),
                   w=1.0, sort_interval=0):
    """Species loaded as a uniform thermal plasma (device loader)."""
    g = sim.g
    np_ = ppc * g.nx * g.ny * g.nz
    sp = Species(name, q, m, np_, g, sort_interval=sort_interval)
    uth = np.broadcast_to(np.asarray(uth, dtype=np.float64), (3,))
    _lib.call("vpic_hip_load_uniform_plasma", sp.p.ptr, ctypes.byref(g), ppc,
              float(uth[0]), float(uth[1]), float(uth[2]), float(drift[0]),
              float(drift[1]), float(drift[2]), float(w), int(seed), None)
    return sim.add_species(sp)
