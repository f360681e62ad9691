"""Host-side mirrors of the C-ABI data types in include/vpic_hip.h.

The layouts are the reference's (lanl/vpic legacy):
particle_t   src/species_advance/species_advance_aos.h:21-32   (32 B)
particle_mover_t                         species_advance_aos.h:37-40   (16 B)
interpolator_t src/sf_interface/sf_interface.h:62-74            (80 B)
accumulator_t  src/sf_interface/sf_interface.h:115-123          (48 B)
field_t        src/field_advance/field_advance.h:152-160        (80 B)
grid_t fields  src/grid/grid.h:73-131 (the subset the path reads)
"""
import ctypes

import numpy as np

PARTICLE = np.dtype([("dx", "<f4"), ("dy", "<f4"), ("dz", "<f4"), ("i", "<i4"),
                     ("ux", "<f4"), ("uy", "<f4"), ("uz", "<f4"), ("w", "<f4")])
MOVER = np.dtype([("dispx", "<f4"), ("dispy", "<f4"), ("dispz", "<f4"),
                  ("i", "<i4")])
INTERP_FIELDS = ["ex", "dexdy", "dexdz", "d2exdydz", "ey", "deydz", "deydx",
                 "d2eydzdx", "ez", "dezdx", "dezdy", "d2ezdxdy", "cbx",
                 "dcbxdx", "cby", "dcbydy", "cbz", "dcbzdz", "pad0", "pad1"]
INTERPOLATOR = np.dtype([(n, "<f4") for n in INTERP_FIELDS])
ACCUMULATOR = np.dtype([("jx", "<f4", (4,)), ("jy", "<f4", (4,)),
                        ("jz", "<f4", (4,))])
FIELD_FLOATS = ["ex", "ey", "ez", "div_e_err", "cbx", "cby", "cbz",
                "div_b_err", "tcax", "tcay", "tcaz", "rhob", "jfx", "jfy",
                "jfz", "rhof"]
FIELD = np.dtype([(n, "<f4") for n in FIELD_FLOATS] +
                 [(n, "<i2") for n in ("ematx", "ematy", "ematz", "nmat",
                                       "fmatx", "fmaty", "fmatz", "cmat")])

assert PARTICLE.itemsize == 32 and MOVER.itemsize == 16
assert INTERPOLATOR.itemsize == 80 and ACCUMULATOR.itemsize == 48
assert FIELD.itemsize == 80

# Boundary codes (src/grid/grid.h:14-38).
BC_PERIODIC = 0
BC_ANTI_SYMMETRIC_FIELDS = -1
BC_PEC_FIELDS = -1
BC_SYMMETRIC_FIELDS = -2
BC_PMC_FIELDS = -3
BC_ABSORB_FIELDS = -4
BC_REFLECT_PARTICLES = -1
BC_ABSORB_PARTICLES = -2


class Grid(ctypes.Structure):
    """vpic_grid_t."""
    _fields_ = [("dt", ctypes.c_float), ("cvac", ctypes.c_float),
                ("eps0", ctypes.c_float),
                ("dx", ctypes.c_float), ("dy", ctypes.c_float),
                ("dz", ctypes.c_float), ("dV", ctypes.c_float),
                ("rdx", ctypes.c_float), ("rdy", ctypes.c_float),
                ("rdz", ctypes.c_float), ("r8V", ctypes.c_float),
                ("nx", ctypes.c_int32), ("ny", ctypes.c_int32),
                ("nz", ctypes.c_int32),
                ("sx", ctypes.c_int32), ("sy", ctypes.c_int32),
                ("sz", ctypes.c_int32), ("nv", ctypes.c_int32),
                ("rangel", ctypes.c_int64), ("rangeh", ctypes.c_int64),
                ("neighbor", ctypes.c_void_p),
                ("pbc", ctypes.c_int32 * 6), ("fbc", ctypes.c_int32 * 6)]

    def copy(self):
        g = Grid()
        ctypes.memmove(ctypes.byref(g), ctypes.byref(self), ctypes.sizeof(Grid))
        return g

    def voxel(self, x, y, z):
        """VOXEL(x,y,z,nx,ny,nz), src/grid/grid.h:136."""
        return x + (self.nx + 2) * (y + (self.ny + 2) * z)


INJECTOR = np.dtype([("dx", "<f4"), ("dy", "<f4"), ("dz", "<f4"),
                     ("i", "<i4"), ("ux", "<f4"), ("uy", "<f4"), ("uz", "<f4"),
                     ("w", "<f4"), ("dispx", "<f4"), ("dispy", "<f4"),
                     ("dispz", "<f4"), ("sp_id", "<i4")])
BOUNDARY_HIT = np.dtype([("p", PARTICLE), ("m", MOVER), ("face", "<i4"),
                         ("code", "<i4"), ("pad", "<i4", (2,))])
BOUNDARY_UNKNOWN = 1

HYDRO_FIELDS = ("jx", "jy", "jz", "rho", "px", "py", "pz", "ke", "txx",
                "tyy", "tzz", "tyz", "tzx", "txy")
HYDRO = np.dtype([(n, np.float32) for n in HYDRO_FIELDS] +
                 [("pad0", np.float32), ("pad1", np.float32)])


class PushSpecies(ctypes.Structure):
    """vpic_push_species_t: one species of vpic_hip_advance_p_species."""
    _fields_ = [("p", ctypes.c_void_p), ("np", ctypes.c_int32),
                ("q", ctypes.c_float), ("m", ctypes.c_float),
                ("pm", ctypes.c_void_p), ("max_nm", ctypes.c_int32),
                ("dev_nm", ctypes.c_void_p), ("partition", ctypes.c_void_p),
                ("keys", ctypes.c_void_p)]


class DumpInfo(ctypes.Structure):
    """vpic_dump_info_t: step, the domain's low corner, rank / nproc."""
    _fields_ = [("step", ctypes.c_int64), ("x0", ctypes.c_float),
                ("y0", ctypes.c_float), ("z0", ctypes.c_float),
                ("rank", ctypes.c_int32), ("nproc", ctypes.c_int32)]


DUMP_BAND = 0
DUMP_BAND_INTERLEAVE = 1


class DumpParams(ctypes.Structure):
    """vpic_dump_params_t: DumpParameters (src/vpic/vpic.h:92-117) of one
    strided dump: directory, file base name, strides, format, and the
    output_vars bit field (bit i = word i of the record)."""
    _fields_ = [("base_dir", ctypes.c_char_p),
                ("base_file_name", ctypes.c_char_p),
                ("stride", ctypes.c_int32 * 3), ("format", ctypes.c_int32),
                ("output_vars", ctypes.c_uint32)]


def dump_params(base_dir, base_file_name, stride=(1, 1, 1), fmt=DUMP_BAND,
                output_vars=0xffffffff):
    """A DumpParams with the reference's defaults (all variables, unit
    strides, band format)."""
    return DumpParams(base_dir.encode(), base_file_name.encode(),
                      (ctypes.c_int32 * 3)(*stride), fmt,
                      output_vars & 0xffffffff)


class GlobalInfo(ctypes.Structure):
    """vpic_global_info_t: the global domain's extents and topology."""
    _fields_ = [(n, ctypes.c_float) for n in
                ("x0", "x1", "y0", "y1", "z0", "z1")] + \
               [(n, ctypes.c_int32) for n in ("px", "py", "pz", "rank")]


class FieldParams(ctypes.Structure):
    """vpic_field_params_t."""
    _fields_ = [(n, ctypes.c_float) for n in
                ("decayx", "drivex", "decayy", "drivey", "decayz", "drivez",
                 "rmux", "rmuy", "rmuz", "epsx", "epsy", "epsz", "damp",
                 "nonconductive")]


def grid_periodic(gx0, gy0, gz0, gx1, gy1, gz1, gnx, gny, gnz, dt, cvac=1.0,
                  eps0=1.0):
    """partition_periodic_box for one rank (src/grid/partition.cc:35-90)
    followed by size_grid's strides (src/grid/ops.cc:25-30)."""
    f32 = np.float32
    g = Grid()
    g.dt, g.cvac, g.eps0 = f32(dt), f32(cvac), f32(eps0)
    g.dx = f32((gx1 - gx0) / float(gnx))
    g.dy = f32((gy1 - gy0) / float(gny))
    g.dz = f32((gz1 - gz0) / float(gnz))
    g.dV = f32(((gx1 - gx0) / float(gnx)) * ((gy1 - gy0) / float(gny)) *
               ((gz1 - gz0) / float(gnz)))
    g.rdx = f32(float(gnx) / (gx1 - gx0))
    g.rdy = f32(float(gny) / (gy1 - gy0))
    g.rdz = f32(float(gnz) / (gz1 - gz0))
    g.r8V = f32((float(gnx) / (gx1 - gx0)) * (float(gny) / (gy1 - gy0)) *
                (float(gnz) / (gz1 - gz0)) * 0.125)
    g.nx, g.ny, g.nz = gnx, gny, gnz
    g.sx = 1
    g.sy = gnx + 2
    g.sz = (gny + 2) * g.sy
    g.nv = (gnz + 2) * g.sz
    g.rangel = 0
    g.rangeh = g.nv - 1
    g.neighbor = None
    for k in range(6):
        g.pbc[k] = BC_PERIODIC
        g.fbc[k] = BC_PERIODIC
    return g


def courant_length(lx, ly, lz, nx, ny, nz):
    """vpic_simulation::courant_length, src/vpic/vpic.h:658-665."""
    w1 = 0.0
    if nx > 1:
        w1 += (nx / lx) ** 2
    if ny > 1:
        w1 += (ny / ly) ** 2
    if nz > 1:
        w1 += (nz / lz) ** 2
    return float(np.sqrt(1.0 / w1))


def vacuum_params(damp=0.0):
    """create_sfa_params for eps = mu = 1, sigma = 0
    (src/field_advance/standard/sfa.cc:100-136)."""
    mp = FieldParams()
    mp.decayx = mp.decayy = mp.decayz = 1.0
    mp.drivex = mp.drivey = mp.drivez = 1.0
    mp.rmux = mp.rmuy = mp.rmuz = 1.0
    mp.epsx = mp.epsy = mp.epsz = 1.0
    mp.damp = damp
    mp.nonconductive = 1.0
    return mp


def push_coefficients(q, m, g):
    """qdt_2mc and cdt_d{x,y,z} in float, advance_p_pipeline.cc:279-283."""
    f32 = np.float32
    q, m = f32(q), f32(m)
    qdt_2mc = (q * f32(g.dt)) / (f32(2) * m * f32(g.cvac))
    c = f32(g.cvac) * f32(g.dt)
    return qdt_2mc, c * f32(g.rdx), c * f32(g.rdy), c * f32(g.rdz)
