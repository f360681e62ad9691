// dropin.cc -- the reference-named functions (include/vpic_dropin.h): the
// signatures of lanl/vpic's particle and field kernels over its own structs,
// forwarding to the vpic_hip_* C-ABI.  Host code only: residency of the
// caller's arrays in HBM, the grid_t -> vpic_grid_t translation and the
// reference's error convention.
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <vector>

#include "../../include/vpic_dropin.h"
#include "../../include/vpic_hip.h"
#include "../../include/vpic_dropin_layout.h"

// Error recording of the C-ABI (api.hip).
int vpic_set_error(int code, const char *fmt, ...);

static_assert(sizeof(particle_t) == sizeof(vpic_particle_t), "particle_t");
static_assert(sizeof(particle_mover_t) == sizeof(vpic_particle_mover_t),
              "particle_mover_t");
static_assert(sizeof(interpolator_t) == sizeof(vpic_interpolator_t),
              "interpolator_t");
static_assert(sizeof(accumulator_t) == sizeof(vpic_accumulator_t),
              "accumulator_t");
static_assert(sizeof(field_t) == sizeof(vpic_field_t), "field_t");
static_assert(sizeof(hydro_t) == sizeof(vpic_hydro_t), "hydro_t");
static_assert(sizeof(particle_injector_t) == sizeof(vpic_particle_injector_t),
              "particle_injector_t");

namespace {

// Device copy of one host array.
struct Mirror {
  void *dev = nullptr;
  size_t cap = 0;          // bytes allocated on the device
  size_t live = 0;         // bytes the device last wrote
  bool dev_valid = false;  // device holds the array's current data
  bool host_stale = false; // device wrote data the host has not seen
};

// Per-species device state that has no host array of its own.
struct SpeciesState {
  int32_t *dpart = nullptr;  // partition of the last sort_p, nv+1 entries
  int part_nv = 0;
  bool sorted = false;
};

struct State {
  std::map<const void *, Mirror> mirrors;
  std::map<const species_t *, SpeciesState> species;
  bool resident = false;
  bool inited = false;
  vpic_hip_error_handler_t handler = nullptr;
  void *aux = nullptr;  // sort_p scratch
  size_t aux_cap = 0;
  void *one = nullptr;  // move_p: {particle, mover, ret} staging
  void *stage = nullptr;  // boundary_p: hits / injectors
  size_t stage_cap = 0;
};

State &st() {
  static State s;
  return s;
}

// The reference's ERROR (util_base.h:267-273): message and exit(1), unless
// a handler is installed.  Returns false so callers can `return fail(..)`.
bool fail(const char *fn, const char *fmt, ...) {
  char msg[600];
  int k = snprintf(msg, sizeof msg, "%s: ", fn);
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(msg + k, sizeof msg - k, fmt, ap);
  va_end(ap);
  if (st().handler) {
    st().handler(msg);
    return false;
  }
  fprintf(stderr, "Error at libvpichip(%s):\n\t%s\n", fn, msg);
  fflush(stderr);
  exit(1);
}

void warn(const char *fn, const char *fmt, ...) {
  char msg[600];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(msg, sizeof msg, fmt, ap);
  va_end(ap);
  fprintf(stderr, "Warning at libvpichip(%s):\n\t%s\n", fn, msg);
}

#define HIPOK(fn, call)                                                  \
  do {                                                                   \
    hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess)                                                \
      return fail(fn, "%s: %s", #call, hipGetErrorString(e_));          \
  } while (0)

// HIPOK for functions returning void.
#define HIPOK_V(fn, call)                                                \
  do {                                                                   \
    hipError_t e_ = (call);                                              \
    if (e_ != hipSuccess) {                                              \
      fail(fn, "%s: %s", #call, hipGetErrorString(e_));                  \
      return;                                                            \
    }                                                                    \
  } while (0)

#define ABI(fn, call)                                                    \
  do {                                                                   \
    if ((call) != 0) return fail(fn, "%s", vpic_hip_last_error());       \
  } while (0)

bool ensure_init(const char *fn) {
  if (st().inited) return true;
  int dev = 0;
  HIPOK(fn, hipGetDevice(&dev));
  ABI(fn, vpic_hip_init(dev));
  HIPOK(fn, hipMalloc(&st().one, 256));
  st().inited = true;
  return true;
}

// Mirror of host array h with at least cap bytes on the device.
bool mirror(const char *fn, const void *h, size_t cap, Mirror **out) {
  Mirror &m = st().mirrors[h];
  if (m.cap < cap || !m.dev) {
    if (m.dev) HIPOK(fn, hipFree(m.dev));
    m = Mirror();
    HIPOK(fn, hipMalloc(&m.dev, cap ? cap : 16));
    m.cap = cap;
  }
  *out = &m;
  return true;
}

// Device copy of h for reading: the first `live` bytes are current.
template <class T>
bool dev_in(const char *fn, const void *h, size_t cap, size_t live, T **d) {
  Mirror *m;
  if (!mirror(fn, h, cap, &m)) return false;
  if (!st().resident || !m->dev_valid) {
    if (live) HIPOK(fn, hipMemcpy(m->dev, h, live, hipMemcpyHostToDevice));
    m->dev_valid = true;
    m->host_stale = false;
  }
  *d = (T *)m->dev;
  return true;
}

// Device copy of h that the call only writes.
template <class T>
bool dev_out(const char *fn, const void *h, size_t cap, T **d) {
  Mirror *m;
  if (!mirror(fn, h, cap, &m)) return false;
  *d = (T *)m->dev;
  return true;
}

// The device wrote the first `live` bytes of h's copy.
bool dev_wrote(const char *fn, const void *h, size_t live) {
  Mirror &m = st().mirrors[h];
  m.dev_valid = true;
  m.live = live;
  if (st().resident) {
    m.host_stale = true;
  } else {
    if (live)
      HIPOK(fn, hipMemcpy((void *)h, m.dev, live, hipMemcpyDeviceToHost));
    m.host_stale = false;
  }
  return true;
}

bool download(const char *fn, const void *h, size_t bytes) {
  auto it = st().mirrors.find(h);
  if (it == st().mirrors.end() || !it->second.dev_valid) return true;
  if (bytes > it->second.cap) bytes = it->second.cap;
  if (bytes)
    HIPOK(fn, hipMemcpy((void *)h, it->second.dev, bytes,
                        hipMemcpyDeviceToHost));
  it->second.host_stale = false;
  return true;
}

bool upload(const char *fn, const void *h, size_t cap, size_t bytes) {
  Mirror *m;
  if (!mirror(fn, h, cap, &m)) return false;
  if (bytes) HIPOK(fn, hipMemcpy(m->dev, h, bytes, hipMemcpyHostToDevice));
  m->dev_valid = true;
  m->host_stale = false;
  return true;
}

// grid_t (grid.h:73-131) -> vpic_grid_t.  The neighbor table is read like
// any other input (once per table in resident mode).  Field boundary codes come from bc[] (a rank number >= 0
// is a communicated face; on one device it must be this domain's own wrap).
bool make_grid(const char *fn, const grid_t *g, vpic_grid_t *G) {
  if (!g) return fail(fn, "Bad args (grid is NULL)");
  if (!g->neighbor) return fail(fn, "grid has no neighbor table");
  memset(G, 0, sizeof *G);
  G->dt = g->dt;
  G->cvac = g->cvac;
  G->eps0 = g->eps0;
  G->dx = g->dx;
  G->dy = g->dy;
  G->dz = g->dz;
  G->dV = g->dV;
  G->rdx = g->rdx;
  G->rdy = g->rdy;
  G->rdz = g->rdz;
  G->r8V = g->r8V;
  G->nx = g->nx;
  G->ny = g->ny;
  G->nz = g->nz;
  G->sx = g->sx;
  G->sy = g->sy;
  G->sz = g->sz;
  G->nv = g->nv;
  G->rangel = g->rangel;
  G->rangeh = g->rangeh;
  if (g->nv <= 0) return fail(fn, "grid has no voxels");
  const size_t nb_bytes = (size_t)6 * g->nv * sizeof(int64_t);
  Mirror *m;
  if (!mirror(fn, g->neighbor, nb_bytes, &m)) return false;
  if (!st().resident || !m->dev_valid) {
    HIPOK(fn, hipMemcpy(m->dev, g->neighbor, nb_bytes,
                        hipMemcpyHostToDevice));
    m->dev_valid = true;
  }
  G->neighbor = (const int64_t *)m->dev;
  static const int face[6][3] = {{-1, 0, 0}, {0, -1, 0}, {0, 0, -1},
                                 {1, 0, 0},  {0, 1, 0},  {0, 0, 1}};
  const int nx = g->nx, ny = g->ny;
  for (int k = 0; k < 6; k++) {
    // an edge voxel on face k, and the table's neighbour across it
    const int x = face[k][0] > 0 ? g->nx : 1;
    const int y = face[k][1] > 0 ? g->ny : 1;
    const int z = face[k][2] > 0 ? g->nz : 1;
    const int64_t v = x + (int64_t)(nx + 2) * (y + (int64_t)(ny + 2) * z);
    const int64_t nb = g->neighbor[6 * v + k];
    const bool local = nb >= g->rangel && nb <= g->rangeh;
    G->pbc[k] = nb < 0 ? (int)nb : VPIC_BC_PERIODIC;
    const int b = g->bc[BOUNDARY(face[k][0], face[k][1], face[k][2])];
    if (b >= 0) {
      if (!local)
        return fail(fn, "face %d is shared with rank %d: domains split over "
                        "several ranks are not supported", k, b);
      G->fbc[k] = VPIC_BC_PERIODIC;
    } else {
      G->fbc[k] = b;
    }
  }
  return true;
}

bool field_params(const char *fn, const field_array_t *fa,
                  vpic_field_params_t *mp) {
  const sfa_params_t *sp = (const sfa_params_t *)fa->params;
  if (!sp || !sp->mc) return fail(fn, "field array has no sfa_params_t");
  const material_coefficient_t &c = sp->mc[0];
  mp->decayx = c.decayx;
  mp->drivex = c.drivex;
  mp->decayy = c.decayy;
  mp->drivey = c.drivey;
  mp->decayz = c.decayz;
  mp->drivez = c.drivez;
  mp->rmux = c.rmux;
  mp->rmuy = c.rmuy;
  mp->rmuz = c.rmuz;
  mp->epsx = c.epsx;
  mp->epsy = c.epsy;
  mp->epsz = c.epsz;
  mp->damp = sp->damp;
  mp->nonconductive = c.nonconductive;
  return true;
}

size_t pbytes(int n) { return (size_t)(n > 0 ? n : 0) * sizeof(particle_t); }
size_t abytes(const grid_t *g) { return (size_t)g->nv * sizeof(accumulator_t); }
size_t ibytes(const grid_t *g) { return (size_t)g->nv * sizeof(interpolator_t); }
size_t fbytes(const grid_t *g) { return (size_t)g->nv * sizeof(field_t); }
size_t hbytes(const grid_t *g) { return (size_t)g->nv * sizeof(hydro_t); }

bool species_ok(const char *fn, const species_t *sp) {
  if (!sp || !sp->g || sp->np < 0 || sp->np > sp->max_np ||
      (sp->np && !sp->p))
    return fail(fn, "Bad args (species)");
  return true;
}

// Field kernel helper: f in/out, grid.
bool field_call(const char *fn, field_array_t *fa, vpic_field_t **f,
                vpic_grid_t *G) {
  if (!fa || !fa->f || !fa->g) return fail(fn, "Bad args");
  if (!ensure_init(fn) || !make_grid(fn, fa->g, G)) return false;
  return dev_in(fn, fa->f, fbytes(fa->g), fbytes(fa->g), f);
}

}  // namespace

// ---------------------------------------------------------------------------
// species kernels
// ---------------------------------------------------------------------------
extern "C" void advance_p(species_t *sp, accumulator_array_t *aa,
                          const interpolator_array_t *ia) {
  const char *fn = "advance_p";
  // advance_p_pipeline.cc:263-270
  if (!sp || !aa || !ia || sp->g != aa->g || sp->g != ia->g || !aa->a ||
      !ia->i) {
    fail(fn, "Bad args");
    return;
  }
  if (!species_ok(fn, sp) || !ensure_init(fn)) return;
  const grid_t *g = sp->g;
  vpic_grid_t G;
  if (!make_grid(fn, g, &G)) return;
  vpic_particle_t *p;
  vpic_accumulator_t *a;
  vpic_interpolator_t *f;
  vpic_particle_mover_t *pm;
  const int max_nm = sp->max_nm > 0 ? sp->max_nm : 0;
  if (!dev_in(fn, sp->p, pbytes(sp->max_np), pbytes(sp->np), &p) ||
      !dev_in(fn, aa->a, abytes(g), abytes(g), &a) ||
      !dev_in(fn, ia->i, ibytes(g), ibytes(g), &f) ||
      !dev_out(fn, sp->pm, (size_t)max_nm * sizeof(particle_mover_t), &pm))
    return;
  const SpeciesState &ss = st().species[sp];
  const int32_t *part =
      (ss.sorted && ss.part_nv == g->nv) ? ss.dpart : nullptr;
  int32_t nm = 0, ignored = 0;
  if (vpic_hip_advance_p(p, sp->np, sp->q, sp->m, pm, max_nm, &nm, &ignored,
                         a, f, &G, part, nullptr) != 0) {
    fail(fn, "%s", vpic_hip_last_error());
    return;
  }
  // advance_p_pipeline.cc:314-329
  if (ignored)
    warn(fn, "Pipelines ran out of storage for %i movers", ignored);
  sp->nm = nm;
  if (!dev_wrote(fn, sp->p, pbytes(sp->np)) ||
      !dev_wrote(fn, aa->a, abytes(g)) ||
      !dev_wrote(fn, sp->pm, (size_t)nm * sizeof(particle_mover_t)))
    return;
}

extern "C" void sort_p(species_t *sp) {
  const char *fn = "sort_p";
  if (!sp) {
    fail(fn, "Bad args.");
    return;
  }
  if (!species_ok(fn, sp) || !ensure_init(fn)) return;
  const grid_t *g = sp->g;
  vpic_grid_t G;
  if (!make_grid(fn, g, &G)) return;
  // sort_p_pipeline.cc:319: the step of the last sort
  sp->last_sorted = g->step;
  SpeciesState &ss = st().species[sp];
  if (!ss.dpart || ss.part_nv != g->nv) {
    if (ss.dpart && hipFree(ss.dpart) != hipSuccess) {
      fail(fn, "hipFree failed");
      return;
    }
    ss.dpart = nullptr;
    if (hipMalloc(&ss.dpart, ((size_t)g->nv + 1) * sizeof(int32_t)) !=
        hipSuccess) {
      fail(fn, "out of device memory for the partition");
      return;
    }
    ss.part_nv = g->nv;
  }
  vpic_particle_t *p;
  if (!dev_in(fn, sp->p, pbytes(sp->max_np), pbytes(sp->np), &p)) return;
  const size_t need = pbytes(sp->np > 0 ? sp->np : 1);
  if (st().aux_cap < need) {
    if (st().aux) (void)hipFree(st().aux);
    st().aux = nullptr;
    st().aux_cap = 0;
    if (hipMalloc(&st().aux, need) != hipSuccess) {
      fail(fn, "out of device memory for the sort scratch");
      return;
    }
    st().aux_cap = need;
  }
  if (vpic_hip_sort_p(p, (vpic_particle_t *)st().aux, sp->np, ss.dpart, &G, 1,
                      nullptr) != 0) {
    fail(fn, "%s", vpic_hip_last_error());
    return;
  }
  ss.sorted = true;
  if (!dev_wrote(fn, sp->p, pbytes(sp->np))) return;
  if (sp->partition && !st().resident)
    if (hipMemcpy(sp->partition, ss.dpart,
                  ((size_t)g->nv + 1) * sizeof(int32_t),
                  hipMemcpyDeviceToHost) != hipSuccess)
      fail(fn, "partition download failed");
}

extern "C" int move_p(particle_t *p0, particle_mover_t *m, accumulator_t *a0,
                      const grid_t *g, const float qsp) {
  const char *fn = "move_p";
  if (!p0 || !m || !a0 || !g || m->i < 0) {
    fail(fn, "Bad args");
    return 0;
  }
  if (!ensure_init(fn)) return 0;
  vpic_grid_t G;
  if (!make_grid(fn, g, &G)) return 0;
  State &S = st();
  auto it = S.mirrors.find(p0);
  const bool on_device =
      S.resident && it != S.mirrors.end() && it->second.dev_valid &&
      (size_t)(m->i + 1) * sizeof(particle_t) <= it->second.cap;
  // staging: [0,32) one particle, [32,48) the mover, [64,68) the flag
  char *one = (char *)S.one;
  vpic_particle_t *dp;
  vpic_particle_mover_t mv;
  memcpy(&mv, m, sizeof mv);
  if (on_device) {
    dp = (vpic_particle_t *)it->second.dev;
  } else {
    dp = (vpic_particle_t *)one;
    if (hipMemcpy(one, p0 + m->i, sizeof(particle_t),
                  hipMemcpyHostToDevice) != hipSuccess) {
      fail(fn, "particle upload failed");
      return 0;
    }
    mv.i = 0;
  }
  if (hipMemcpy(one + 32, &mv, sizeof mv, hipMemcpyHostToDevice) !=
      hipSuccess) {
    fail(fn, "mover upload failed");
    return 0;
  }
  vpic_accumulator_t *a;
  if (!dev_in(fn, a0, abytes(g), abytes(g), &a)) return 0;
  if (vpic_hip_move_p(dp, (vpic_particle_mover_t *)(one + 32), 1, a, &G, qsp,
                      (int32_t *)(one + 64), nullptr) != 0) {
    fail(fn, "%s", vpic_hip_last_error());
    return 0;
  }
  int32_t ret = 0;
  if (hipMemcpy(&mv, one + 32, sizeof mv, hipMemcpyDeviceToHost) !=
          hipSuccess ||
      hipMemcpy(&ret, one + 64, sizeof ret, hipMemcpyDeviceToHost) !=
          hipSuccess) {
    fail(fn, "download failed");
    return 0;
  }
  m->dispx = mv.dispx;
  m->dispy = mv.dispy;
  m->dispz = mv.dispz;
  if (on_device) {
    it->second.host_stale = true;
  } else if (hipMemcpy(p0 + m->i, one, sizeof(particle_t),
                       hipMemcpyDeviceToHost) != hipSuccess) {
    fail(fn, "particle download failed");
    return 0;
  }
  if (!dev_wrote(fn, a0, abytes(g))) return 0;
  return ret;
}

extern "C" double energy_p(const species_t *sp,
                           const interpolator_array_t *ia) {
  const char *fn = "energy_p";
  if (!sp || !ia || sp->g != ia->g || !ia->i) {
    fail(fn, "Bad args");
    return 0;
  }
  if (!species_ok(fn, sp) || !ensure_init(fn)) return 0;
  vpic_grid_t G;
  if (!make_grid(fn, sp->g, &G)) return 0;
  vpic_particle_t *p;
  vpic_interpolator_t *f;
  if (!dev_in(fn, sp->p, pbytes(sp->max_np), pbytes(sp->np), &p) ||
      !dev_in(fn, ia->i, ibytes(sp->g), ibytes(sp->g), &f))
    return 0;
  double en = 0;
  if (vpic_hip_energy_p(p, sp->np, sp->q, sp->m, f, &G, &en, nullptr) != 0) {
    fail(fn, "%s", vpic_hip_last_error());
    return 0;
  }
  return en;
}

extern "C" void accumulate_rho_p(field_array_t *fa, const species_t *sp) {
  const char *fn = "accumulate_rho_p";
  if (!fa || !sp || fa->g != sp->g || !fa->f) {
    fail(fn, "Bad args");
    return;
  }
  if (!species_ok(fn, sp)) return;
  vpic_field_t *f;
  vpic_grid_t G;
  vpic_particle_t *p;
  if (!field_call(fn, fa, &f, &G) ||
      !dev_in(fn, sp->p, pbytes(sp->max_np), pbytes(sp->np), &p))
    return;
  if (vpic_hip_accumulate_rho_p(f, p, sp->np, sp->q, &G, nullptr) != 0) {
    fail(fn, "%s", vpic_hip_last_error());
    return;
  }
  dev_wrote(fn, fa->f, fbytes(fa->g));
}

// accumulate_rhob (rho_p.cc:126-173) of one host particle into the host
// field array f: the particle is staged to the device and deposited into
// f's device copy.
extern "C" void accumulate_rhob(field_t *f, const particle_t *p,
                                const grid_t *g, const float qsp) {
  const char *fn = "accumulate_rhob";
  if (!f || !p || !g) {
    fail(fn, "Bad args");
    return;
  }
  if (!ensure_init(fn)) return;
  vpic_grid_t G;
  if (!make_grid(fn, g, &G)) return;
  const size_t fb = sizeof(field_t) * (size_t)g->nv;
  vpic_field_t *fd;
  if (!dev_in(fn, f, fb, fb, &fd)) return;
  HIPOK_V(fn, hipMemcpy(st().one, p, sizeof(particle_t),
                        hipMemcpyHostToDevice));
  if (vpic_hip_accumulate_rhob(fd, (const vpic_particle_t *)st().one, 1, qsp,
                               &G, nullptr) != 0) {
    fail(fn, "%s", vpic_hip_last_error());
    return;
  }
  dev_wrote(fn, f, fb);
}

static void center(const char *fn, species_t *sp,
                   const interpolator_array_t *ia, int uncenter) {
  if (!sp || !ia || sp->g != ia->g || !ia->i) {
    fail(fn, "Bad args.");
    return;
  }
  if (!species_ok(fn, sp) || !ensure_init(fn)) return;
  vpic_grid_t G;
  if (!make_grid(fn, sp->g, &G)) return;
  vpic_particle_t *p;
  vpic_interpolator_t *f;
  if (!dev_in(fn, sp->p, pbytes(sp->max_np), pbytes(sp->np), &p) ||
      !dev_in(fn, ia->i, ibytes(sp->g), ibytes(sp->g), &f))
    return;
  if (vpic_hip_center_p(p, sp->np, sp->q, sp->m, f, &G, uncenter, nullptr)) {
    fail(fn, "%s", vpic_hip_last_error());
    return;
  }
  dev_wrote(fn, sp->p, pbytes(sp->np));
}

extern "C" void center_p(species_t *sp, const interpolator_array_t *ia) {
  center("center_p", sp, ia, 0);
}

extern "C" void uncenter_p(species_t *sp, const interpolator_array_t *ia) {
  center("uncenter_p", sp, ia, 1);
}

extern "C" void accumulate_hydro_p(hydro_array_t *ha, const species_t *sp,
                                   const interpolator_array_t *ia) {
  const char *fn = "accumulate_hydro_p";
  // hydro_p_pipeline.cc:224-231
  if (!ha || !sp || !ia || ha->g != sp->g || ha->g != ia->g || !ha->h ||
      !ia->i) {
    fail(fn, "Bad args.");
    return;
  }
  if (!species_ok(fn, sp) || !ensure_init(fn)) return;
  const grid_t *g = sp->g;
  vpic_grid_t G;
  if (!make_grid(fn, g, &G)) return;
  vpic_hydro_t *h;
  vpic_particle_t *p;
  vpic_interpolator_t *f;
  if (!dev_in(fn, ha->h, hbytes(g), hbytes(g), &h) ||
      !dev_in(fn, sp->p, pbytes(sp->max_np), pbytes(sp->np), &p) ||
      !dev_in(fn, ia->i, ibytes(g), ibytes(g), &f))
    return;
  const SpeciesState &ss = st().species[sp];
  const int32_t *part =
      (ss.sorted && ss.part_nv == g->nv) ? ss.dpart : nullptr;
  if (vpic_hip_accumulate_hydro_p(h, p, sp->np, sp->q, sp->m, f, &G, part,
                                  nullptr) != 0) {
    fail(fn, "%s", vpic_hip_last_error());
    return;
  }
  dev_wrote(fn, ha->h, hbytes(g));
}

// Device staging buffer of at least `bytes`.
static bool stage(const char *fn, size_t bytes, void **out) {
  State &S = st();
  if (S.stage_cap < bytes) {
    if (S.stage) HIPOK(fn, hipFree(S.stage));
    S.stage = nullptr;
    S.stage_cap = 0;
    HIPOK(fn, hipMalloc(&S.stage, bytes));
    S.stage_cap = bytes;
  }
  *out = S.stage;
  return true;
}

extern "C" void boundary_p(particle_bc_t *pbc_list, species_t *sp_list,
                           field_array_t *fa, accumulator_array_t *aa) {
  const char *fn = "boundary_p";
  if (!sp_list) return;  // boundary_p.cc:98-101: nothing to do
  if (!fa || !aa || !fa->f || !aa->a || sp_list->g != aa->g ||
      fa->g != aa->g) {
    fail(fn, "Bad args.");
    return;
  }
  // boundary_p.cc:111-127: the custom boundaries by code (id = -3 - k)
  enum { MAX_PBC = 32 };
  particle_bc_t *pbc_of[MAX_PBC] = {nullptr};
  int nb = 0;
  for (particle_bc_t *b = pbc_list; b; b = b->next) nb++;
  if (nb > MAX_PBC) {
    fail(fn, "Update this to support more particle boundary conditions.");
    return;
  }
  for (particle_bc_t *b = pbc_list; b; b = b->next) {
    const int64_t k = -b->id - 3;
    if (k < 0 || k >= MAX_PBC) {
      fail(fn, "particle boundary id %lld is not a custom boundary id",
           (long long)b->id);
      return;
    }
    pbc_of[k] = b;
  }
  if (!ensure_init(fn)) return;
  const grid_t *g = fa->g;
  vpic_grid_t G;
  if (!make_grid(fn, g, &G)) return;
  vpic_field_t *f;
  if (!dev_in(fn, fa->f, fbytes(g), fbytes(g), &f)) return;
  std::vector<particle_injector_t> ci;
  std::vector<vpic_boundary_hit_t> hits;
  bool handled = false, f_on_host = false;
  for (species_t *sp = sp_list; sp; sp = sp->next) {
    if (!species_ok(fn, sp) || sp->g != g) {
      fail(fn, "Bad args.");
      return;
    }
    const int nm = sp->nm;
    if (nm <= 0) {
      sp->nm = 0;
      continue;
    }
    vpic_particle_t *p;
    vpic_particle_mover_t *pm;
    void *dh;
    if (!dev_in(fn, sp->p, pbytes(sp->max_np), pbytes(sp->np), &p) ||
        !dev_in(fn, sp->pm, (size_t)sp->max_nm * sizeof(particle_mover_t),
                (size_t)nm * sizeof(particle_mover_t), &pm) ||
        !stage(fn, (size_t)nm * sizeof(vpic_boundary_hit_t), &dh))
      return;
    int32_t np = sp->np, counts[2] = {0, 0};
    if (vpic_hip_boundary_p(p, &np, pm, nm, f, sp->q, &G,
                            (vpic_boundary_hit_t *)dh, counts, nullptr) != 0) {
      fail(fn, "%s", vpic_hip_last_error());
      return;
    }
    sp->np = np;
    sp->nm = 0;
    if (!dev_wrote(fn, sp->p, pbytes(np))) return;
    // boundary_p.cc:348-352, once per species with the count
    if (counts[1])
      warn(fn, "Unknown boundary interaction ... dropping %d particles "
               "(species=%s)", counts[1], sp->name ? sp->name : "");
    if (!counts[0]) continue;
    // custom boundaries: the handlers run on the host, last mover first,
    // with the host field array current (they may deposit into rhob)
    hits.resize(nm);
    HIPOK_V(fn, hipMemcpy(hits.data(), dh, (size_t)nm * sizeof(hits[0]),
                          hipMemcpyDeviceToHost));
    if (!f_on_host) {
      if (!download(fn, fa->f, fbytes(g))) return;
      f_on_host = true;
    }
    for (int m = nm - 1; m >= 0; m--) {
      vpic_boundary_hit_t &h = hits[m];
      if (h.code > -3) continue;
      const int64_t k = -(int64_t)h.code - 3;
      particle_bc_t *b = k < MAX_PBC ? pbc_of[k] : nullptr;
      if (!b || !b->interact) {
        warn(fn, "Unknown boundary interaction ... dropping particle "
                 "(species=%s)", sp->name ? sp->name : "");
        continue;
      }
      ci.resize(ci.size() + 1);
      const int n = b->interact(b->params, sp, (particle_t *)&h.p,
                                (particle_mover_t *)&h.m,
                                &ci[ci.size() - 1], 1, h.face);
      ci.resize(ci.size() - 1 + (n > 0 ? 1 : 0));
      handled = true;
    }
  }
  if (f_on_host) {
    // the handlers' rhob deposits (and the absorbed charges) go back
    if (!upload(fn, fa->f, fbytes(g), fbytes(g))) return;
  } else if (!dev_wrote(fn, fa->f, fbytes(g))) {
    return;
  }
  (void)handled;
  if (ci.empty()) return;
  // boundary_p.cc:595-711: the injectors last first, per species in the
  // order the handlers produced them; move_p deposits into aa's copy 0
  vpic_accumulator_t *a;
  if (!dev_in(fn, aa->a, abytes(g), abytes(g), &a)) return;
  for (species_t *sp = sp_list; sp; sp = sp->next) {
    std::vector<particle_injector_t> mine;
    for (const particle_injector_t &c : ci)
      if (c.sp_id == sp->id) mine.push_back(c);
    if (mine.empty()) continue;
    vpic_particle_t *p;
    vpic_particle_mover_t *pm;
    void *dpi;
    const size_t pib = mine.size() * sizeof(particle_injector_t);
    if (!dev_in(fn, sp->p, pbytes(sp->max_np), pbytes(sp->np), &p) ||
        !dev_in(fn, sp->pm, (size_t)sp->max_nm * sizeof(particle_mover_t),
                (size_t)sp->nm * sizeof(particle_mover_t), &pm) ||
        !stage(fn, pib, &dpi))
      return;
    HIPOK_V(fn, hipMemcpy(dpi, mine.data(), pib, hipMemcpyHostToDevice));
    int32_t np = sp->np, nm = sp->nm, counts[2] = {0, 0};
    if (vpic_hip_inject_p(p, &np, sp->max_np, pm, &nm, sp->max_nm,
                          (const vpic_particle_injector_t *)dpi,
                          (int32_t)mine.size(), a, sp->q, &G, counts,
                          nullptr) != 0) {
      fail(fn, "%s", vpic_hip_last_error());
      return;
    }
    // boundary_p.cc:716-738 (the fixed-size build's warnings)
    if (counts[0])
      warn(fn, "Dropped %i particles from species \"%s\".  Use a larger "
               "local particle allocation in your simulation setup for this "
               "species on this node.", counts[0], sp->name ? sp->name : "");
    if (counts[1])
      warn(fn, "%i particles were not completed moved to their final "
               "location this timestep for species \"%s\".  Use a larger "
               "local particle mover buffer in your simulation setup for "
               "this species on this node.", counts[1],
           sp->name ? sp->name : "");
    sp->np = np;
    sp->nm = nm;
    if (!dev_wrote(fn, sp->p, pbytes(np)) ||
        !dev_wrote(fn, sp->pm, (size_t)nm * sizeof(particle_mover_t)))
      return;
  }
  dev_wrote(fn, aa->a, abytes(g));
}

// ---------------------------------------------------------------------------
// field-particle interface
// ---------------------------------------------------------------------------
extern "C" void load_interpolator_array(interpolator_array_t *ia,
                                        const field_array_t *fa) {
  const char *fn = "load_interpolator_array";
  if (!ia || !fa || ia->g != fa->g || !ia->i) {
    fail(fn, "Bad args");
    return;
  }
  vpic_field_t *f;
  vpic_grid_t G;
  vpic_interpolator_t *fi;
  if (!field_call(fn, (field_array_t *)fa, &f, &G) ||
      !dev_in(fn, ia->i, ibytes(ia->g), ibytes(ia->g), &fi))
    return;
  if (vpic_hip_load_interpolator_array(fi, f, &G, nullptr) != 0) {
    fail(fn, "%s", vpic_hip_last_error());
    return;
  }
  dev_wrote(fn, ia->i, ibytes(ia->g));
}

extern "C" void clear_accumulator_array(accumulator_array_t *aa) {
  const char *fn = "clear_accumulator_array";
  if (!aa || !aa->a || !aa->g) {
    fail(fn, "Bad args.");
    return;
  }
  if (!ensure_init(fn)) return;
  vpic_grid_t G;
  if (!make_grid(fn, aa->g, &G)) return;
  vpic_accumulator_t *a;
  if (!dev_out(fn, aa->a, abytes(aa->g), &a)) return;
  if (vpic_hip_clear_accumulator_array(a, &G, nullptr) != 0) {
    fail(fn, "%s", vpic_hip_last_error());
    return;
  }
  if (!dev_wrote(fn, aa->a, abytes(aa->g))) return;
  // The other copies (pipelines' copies of a reference-built array) get
  // the reference's clear range, clear_array_pipeline.cc:53-56.
  if (!st().resident && aa->n_pipeline > 0) {
    const grid_t *g = aa->g;
    const int nx = g->nx, ny = g->ny, nz = g->nz;
    const int vl = 1 + (nx + 2) * (1 + (ny + 2) * 1);
    const int vh = nx + (nx + 2) * (ny + (ny + 2) * nz);
    const int i0 = (vl / 2) * 2;
    const int na = (((vh - i0 + 1) + 1) / 2) * 2;
    for (int c = 1; c <= aa->n_pipeline; c++)
      memset(aa->a + (size_t)c * aa->stride + i0, 0,
             (size_t)na * sizeof(accumulator_t));
  }
}

extern "C" void reduce_accumulator_array(accumulator_array_t *aa) {
  const char *fn = "reduce_accumulator_array";
  if (!aa || !aa->a || !aa->g) {
    fail(fn, "Bad args.");
    return;
  }
  // One device copy: the currents are already in copy 0.
}

extern "C" void unload_accumulator_array(field_array_t *fa,
                                         const accumulator_array_t *aa) {
  const char *fn = "unload_accumulator_array";
  if (!fa || !aa || fa->g != aa->g || !aa->a) {
    fail(fn, "Bad args");
    return;
  }
  vpic_field_t *f;
  vpic_grid_t G;
  vpic_accumulator_t *a;
  if (!field_call(fn, fa, &f, &G) ||
      !dev_in(fn, aa->a, abytes(aa->g), abytes(aa->g), &a))
    return;
  if (vpic_hip_unload_accumulator_array(f, a, &G, nullptr) != 0) {
    fail(fn, "%s", vpic_hip_last_error());
    return;
  }
  dev_wrote(fn, fa->f, fbytes(fa->g));
}

extern "C" void clear_hydro_array(hydro_array_t *ha) {
  const char *fn = "clear_hydro_array";
  if (!ha || !ha->h || !ha->g) {
    fail(fn, "Bad args");
    return;
  }
  if (!ensure_init(fn)) return;
  vpic_grid_t G;
  if (!make_grid(fn, ha->g, &G)) return;
  vpic_hydro_t *h;
  if (!dev_out(fn, ha->h, hbytes(ha->g), &h)) return;
  if (vpic_hip_clear_hydro_array(h, &G, nullptr) != 0) {
    fail(fn, "%s", vpic_hip_last_error());
    return;
  }
  if (!dev_wrote(fn, ha->h, hbytes(ha->g))) return;
  // clear_array_pipeline.cc:72-94 clears nv entries of every copy; the
  // host-only copies 1..n_pipeline of a reference-built array get the same.
  if (!st().resident)
    for (int c = 1; c <= ha->n_pipeline; c++)
      memset(ha->h + (size_t)c * ha->stride, 0, hbytes(ha->g));
}

extern "C" void reduce_hydro_array(hydro_array_t *ha) {
  const char *fn = "reduce_hydro_array";
  if (!ha || !ha->h || !ha->g) {
    fail(fn, "Bad args.");
    return;
  }
  // One device copy: the moments are already in copy 0.
}

extern "C" void synchronize_hydro_array(hydro_array_t *ha) {
  const char *fn = "synchronize_hydro_array";
  if (!ha || !ha->h || !ha->g) {
    fail(fn, "NULL hydro array.");
    return;
  }
  if (!ensure_init(fn)) return;
  vpic_grid_t G;
  if (!make_grid(fn, ha->g, &G)) return;
  vpic_hydro_t *h;
  if (!dev_in(fn, ha->h, hbytes(ha->g), hbytes(ha->g), &h)) return;
  if (vpic_hip_synchronize_hydro_array(h, &G, nullptr) != 0) {
    fail(fn, "%s", vpic_hip_last_error());
    return;
  }
  dev_wrote(fn, ha->h, hbytes(ha->g));
}

// ---------------------------------------------------------------------------
// field kernels (the entries of field_advance_kernels_t)
// ---------------------------------------------------------------------------
extern "C" void vpic_hip_fa_advance_b(field_array_t *fa, float frac) {
  const char *fn = "advance_b";
  vpic_field_t *f;
  vpic_grid_t G;
  if (!field_call(fn, fa, &f, &G)) return;
  if (vpic_hip_advance_b(f, &G, frac, nullptr) != 0) {
    fail(fn, "%s", vpic_hip_last_error());
    return;
  }
  dev_wrote(fn, fa->f, fbytes(fa->g));
}

extern "C" void vpic_hip_fa_advance_e(field_array_t *fa, float frac) {
  const char *fn = "vacuum_advance_e";
  if (!fa) {
    fail(fn, "Bad args");
    return;
  }
  // vacuum_advance_e.cc:20-23
  if (frac != 1) {
    fail(fn, "standard advance_e does not support frac != 1 yet");
    return;
  }
  vpic_field_params_t mp;
  vpic_field_t *f;
  vpic_grid_t G;
  if (!field_params(fn, fa, &mp) || !field_call(fn, fa, &f, &G)) return;
  if (vpic_hip_vacuum_advance_e(f, &G, &mp, frac, nullptr) != 0) {
    fail(fn, "%s", vpic_hip_last_error());
    return;
  }
  dev_wrote(fn, fa->f, fbytes(fa->g));
}

extern "C" void vpic_hip_fa_energy_f(double *en, const field_array_t *fa) {
  const char *fn = "vacuum_energy_f";
  if (!en || !fa) {
    fail(fn, "Bad args");
    return;
  }
  vpic_field_params_t mp;
  vpic_field_t *f;
  vpic_grid_t G;
  if (!field_params(fn, fa, &mp) ||
      !field_call(fn, (field_array_t *)fa, &f, &G))
    return;
  if (vpic_hip_energy_f(en, f, &G, &mp, nullptr) != 0)
    fail(fn, "%s", vpic_hip_last_error());
}

#define FIELD_SIMPLE(NAME, ABI_FN)                                      \
  extern "C" void vpic_hip_fa_##NAME(field_array_t *fa) {               \
    const char *fn = #NAME;                                             \
    vpic_field_t *f;                                                    \
    vpic_grid_t G;                                                      \
    if (!field_call(fn, fa, &f, &G)) return;                            \
    if (ABI_FN(f, &G, nullptr) != 0) {                                  \
      fail(fn, "%s", vpic_hip_last_error());                            \
      return;                                                           \
    }                                                                   \
    dev_wrote(fn, fa->f, fbytes(fa->g));                                \
  }

FIELD_SIMPLE(clear_jf, vpic_hip_clear_jf)
FIELD_SIMPLE(clear_rhof, vpic_hip_clear_rhof)
FIELD_SIMPLE(synchronize_jf, vpic_hip_synchronize_jf)
FIELD_SIMPLE(synchronize_rho, vpic_hip_synchronize_rho)
FIELD_SIMPLE(compute_div_b_err, vpic_hip_compute_div_b_err)
FIELD_SIMPLE(clean_div_b, vpic_hip_clean_div_b)

#define FIELD_WITH_PARAMS(NAME, ABI_FN)                                 \
  extern "C" void vpic_hip_fa_##NAME(field_array_t *fa) {               \
    const char *fn = #NAME;                                             \
    if (!fa) {                                                          \
      fail(fn, "Bad args");                                             \
      return;                                                           \
    }                                                                   \
    vpic_field_params_t mp;                                             \
    vpic_field_t *f;                                                    \
    vpic_grid_t G;                                                      \
    if (!field_params(fn, fa, &mp) || !field_call(fn, fa, &f, &G))      \
      return;                                                           \
    if (ABI_FN(f, &G, &mp, nullptr) != 0) {                             \
      fail(fn, "%s", vpic_hip_last_error());                            \
      return;                                                           \
    }                                                                   \
    dev_wrote(fn, fa->f, fbytes(fa->g));                                \
  }

FIELD_WITH_PARAMS(vacuum_compute_div_e_err, vpic_hip_compute_div_e_err)
FIELD_WITH_PARAMS(vacuum_clean_div_e, vpic_hip_clean_div_e)
FIELD_WITH_PARAMS(vacuum_compute_rhob, vpic_hip_compute_rhob)
FIELD_WITH_PARAMS(vacuum_compute_curl_b, vpic_hip_compute_curl_b)

#define FIELD_RMS(NAME, ABI_FN)                                         \
  extern "C" double vpic_hip_fa_##NAME(const field_array_t *fa) {       \
    const char *fn = #NAME;                                             \
    vpic_field_t *f;                                                    \
    vpic_grid_t G;                                                      \
    double r = 0;                                                       \
    if (!field_call(fn, (field_array_t *)fa, &f, &G)) return 0;         \
    if (ABI_FN(&r, f, &G, nullptr) != 0)                                \
      fail(fn, "%s", vpic_hip_last_error());                            \
    return r;                                                           \
  }

FIELD_RMS(compute_rms_div_e_err, vpic_hip_compute_rms_div_e_err)
FIELD_RMS(compute_rms_div_b_err, vpic_hip_compute_rms_div_b_err)

extern "C" double vpic_hip_fa_synchronize_tang_e_norm_b(field_array_t *fa) {
  const char *fn = "synchronize_tang_e_norm_b";
  vpic_field_t *f;
  vpic_grid_t G;
  double err = 0;
  if (!field_call(fn, fa, &f, &G)) return 0;
  if (vpic_hip_synchronize_tang_e_norm_b(&err, f, &G, nullptr) != 0) {
    fail(fn, "%s", vpic_hip_last_error());
    return 0;
  }
  dev_wrote(fn, fa->f, fbytes(fa->g));
  return err;
}

static void fa_advance_e_entry(field_array_t *fa, float frac) {
  vpic_hip_fa_advance_e(fa, frac);
}

extern "C" int vpic_hip_install_field_kernels(field_array_t *fa) {
  if (!fa || !fa->g || !fa->f)
    return vpic_set_error(-1, "install_field_kernels: Bad args");
  const sfa_params_t *sp = (const sfa_params_t *)fa->params;
  if (!sp || !sp->mc)
    return vpic_set_error(-1, 
        "install_field_kernels: field array has no sfa_params_t");
  // sfa.cc:205-214: the vacuum kernels are the one-material kernels
  if (sp->n_mc != 1)
    return vpic_set_error(-1, 
        "install_field_kernels: %d materials; the device path implements "
        "the one-material (vacuum) kernels only", sp->n_mc);
  field_advance_kernels_t *k = fa->kernel;
  k->advance_b = vpic_hip_fa_advance_b;
  k->advance_e = fa_advance_e_entry;
  k->energy_f = vpic_hip_fa_energy_f;
  k->clear_jf = vpic_hip_fa_clear_jf;
  k->synchronize_jf = vpic_hip_fa_synchronize_jf;
  k->clear_rhof = vpic_hip_fa_clear_rhof;
  k->synchronize_rho = vpic_hip_fa_synchronize_rho;
  k->compute_rhob = vpic_hip_fa_vacuum_compute_rhob;
  k->compute_curl_b = vpic_hip_fa_vacuum_compute_curl_b;
  k->synchronize_tang_e_norm_b = vpic_hip_fa_synchronize_tang_e_norm_b;
  k->compute_div_e_err = vpic_hip_fa_vacuum_compute_div_e_err;
  k->compute_rms_div_e_err = vpic_hip_fa_compute_rms_div_e_err;
  k->clean_div_e = vpic_hip_fa_vacuum_clean_div_e;
  k->compute_div_b_err = vpic_hip_fa_compute_div_b_err;
  k->compute_rms_div_b_err = vpic_hip_fa_compute_rms_div_b_err;
  k->clean_div_b = vpic_hip_fa_clean_div_b;
  return 0;
}

// ---------------------------------------------------------------------------
// residency control
// ---------------------------------------------------------------------------
extern "C" int vpic_hip_dropin_resident(int on) {
  State &S = st();
  if (S.resident && !on) {
    for (auto &kv : S.mirrors)
      if (kv.second.host_stale && kv.second.live &&
          hipMemcpy((void *)kv.first, kv.second.dev, kv.second.live,
                    hipMemcpyDeviceToHost) != hipSuccess)
        return vpic_set_error(-1, "dropin_resident: download failed");
      else
        kv.second.host_stale = false;
  }
  S.resident = on != 0;
  return 0;
}

#define SYNC_FAIL(what) return vpic_set_error(-1, "%s", what)

extern "C" int vpic_hip_species_sync(species_t *sp, int to_device) {
  if (!sp || !sp->g || sp->np < 0 || sp->np > sp->max_np)
    SYNC_FAIL("species_sync: Bad args");
  if (!ensure_init("species_sync")) SYNC_FAIL("species_sync: init failed");
  const size_t pm_cap = (size_t)(sp->max_nm > 0 ? sp->max_nm : 0) *
                        sizeof(particle_mover_t);
  const size_t pm_live = (size_t)(sp->nm > 0 ? sp->nm : 0) *
                         sizeof(particle_mover_t);
  if (to_device) {
    if (!upload("species_sync", sp->p, pbytes(sp->max_np), pbytes(sp->np)) ||
        (sp->pm && !upload("species_sync", sp->pm, pm_cap, pm_live)))
      SYNC_FAIL("species_sync: upload failed");
    return 0;
  }
  if (!download("species_sync", sp->p, pbytes(sp->np)) ||
      (sp->pm && !download("species_sync", sp->pm, pm_live)))
    SYNC_FAIL("species_sync: download failed");
  auto it = st().species.find(sp);
  if (sp->partition && it != st().species.end() && it->second.sorted &&
      it->second.part_nv == sp->g->nv &&
      hipMemcpy(sp->partition, it->second.dpart,
                ((size_t)sp->g->nv + 1) * sizeof(int32_t),
                hipMemcpyDeviceToHost) != hipSuccess)
    SYNC_FAIL("species_sync: partition download failed");
  return 0;
}

extern "C" int vpic_hip_field_array_sync(field_array_t *fa, int to_device) {
  if (!fa || !fa->f || !fa->g) SYNC_FAIL("field_array_sync: Bad args");
  if (!ensure_init("field_array_sync")) SYNC_FAIL("init failed");
  bool ok = to_device ? upload("field_array_sync", fa->f, fbytes(fa->g),
                               fbytes(fa->g))
                      : download("field_array_sync", fa->f, fbytes(fa->g));
  return ok ? 0 : vpic_set_error(-1, "field_array_sync failed");
}

extern "C" int vpic_hip_interpolator_array_sync(interpolator_array_t *ia,
                                                int to_device) {
  if (!ia || !ia->i || !ia->g) SYNC_FAIL("interpolator_array_sync: Bad args");
  if (!ensure_init("interpolator_array_sync")) SYNC_FAIL("init failed");
  bool ok = to_device ? upload("interpolator_array_sync", ia->i,
                               ibytes(ia->g), ibytes(ia->g))
                      : download("interpolator_array_sync", ia->i,
                                 ibytes(ia->g));
  return ok ? 0 : vpic_set_error(-1, "interpolator_array_sync failed");
}

extern "C" int vpic_hip_accumulator_array_sync(accumulator_array_t *aa,
                                               int to_device) {
  if (!aa || !aa->a || !aa->g) SYNC_FAIL("accumulator_array_sync: Bad args");
  if (!ensure_init("accumulator_array_sync")) SYNC_FAIL("init failed");
  bool ok = to_device ? upload("accumulator_array_sync", aa->a,
                               abytes(aa->g), abytes(aa->g))
                      : download("accumulator_array_sync", aa->a,
                                 abytes(aa->g));
  return ok ? 0 : vpic_set_error(-1, "accumulator_array_sync failed");
}

extern "C" int vpic_hip_hydro_array_sync(hydro_array_t *ha, int to_device) {
  if (!ha || !ha->h || !ha->g) SYNC_FAIL("hydro_array_sync: Bad args");
  if (!ensure_init("hydro_array_sync")) SYNC_FAIL("init failed");
  bool ok = to_device ? upload("hydro_array_sync", ha->h, hbytes(ha->g),
                               hbytes(ha->g))
                      : download("hydro_array_sync", ha->h, hbytes(ha->g));
  return ok ? 0 : vpic_set_error(-1, "hydro_array_sync failed");
}

extern "C" int vpic_hip_dropin_release(const void *host_array) {
  State &S = st();
  auto it = S.mirrors.find(host_array);
  if (it != S.mirrors.end()) {
    if (it->second.dev) (void)hipFree(it->second.dev);
    S.mirrors.erase(it);
  }
  auto js = S.species.find((const species_t *)host_array);
  if (js != S.species.end()) {
    if (js->second.dpart) (void)hipFree(js->second.dpart);
    S.species.erase(js);
  }
  return 0;
}

extern "C" int vpic_hip_dropin_release_all(void) {
  State &S = st();
  for (auto &kv : S.mirrors)
    if (kv.second.dev) (void)hipFree(kv.second.dev);
  for (auto &kv : S.species)
    if (kv.second.dpart) (void)hipFree(kv.second.dpart);
  S.mirrors.clear();
  S.species.clear();
  if (S.aux) (void)hipFree(S.aux);
  S.aux = nullptr;
  S.aux_cap = 0;
  return 0;
}

extern "C" int vpic_hip_dropin_error_handler(vpic_hip_error_handler_t h) {
  st().handler = h;
  return 0;
}

extern "C" int vpic_hip_dropin_layout(int64_t *out, int n) {
  const int64_t table[] = {VPIC_DROPIN_LAYOUT(VPIC_DROPIN_SIZE,
                                              VPIC_DROPIN_OFFSET)};
  const int count = (int)(sizeof table / sizeof table[0]);
  for (int k = 0; k < count && k < n; k++) out[k] = table[k];
  return count;
}
