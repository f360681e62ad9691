// api.hip -- the C-ABI of libvpichip (include/vpic_hip.h): argument checks,
// error reporting, per-device scratch, push timing and the host set-up
// helpers.  Every entry point returns 0 or a negative status and leaves a
// message for vpic_hip_last_error().
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <cmath>
#include <mutex>
#include <vector>

#include "common.h"

// ---------------------------------------------------------------------------
// errors
// ---------------------------------------------------------------------------
static thread_local char g_err[512];

int vpic_set_error(int code, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

int vpic_check_hip(hipError_t e, const char *what) {
  return vpic_set_error(VPIC_EHIP, "%s: %s", what, hipGetErrorString(e));
}

int scratch_reserve(DeviceScratch &s, size_t bytes) {
  if (bytes <= s.bytes && s.ptr) return VPIC_OK;
  if (s.ptr) {
    hipError_t e = hipFree(s.ptr);
    s.ptr = nullptr;
    s.bytes = 0;
    if (e != hipSuccess) return vpic_check_hip(e, "hipFree(scratch)");
  }
  size_t want = bytes + bytes / 8 + 256;
  hipError_t e = hipMalloc(&s.ptr, want);
  if (e != hipSuccess) {
    s.ptr = nullptr;
    return vpic_set_error(VPIC_ENOMEM, "hipMalloc(%zu bytes of scratch): %s",
                          want, hipGetErrorString(e));
  }
  s.bytes = want;
  return VPIC_OK;
}

#define CHECK(cond, ...)                                       \
  do {                                                         \
    if (!(cond)) return vpic_set_error(VPIC_EINVAL, __VA_ARGS__); \
  } while (0)

static bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

static int check_grid(const vpic_grid_t *g) {
  CHECK(g, "grid is NULL");
  CHECK(g->nx >= 1 && g->ny >= 1 && g->nz >= 1, "bad grid resolution %dx%dx%d",
        g->nx, g->ny, g->nz);
  CHECK(g->sx == 1 && g->sy == g->nx + 2 && g->sz == (g->ny + 2) * g->sy &&
            (int64_t)g->nv == (int64_t)(g->nz + 2) * g->sz,
        "grid strides inconsistent with %dx%dx%d", g->nx, g->ny, g->nz);
  CHECK(g->rangeh - g->rangel + 1 == g->nv, "grid range [%lld,%lld] != nv",
        (long long)g->rangel, (long long)g->rangeh);
  for (int k = 0; k < 6; k++) {
    const int b = g->fbc[k];
    CHECK(b == VPIC_BC_PERIODIC || b == VPIC_BC_ANTI_SYMMETRIC_FIELDS ||
              b == VPIC_BC_SYMMETRIC_FIELDS || b == VPIC_BC_PMC_FIELDS ||
              b == VPIC_BC_ABSORB_FIELDS,
          "unknown field boundary code %d on face %d", b, k);
  }
  for (int k = 0; k < 3; k++)
    CHECK((g->fbc[k] == VPIC_BC_PERIODIC) == (g->fbc[k + 3] == VPIC_BC_PERIODIC),
          "axis %d: a periodic field face needs a periodic partner", k);
  for (int k = 0; k < 3; k++)
    CHECK((g->pbc[k] == VPIC_BC_PERIODIC) == (g->pbc[k + 3] == VPIC_BC_PERIODIC),
          "axis %d: a periodic particle face needs a periodic partner", k);
  return VPIC_OK;
}

static int check_fields_supported(const vpic_grid_t *g) {
  for (int k = 0; k < 6; k++)
    CHECK(g->fbc[k] != VPIC_BC_ABSORB_FIELDS,
          "absorbing field boundaries are not supported by the device path");
  return VPIC_OK;
}

static hipStream_t S(void *stream) { return (hipStream_t)stream; }

// ---------------------------------------------------------------------------
// library state
// ---------------------------------------------------------------------------
extern "C" int vpic_hip_init(int device) {
  int n = 0;
  VPIC_HIP(hipGetDeviceCount(&n));
  CHECK(device >= 0 && device < n, "device %d out of range (%d devices)",
        device, n);
  VPIC_HIP(hipSetDevice(device));
  VPIC_HIP(hipFree(nullptr));
  // the device address of the no-allocation push's fixed scratch
  return push_noalloc_setup();
}

extern "C" const char *vpic_hip_last_error(void) { return g_err; }

extern "C" const char *vpic_hip_version(void) {
  return "libvpichip 0.1 (gfx950)";
}

// ---------------------------------------------------------------------------
// push timing
// ---------------------------------------------------------------------------
namespace {
struct Timing {
  std::mutex mu;
  bool on = false;
  std::vector<hipEvent_t> ev;  // begin/end pairs
  int64_t launches = 0, particles = 0;
  double total_ms = 0;
};
Timing &timing() {
  static Timing t;
  return t;
}
}  // namespace

bool push_timing_on() { return timing().on; }

void push_timing_record(hipStream_t s, int np, bool begin) {
  Timing &t = timing();
  std::lock_guard<std::mutex> lk(t.mu);
  hipEvent_t e;
  if (hipEventCreate(&e) != hipSuccess) return;
  (void)hipEventRecord(e, s);
  t.ev.push_back(e);
  if (!begin) {
    t.launches++;
    t.particles += np;
  }
}

extern "C" int vpic_hip_push_timing_enable(int on) {
  timing().on = on != 0;
  return VPIC_OK;
}

extern "C" int vpic_hip_push_timing(int64_t *launches, double *total_ms,
                                    int64_t *particles) {
  Timing &t = timing();
  std::lock_guard<std::mutex> lk(t.mu);
  for (size_t k = 0; k + 1 < t.ev.size(); k += 2) {
    float ms = 0;
    VPIC_HIP(hipEventSynchronize(t.ev[k + 1]));
    VPIC_HIP(hipEventElapsedTime(&ms, t.ev[k], t.ev[k + 1]));
    t.total_ms += ms;
  }
  for (hipEvent_t e : t.ev) (void)hipEventDestroy(e);
  t.ev.clear();
  if (launches) *launches = t.launches;
  if (total_ms) *total_ms = t.total_ms;
  if (particles) *particles = t.particles;
  t.launches = 0;
  t.particles = 0;
  t.total_ms = 0;
  return VPIC_OK;
}

// ---------------------------------------------------------------------------
// species advance
// ---------------------------------------------------------------------------
static int check_push_args(vpic_particle_t *p, int32_t np,
                           vpic_particle_mover_t *pm, int32_t max_nm,
                           vpic_accumulator_t *a, const vpic_interpolator_t *f,
                           const vpic_grid_t *g) {
  int rc = check_grid(g);
  if (rc) return rc;
  // advance_p_pipeline.cc:263-270 ("Bad args.")
  CHECK(np >= 0, "np < 0");
  CHECK(np == 0 || p, "particle array is NULL");
  CHECK(a && f, "accumulator or interpolator array is NULL");
  CHECK(max_nm >= 0, "max_nm < 0");
  CHECK(max_nm == 0 || pm, "mover array is NULL");
  CHECK(aligned16(p) && aligned16(a) && aligned16(f),
        "device arrays must be 16-byte aligned");
  return VPIC_OK;
}

namespace {
DeviceScratch &nm_scratch() {
  static thread_local DeviceScratch s[16];
  int dev = 0;
  (void)hipGetDevice(&dev);
  return s[dev & 15];
}
}  // namespace

extern "C" int vpic_hip_advance_p_async(
    vpic_particle_t *p, int32_t np, float q, float m,
    vpic_particle_mover_t *pm, int32_t max_nm, int32_t *dev_nm,
    vpic_accumulator_t *a, const vpic_interpolator_t *f,
    const vpic_grid_t *g, const int32_t *partition, void *stream) {
  int rc = check_push_args(p, np, pm, max_nm, a, f, g);
  if (rc) return rc;
  CHECK(dev_nm, "dev_nm is NULL");
  return launch_advance_p(p, np, q, m, pm, max_nm, dev_nm, a, f, g, partition,
                          nullptr, S(stream));
}

extern "C" int vpic_hip_advance_p_keys(
    vpic_particle_t *p, int32_t np, float q, float m,
    vpic_particle_mover_t *pm, int32_t max_nm, int32_t *dev_nm,
    vpic_accumulator_t *a, const vpic_interpolator_t *f,
    const vpic_grid_t *g, const int32_t *partition, int32_t *keys,
    void *stream) {
  int rc = check_push_args(p, np, pm, max_nm, a, f, g);
  if (rc) return rc;
  CHECK(dev_nm, "dev_nm is NULL");
  return launch_advance_p(p, np, q, m, pm, max_nm, dev_nm, a, f, g, partition,
                          keys, S(stream));
}

extern "C" int vpic_hip_advance_p_noalloc(
    vpic_particle_t *p, int32_t np, float q, float m,
    vpic_particle_mover_t *pm, int32_t max_nm, int32_t *dev_nm,
    vpic_accumulator_t *a, const vpic_interpolator_t *f,
    const vpic_grid_t *g, void *stream) {
  int rc = check_push_args(p, np, pm, max_nm, a, f, g);
  if (rc) return rc;
  CHECK(dev_nm, "dev_nm is NULL");
  return launch_advance_p_noalloc(p, np, q, m, pm, max_nm, dev_nm, a, f, g,
                                  S(stream));
}

extern "C" int vpic_hip_push_group(int32_t gs) {
  CHECK(push_set_group(gs) == VPIC_OK,
        "push lane group %d not one of 1, 2, 4, 8, 16, 32, 64", gs);
  return VPIC_OK;
}

extern "C" int vpic_hip_push_bin_mode(int32_t mode) {
  CHECK(push_set_bin_mode(mode) == VPIC_OK,
        "push bin mode %d not one of 0, 1, 2", mode);
  return VPIC_OK;
}

extern "C" int vpic_hip_push_direct_runs(int32_t runs) {
  CHECK(runs >= 2 && runs <= 65, "direct-deposit run threshold %d outside "
        "[2,65]", runs);
  return push_set_direct_runs(runs);
}

extern "C" int vpic_hip_push_tile(int32_t tx, int32_t ty, int32_t tz) {
  CHECK(push_set_tile(tx, ty, tz, 1) == VPIC_OK,
        "push tile %dx%dx%d rejected (window of 96 B per cell plus 72 B per "
        "tile cell must fit 140 KiB of LDS, ty*tz <= 256, tx*ty*tz <= 512)",
        tx, ty, tz);
  return VPIC_OK;
}

extern "C" int vpic_hip_push_get_tile(int32_t *out) {
  CHECK(out, "out is NULL");
  int t[4];
  push_get_tile(t);
  for (int k = 0; k < 4; k++) out[k] = t[k];
  return VPIC_OK;
}

extern "C" int vpic_hip_push_tile_window(int32_t tx, int32_t ty, int32_t tz,
                                         int32_t halo) {
  CHECK(halo == 1 || halo == 2, "push window halo %d is not 1 or 2", halo);
  CHECK(push_set_tile(tx, ty, tz, halo) == VPIC_OK,
        "push tile %dx%dx%d with halo %d rejected (window of 96 B per cell "
        "plus 72 B per tile cell must fit 140 KiB of LDS, ty*tz <= 256, "
        "tx*ty*tz <= 512)", tx, ty, tz, halo);
  return VPIC_OK;
}

extern "C" int vpic_hip_advance_p(vpic_particle_t *p, int32_t np, float q,
                                  float m, vpic_particle_mover_t *pm,
                                  int32_t max_nm, int32_t *host_nm,
                                  int32_t *host_n_ignored,
                                  vpic_accumulator_t *a,
                                  const vpic_interpolator_t *f,
                                  const vpic_grid_t *g,
                                  const int32_t *partition, void *stream) {
  int rc = check_push_args(p, np, pm, max_nm, a, f, g);
  if (rc) return rc;
  CHECK(host_nm, "host_nm is NULL");
  DeviceScratch &ns = nm_scratch();
  if ((rc = scratch_reserve(ns, 64))) return rc;
  rc = launch_advance_p(p, np, q, m, pm, max_nm, (int *)ns.ptr, a, f, g,
                        partition, nullptr, S(stream));
  if (rc) return rc;
  int32_t out[3];
  VPIC_HIP(hipMemcpyAsync(out, ns.ptr, sizeof(out), hipMemcpyDeviceToHost,
                          S(stream)));
  VPIC_HIP(hipStreamSynchronize(S(stream)));
  *host_nm = out[0];
  if (host_n_ignored) *host_n_ignored = out[1];
  if (out[2])
    return vpic_set_error(VPIC_EINVAL,
                          "advance_p: %d particles have a voxel index outside "
                          "[0,nv) and were not advanced", out[2]);
  return VPIC_OK;
}

extern "C" int vpic_hip_move_p(vpic_particle_t *p, vpic_particle_mover_t *pm,
                               int32_t nm, vpic_accumulator_t *a,
                               const vpic_grid_t *g, float q, int32_t *dev_ret,
                               void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(nm >= 0, "nm < 0");
  if (nm == 0) return VPIC_OK;
  CHECK(p && pm && a && dev_ret, "NULL argument");
  CHECK(aligned16(p), "particle array must be 16-byte aligned");
  return launch_move_p_list(p, pm, nm, a, g, q, dev_ret, S(stream));
}

extern "C" int vpic_hip_boundary_p(vpic_particle_t *p, int32_t *np,
                                   const vpic_particle_mover_t *pm, int32_t nm,
                                   vpic_field_t *f, float q,
                                   const vpic_grid_t *g,
                                   vpic_boundary_hit_t *hits, int32_t *counts,
                                   void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(np && counts, "NULL argument");
  CHECK(*np >= 0 && nm >= 0, "np < 0 or nm < 0");
  CHECK(nm <= *np, "more movers than particles");
  counts[0] = counts[1] = 0;
  if (nm == 0) return VPIC_OK;
  CHECK(p && pm && f, "NULL argument");
  CHECK(aligned16(p), "particle array must be 16-byte aligned");
  return launch_boundary_p(p, np, pm, nm, f, q, g, hits, counts, S(stream));
}

extern "C" int vpic_hip_accumulate_rhob(vpic_field_t *f,
                                        const vpic_particle_t *p, int32_t n,
                                        float qsp, const vpic_grid_t *g,
                                        void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(n >= 0, "n < 0");
  if (n == 0) return VPIC_OK;
  CHECK(f && p, "NULL argument");
  DeviceScratch &ns = nm_scratch();
  if ((rc = scratch_reserve(ns, 64))) return rc;
  VPIC_HIP(hipMemsetAsync(ns.ptr, 0, sizeof(int), S(stream)));
  if ((rc = launch_accumulate_rhob(f, p, n, qsp, g, (int *)ns.ptr,
                                   S(stream))))
    return rc;
  int bad = 0;
  VPIC_HIP(hipMemcpyAsync(&bad, ns.ptr, sizeof(int), hipMemcpyDeviceToHost,
                          S(stream)));
  VPIC_HIP(hipStreamSynchronize(S(stream)));
  CHECK(bad == 0, "accumulate_rhob: %d particles with a voxel whose nodes "
        "are off the grid were skipped", bad);
  return VPIC_OK;
}

extern "C" int vpic_hip_inject_p(vpic_particle_t *p, int32_t *np,
                                 int32_t max_np, vpic_particle_mover_t *pm,
                                 int32_t *nm, int32_t max_nm,
                                 const vpic_particle_injector_t *pi, int32_t n,
                                 vpic_accumulator_t *a, float q,
                                 const vpic_grid_t *g, int32_t *counts,
                                 void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(np && nm && counts, "NULL argument");
  CHECK(*np >= 0 && *nm >= 0 && n >= 0, "negative count");
  CHECK(*np <= max_np && *nm <= max_nm, "count above its capacity");
  counts[0] = counts[1] = 0;
  if (n == 0) return VPIC_OK;
  CHECK(p && pm && pi && a, "NULL argument");
  CHECK(aligned16(p), "particle array must be 16-byte aligned");
  return launch_inject_p(p, np, max_np, pm, nm, max_nm, pi, n, a, q, g,
                         counts, S(stream));
}

extern "C" int vpic_hip_takizuka_abe(
    vpic_particle_t *pi, const int32_t *part_i, int32_t np_i, float qi,
    float mi, vpic_particle_t *pj, const int32_t *part_j, int32_t np_j,
    float qj, float mj, double cvar0, int32_t interval, uint64_t seed,
    uint64_t call, const vpic_grid_t *g, void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(pi && pj && part_i && part_j, "NULL argument");
  CHECK(np_i >= 0 && np_j >= 0, "negative particle count");
  CHECK((pi == pj) == (part_i == part_j) && (pi != pj || np_i == np_j),
        "one species: the same particles, partition and count");
  CHECK(mi > 0 && mj > 0, "masses must be positive");
  CHECK(interval >= 1, "interval must be at least 1");
  CHECK(aligned16(pi) && aligned16(pj), "particle arrays must be 16-byte "
        "aligned");
  return launch_takizuka_abe(pi, part_i, np_i, qi, mi, pj, part_j, np_j, qj,
                             mj, cvar0, interval, seed, call, g, S(stream));
}

extern "C" int vpic_hip_device_synchronize(void) {
  VPIC_HIP(hipDeviceSynchronize());
  return VPIC_OK;
}

extern "C" int vpic_hip_advance_p_species(int32_t nsp,
                                          const vpic_push_species_t *sp,
                                          vpic_accumulator_t *a,
                                          const vpic_interpolator_t *f,
                                          const vpic_grid_t *g, void *stream) {
  CHECK(sp, "sp is NULL");
  CHECK(nsp >= 1 && nsp <= 2, "%d species in one advance_p (1 or 2)", nsp);
  for (int k = 0; k < nsp; k++) {
    int rc = check_push_args(sp[k].p, sp[k].np, sp[k].pm, sp[k].max_nm, a, f,
                             g);
    if (rc) return rc;
    CHECK(sp[k].dev_nm, "species %d: dev_nm is NULL", k);
    for (int j = 0; j < k; j++) {
      CHECK(sp[j].p != sp[k].p || sp[k].np == 0,
            "species %d and %d share a particle array", j, k);
      CHECK(sp[j].dev_nm != sp[k].dev_nm,
            "species %d and %d share their dev_nm counters", j, k);
      CHECK(sp[j].pm != sp[k].pm || sp[k].max_nm == 0,
            "species %d and %d share a mover array", j, k);
      CHECK(sp[j].keys != sp[k].keys || !sp[k].keys,
            "species %d and %d share a key array", j, k);
    }
  }
  return launch_advance_p_set(nsp, sp, a, f, g, S(stream));
}

extern "C" int vpic_hip_sort_p(vpic_particle_t *p, vpic_particle_t *aux,
                               int32_t np, int32_t *partition,
                               const vpic_grid_t *g, int in_place,
                               void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(np >= 0, "np < 0");
  CHECK(partition, "partition is NULL");
  CHECK(np == 0 || (p && aux && p != aux), "particle/aux arrays invalid");
  CHECK(aligned16(p) && aligned16(aux), "arrays must be 16-byte aligned");
  return launch_sort_p(p, aux, np, partition, g, in_place, S(stream));
}

extern "C" int vpic_hip_sort_p_keys(vpic_particle_t *p, vpic_particle_t *aux,
                                    int32_t np, int32_t *partition,
                                    const int32_t *keys, const vpic_grid_t *g,
                                    int in_place, void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(np >= 0, "np < 0");
  CHECK(partition, "partition is NULL");
  CHECK(np == 0 || (p && aux && p != aux), "particle/aux arrays invalid");
  CHECK(aligned16(p) && aligned16(aux), "arrays must be 16-byte aligned");
  return launch_sort_p_keys(p, aux, np, partition, keys, g, in_place,
                            S(stream));
}

extern "C" int vpic_hip_sort_keys_stale(int64_t *count) {
  CHECK(count, "count is NULL");
  *count = sort_keys_stale_count();
  return VPIC_OK;
}

extern "C" int vpic_hip_energy_p(const vpic_particle_t *p, int32_t np,
                                 float q, float m,
                                 const vpic_interpolator_t *f,
                                 const vpic_grid_t *g, double *host_en,
                                 void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(host_en && f && (np == 0 || p), "NULL argument");
  CHECK(np >= 0, "np < 0");
  return launch_energy_p(p, np, q, m, f, g, host_en, S(stream));
}

extern "C" int vpic_hip_center_p(vpic_particle_t *p, int32_t np, float q,
                                 float m, const vpic_interpolator_t *f,
                                 const vpic_grid_t *g, int32_t uncenter,
                                 void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(f && (np == 0 || p), "NULL argument");
  CHECK(np >= 0, "np < 0");
  return launch_center_p(p, np, q, m, f, g, uncenter != 0, S(stream));
}

extern "C" int vpic_hip_accumulate_rho_p(vpic_field_t *f,
                                         const vpic_particle_t *p, int32_t np,
                                         float q, const vpic_grid_t *g,
                                         void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(f && (np == 0 || p), "NULL argument");
  return launch_accumulate_rho_p(f, p, np, q, g, S(stream));
}

// ---------------------------------------------------------------------------
// field / particle interface and field advance
// ---------------------------------------------------------------------------
extern "C" int vpic_hip_load_interpolator_array(vpic_interpolator_t *fi,
                                                const vpic_field_t *f,
                                                const vpic_grid_t *g,
                                                void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(fi && f && aligned16(fi) && aligned16(f), "bad array argument");
  return launch_load_interpolator(fi, f, g, S(stream));
}

extern "C" int vpic_hip_clear_accumulator_array(vpic_accumulator_t *a,
                                                const vpic_grid_t *g,
                                                void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(a && aligned16(a), "bad accumulator array");
  return launch_clear_accumulator(a, g, S(stream));
}

extern "C" int vpic_hip_reduce_accumulator_array(vpic_accumulator_t *a,
                                                 const vpic_grid_t *g,
                                                 void *stream) {
  (void)stream;
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(a, "accumulator array is NULL");
  return VPIC_OK;
}

extern "C" int vpic_hip_unload_accumulator_array(vpic_field_t *f,
                                                 const vpic_accumulator_t *a,
                                                 const vpic_grid_t *g,
                                                 void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(f && a && aligned16(f), "bad array argument");
  return launch_unload_accumulator(f, a, g, S(stream));
}

extern "C" int vpic_hip_clear_jf_unload_accumulator_array(
    vpic_field_t *f, const vpic_accumulator_t *a, const vpic_grid_t *g,
    void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(f && a && aligned16(f), "bad array argument");
  return launch_clear_unload_accumulator(f, a, g, S(stream));
}

extern "C" int vpic_hip_clear_jf(vpic_field_t *f, const vpic_grid_t *g,
                                 void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(f, "field array is NULL");
  return launch_clear_jf(f, g, S(stream));
}

extern "C" int vpic_hip_clear_rhof(vpic_field_t *f, const vpic_grid_t *g,
                                   void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(f, "field array is NULL");
  return launch_clear_rhof(f, g, S(stream));
}

extern "C" int vpic_hip_synchronize_jf(vpic_field_t *f, const vpic_grid_t *g,
                                       void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(f, "field array is NULL");
  return launch_synchronize_jf(f, g, S(stream));
}

extern "C" int vpic_hip_synchronize_rho(vpic_field_t *f, const vpic_grid_t *g,
                                        void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(f, "field array is NULL");
  return launch_synchronize_rho(f, g, S(stream));
}

extern "C" int vpic_hip_advance_b(vpic_field_t *f, const vpic_grid_t *g,
                                  float frac, void *stream) {
  int rc = check_grid(g);
  if (rc || (rc = check_fields_supported(g))) return rc;
  CHECK(f && aligned16(f), "bad field array");
  return launch_advance_b(f, g, frac, S(stream));
}

extern "C" int vpic_hip_vacuum_advance_e(vpic_field_t *f, const vpic_grid_t *g,
                                         const vpic_field_params_t *mp,
                                         float frac, void *stream) {
  int rc = check_grid(g);
  if (rc || (rc = check_fields_supported(g))) return rc;
  CHECK(f && mp && aligned16(f), "bad argument");
  // vacuum_advance_e_pipeline.cc:58-61
  CHECK(frac == 1, "standard advance_e does not support frac != 1 yet");
  return launch_vacuum_advance_e(f, g, mp, S(stream));
}

extern "C" int vpic_hip_energy_f(double host_en[6], const vpic_field_t *f,
                                 const vpic_grid_t *g,
                                 const vpic_field_params_t *mp, void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(host_en && f && mp, "NULL argument");
  return launch_energy_f(host_en, f, g, mp, S(stream));
}

// Divergence cleaning and initialization kernels of the vacuum field table.
extern "C" int vpic_hip_compute_div_e_err(vpic_field_t *f, const vpic_grid_t *g,
                                          const vpic_field_params_t *mp,
                                          void *stream) {
  int rc = check_grid(g);
  if (rc || (rc = check_fields_supported(g))) return rc;
  CHECK(f && mp && aligned16(f), "bad argument");
  return launch_compute_div_e_err(f, g, mp, 0, S(stream));
}

extern "C" int vpic_hip_compute_rhob(vpic_field_t *f, const vpic_grid_t *g,
                                     const vpic_field_params_t *mp,
                                     void *stream) {
  int rc = check_grid(g);
  if (rc || (rc = check_fields_supported(g))) return rc;
  CHECK(f && mp && aligned16(f), "bad argument");
  return launch_compute_div_e_err(f, g, mp, 1, S(stream));
}

extern "C" int vpic_hip_compute_rms_div_e_err(double *host_rms,
                                              const vpic_field_t *f,
                                              const vpic_grid_t *g,
                                              void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(host_rms && f, "NULL argument");
  return launch_compute_rms_div_err(host_rms, f, g, 0, S(stream));
}

extern "C" int vpic_hip_clean_div_e(vpic_field_t *f, const vpic_grid_t *g,
                                    const vpic_field_params_t *mp,
                                    void *stream) {
  int rc = check_grid(g);
  if (rc || (rc = check_fields_supported(g))) return rc;
  CHECK(f && mp && aligned16(f), "bad argument");
  return launch_clean_div_e(f, g, mp, S(stream));
}

extern "C" int vpic_hip_compute_div_b_err(vpic_field_t *f, const vpic_grid_t *g,
                                          void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(f && aligned16(f), "bad argument");
  return launch_compute_div_b_err(f, g, S(stream));
}

extern "C" int vpic_hip_compute_rms_div_b_err(double *host_rms,
                                              const vpic_field_t *f,
                                              const vpic_grid_t *g,
                                              void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(host_rms && f, "NULL argument");
  return launch_compute_rms_div_err(host_rms, f, g, 1, S(stream));
}

extern "C" int vpic_hip_clean_div_b(vpic_field_t *f, const vpic_grid_t *g,
                                    void *stream) {
  int rc = check_grid(g);
  if (rc || (rc = check_fields_supported(g))) return rc;
  CHECK(f && aligned16(f), "bad argument");
  return launch_clean_div_b(f, g, S(stream));
}

extern "C" int vpic_hip_compute_curl_b(vpic_field_t *f, const vpic_grid_t *g,
                                       const vpic_field_params_t *mp,
                                       void *stream) {
  int rc = check_grid(g);
  if (rc || (rc = check_fields_supported(g))) return rc;
  CHECK(f && mp && aligned16(f), "bad argument");
  return launch_compute_curl_b(f, g, mp, S(stream));
}

extern "C" int vpic_hip_synchronize_tang_e_norm_b(double *host_err,
                                                  vpic_field_t *f,
                                                  const vpic_grid_t *g,
                                                  void *stream) {
  int rc = check_grid(g);
  if (rc || (rc = check_fields_supported(g))) return rc;
  CHECK(host_err && f && aligned16(f), "bad argument");
  return launch_synchronize_tang_e_norm_b(host_err, f, g, S(stream));
}

// Hydro moments.
extern "C" int vpic_hip_clear_hydro_array(vpic_hydro_t *h, const vpic_grid_t *g,
                                          void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(h && aligned16(h), "bad argument");
  return launch_clear_hydro(h, g, S(stream));
}

extern "C" int vpic_hip_accumulate_hydro_p(vpic_hydro_t *h,
                                           const vpic_particle_t *p,
                                           int32_t np, float q, float m,
                                           const vpic_interpolator_t *f,
                                           const vpic_grid_t *g,
                                           const int32_t *partition,
                                           void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(h && f && aligned16(h) && (np == 0 || p) && np >= 0,
        "Bad args.");
  CHECK(m != 0, "Bad args: m == 0");
  return launch_accumulate_hydro_p(h, p, np, q, m, f, g,
                                   (const int *)partition, S(stream));
}

extern "C" int vpic_hip_synchronize_hydro_array(vpic_hydro_t *h,
                                                const vpic_grid_t *g,
                                                void *stream) {
  int rc = check_grid(g);
  if (rc || (rc = check_fields_supported(g))) return rc;
  CHECK(h && aligned16(h), "NULL hydro array.");
  return launch_synchronize_hydro(h, g, S(stream));
}

// ---------------------------------------------------------------------------
// set-up helpers
// ---------------------------------------------------------------------------
extern "C" int vpic_hip_grid_periodic(vpic_grid_t *g, double gx0, double gy0,
                                      double gz0, double gx1, double gy1,
                                      double gz1, int32_t gnx, int32_t gny,
                                      int32_t gnz, float dt, float cvac,
                                      float eps0) {
  CHECK(g, "grid is NULL");
  CHECK(gnx >= 1 && gny >= 1 && gnz >= 1, "bad resolution");
  CHECK(gx1 > gx0 && gy1 > gy0 && gz1 > gz0, "bad extent");
  CHECK((int64_t)(gnx + 2) * (gny + 2) * (gnz + 2) < (int64_t)(1 << 28),
        "too many voxels for 8*voxel+face indexing");
  memset(g, 0, sizeof(*g));
  g->dt = dt;
  g->cvac = cvac;
  g->eps0 = eps0;
  // partition.cc:58-70
  g->dx = (float)((gx1 - gx0) / (double)gnx);
  g->dy = (float)((gy1 - gy0) / (double)gny);
  g->dz = (float)((gz1 - gz0) / (double)gnz);
  g->dV = (float)(((gx1 - gx0) / (double)gnx) * ((gy1 - gy0) / (double)gny) *
                  ((gz1 - gz0) / (double)gnz));
  g->rdx = (float)((double)gnx / (gx1 - gx0));
  g->rdy = (float)((double)gny / (gy1 - gy0));
  g->rdz = (float)((double)gnz / (gz1 - gz0));
  g->r8V = (float)(((double)gnx / (gx1 - gx0)) * ((double)gny / (gy1 - gy0)) *
                   ((double)gnz / (gz1 - gz0)) * 0.125);
  // ops.cc:25-30
  g->nx = gnx;
  g->ny = gny;
  g->nz = gnz;
  g->sx = 1;
  g->sy = gnx + 2;
  g->sz = (gny + 2) * g->sy;
  g->nv = (gnz + 2) * g->sz;
  g->rangel = 0;
  g->rangeh = g->nv - 1;
  g->neighbor = nullptr;
  for (int k = 0; k < 6; k++) {
    g->pbc[k] = VPIC_BC_PERIODIC;
    g->fbc[k] = VPIC_BC_PERIODIC;
  }
  return VPIC_OK;
}

extern "C" int vpic_hip_field_params_vacuum(vpic_field_params_t *mp,
                                            float damp) {
  CHECK(mp, "params is NULL");
  // sfa.cc:111-126 with sigma = 0, eps = mu = 1.
  mp->decayx = mp->decayy = mp->decayz = (float)std::exp(-0.0);
  mp->drivex = mp->drivey = mp->drivez = (float)(1. / 1.0);
  mp->rmux = mp->rmuy = mp->rmuz = (float)(1. / 1.0);
  mp->epsx = mp->epsy = mp->epsz = 1.0f;
  mp->damp = damp;
  mp->nonconductive = 1.0f;  // sfa.cc:144 with ax = ay = az = 0
  return VPIC_OK;
}

extern "C" int vpic_hip_copy(void *dst, const void *src, int64_t bytes,
                             void *stream) {
  CHECK(bytes >= 0 && (bytes == 0 || (dst && src)), "copy: bad arguments");
  if (!bytes) return VPIC_OK;
  VPIC_HIP(hipMemcpyAsync(dst, src, (size_t)bytes, hipMemcpyDefault,
                          S(stream)));
  VPIC_HIP(hipStreamSynchronize(S(stream)));
  return VPIC_OK;
}

extern "C" int vpic_hip_grid_neighbor_table(const vpic_grid_t *g,
                                            int64_t *nb) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(nb, "table is NULL");
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const int64_t rl = g->rangel;
  // ops.cc:45-76 then join_grid / set_pbc for the six domain faces.
  for (int z = 0; z <= nz + 1; z++)
    for (int y = 0; y <= ny + 1; y++)
      for (int x = 0; x <= nx + 1; x++) {
        int64_t *n = nb + 6 * (int64_t)vox(x, y, z, nx, ny);
        const bool ghost = x == 0 || x == nx + 1 || y == 0 || y == ny + 1 ||
                           z == 0 || z == nz + 1;
        if (ghost) {
          for (int k = 0; k < 6; k++) n[k] = VPIC_BC_REFLECT_PARTICLES;
          continue;
        }
        n[0] = x == 1 ? (g->pbc[0] == VPIC_BC_PERIODIC ? rl + vox(nx, y, z, nx, ny) : g->pbc[0])
                      : rl + vox(x - 1, y, z, nx, ny);
        n[1] = y == 1 ? (g->pbc[1] == VPIC_BC_PERIODIC ? rl + vox(x, ny, z, nx, ny) : g->pbc[1])
                      : rl + vox(x, y - 1, z, nx, ny);
        n[2] = z == 1 ? (g->pbc[2] == VPIC_BC_PERIODIC ? rl + vox(x, y, nz, nx, ny) : g->pbc[2])
                      : rl + vox(x, y, z - 1, nx, ny);
        n[3] = x == nx ? (g->pbc[3] == VPIC_BC_PERIODIC ? rl + vox(1, y, z, nx, ny) : g->pbc[3])
                       : rl + vox(x + 1, y, z, nx, ny);
        n[4] = y == ny ? (g->pbc[4] == VPIC_BC_PERIODIC ? rl + vox(x, 1, z, nx, ny) : g->pbc[4])
                       : rl + vox(x, y + 1, z, nx, ny);
        n[5] = z == nz ? (g->pbc[5] == VPIC_BC_PERIODIC ? rl + vox(x, y, 1, nx, ny) : g->pbc[5])
                       : rl + vox(x, y, z + 1, nx, ny);
      }
  return VPIC_OK;
}

extern "C" int vpic_hip_load_uniform_plasma(vpic_particle_t *p,
                                            const vpic_grid_t *g, int32_t ppc,
                                            float uth_x, float uth_y,
                                            float uth_z, float ud_x,
                                            float ud_y, float ud_z, float w,
                                            uint64_t seed, void *stream) {
  int rc = check_grid(g);
  if (rc) return rc;
  CHECK(p && aligned16(p), "bad particle array");
  CHECK(ppc >= 1, "ppc < 1");
  CHECK((int64_t)ppc * g->nx * g->ny * g->nz < (int64_t)INT32_MAX,
        "too many particles for int32 indexing");
  return launch_load_uniform_plasma(p, g, ppc, uth_x, uth_y, uth_z, ud_x,
                                    ud_y, ud_z, w, seed, S(stream));
}
