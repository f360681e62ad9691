// common.h -- shared device/host definitions of libvpichip (gfx950).
//
// Compiled with -ffp-contract=off and IEEE-correct division and sqrt (the
// HIP defaults), no fast-math: every FP32 expression of the particle push
// and the field stencils rounds exactly as the reference's scalar build.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/vpic_hip.h"

#define VPIC_WAVE 64

// Voxel index, src/grid/grid.h:136.
__host__ __device__ inline int vox(int x, int y, int z, int nx, int ny) {
  return x + (nx + 2) * (y + (ny + 2) * z);
}

// Grid data every kernel needs, passed by value.
struct GridDev {
  int nx, ny, nz, sy, sz, nv;
  int64_t rangel, rangeh;
  const int64_t *neighbor;
  int pbc[6];
  int fbc[6];
};

inline GridDev make_grid_dev(const vpic_grid_t *g) {
  GridDev d;
  d.nx = g->nx;
  d.ny = g->ny;
  d.nz = g->nz;
  d.sy = g->sy;
  d.sz = g->sz;
  d.nv = g->nv;
  d.rangel = g->rangel;
  d.rangeh = g->rangeh;
  d.neighbor = g->neighbor;
  for (int k = 0; k < 6; k++) {
    d.pbc[k] = g->pbc[k];
    d.fbc[k] = g->fbc[k];
  }
  return d;
}

// Neighbour of voxel v across face (order -x,-y,-z,+x,+y,+z) as the
// reference's table holds it (src/grid/ops.cc:45-76 + join_grid/set_pbc):
// the explicit table when given, else derived from the coordinates.
__device__ inline int64_t grid_neighbor(const GridDev &G, int v, int face) {
  if (G.neighbor) return G.neighbor[6 * (int64_t)v + face];
  const int sx1 = G.nx + 2, sy1 = G.ny + 2;
  int x = v % sx1;
  int t = v / sx1;
  int y = t % sy1;
  int z = t / sy1;
  if (x == 0 || x == G.nx + 1 || y == 0 || y == G.ny + 1 || z == 0 ||
      z == G.nz + 1)
    return VPIC_BC_REFLECT_PARTICLES;
  switch (face) {
    case 0:
      if (x == 1)
        return G.pbc[0] == VPIC_BC_PERIODIC ? G.rangel + (v + G.nx - 1)
                                            : (int64_t)G.pbc[0];
      return G.rangel + (v - 1);
    case 1:
      if (y == 1)
        return G.pbc[1] == VPIC_BC_PERIODIC
                   ? G.rangel + (v + (int64_t)(G.ny - 1) * G.sy)
                   : (int64_t)G.pbc[1];
      return G.rangel + (v - G.sy);
    case 2:
      if (z == 1)
        return G.pbc[2] == VPIC_BC_PERIODIC
                   ? G.rangel + (v + (int64_t)(G.nz - 1) * G.sz)
                   : (int64_t)G.pbc[2];
      return G.rangel + (v - G.sz);
    case 3:
      if (x == G.nx)
        return G.pbc[3] == VPIC_BC_PERIODIC ? G.rangel + (v - (G.nx - 1))
                                            : (int64_t)G.pbc[3];
      return G.rangel + (v + 1);
    case 4:
      if (y == G.ny)
        return G.pbc[4] == VPIC_BC_PERIODIC
                   ? G.rangel + (v - (int64_t)(G.ny - 1) * G.sy)
                   : (int64_t)G.pbc[4];
      return G.rangel + (v + G.sy);
    default:
      if (z == G.nz)
        return G.pbc[5] == VPIC_BC_PERIODIC
                   ? G.rangel + (v - (int64_t)(G.nz - 1) * G.sz)
                   : (int64_t)G.pbc[5];
      return G.rangel + (v + G.sz);
  }
}

// Hardware float atomic add (global_atomic_add_f32, no return).
__device__ inline void atomic_add_f32(float *a, float v) {
  __hip_atomic_fetch_add(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline void atomic_add_f64(double *a, double v) {
  __hip_atomic_fetch_add(a, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ inline int lane_id() { return __lane_id(); }
// Number of bits of the 64-bit mask m set below this lane (v_mbcnt: no
// lane index or lane mask kept in registers).
__device__ __forceinline__ int lanes_below(uint64_t m) {
  return (int)__builtin_amdgcn_mbcnt_hi(
      (unsigned)(m >> 32), __builtin_amdgcn_mbcnt_lo((unsigned)m, 0u));
}

// ---------------------------------------------------------------------------
// Host-side helpers shared by the launchers (defined in api.hip)
// ---------------------------------------------------------------------------

// Record an error message for vpic_hip_last_error and return `code`.
int vpic_set_error(int code, const char *fmt, ...);
// Check a HIP status; on failure record it and return a negative code.
int vpic_check_hip(hipError_t e, const char *what);

#define VPIC_HIP(call)                                   \
  do {                                                   \
    hipError_t e_ = (call);                              \
    if (e_ != hipSuccess) return vpic_check_hip(e_, #call); \
  } while (0)

#define VPIC_LAUNCHED(name)                                   \
  do {                                                        \
    hipError_t e_ = hipGetLastError();                        \
    if (e_ != hipSuccess) return vpic_check_hip(e_, name);    \
  } while (0)

enum {
  VPIC_OK = 0,
  VPIC_EINVAL = -1,
  VPIC_EHIP = -2,
  VPIC_ENOMEM = -3,
  VPIC_EUNSUPPORTED = -4
};

// vacuum_advance_e couplings (vacuum_advance_e_pipeline.h:27-33).
struct ECoef {
  float decayx, drivex, decayy, drivey, decayz, drivez, damp;
  float px_muz, px_muy, py_mux, py_muz, pz_muy, pz_mux, cj;
};

// A box of a batched ensemble (packed arrays: its voxels start at voff).
struct EnsFieldBox {
  int nx, ny, nz;
  int64_t voff;
  float bp[3];  // advance_b(0.5) couplings
  float uc[3];  // unload couplings
  ECoef ec;     // vacuum_advance_e couplings
};

// A work item of a batched phase: voxels (axis < 0), face elements of one
// axis, or particles of one box; linear indices [first, first + count).
struct EnsItem {
  int box, axis, first, count;
};

enum { ENS_UNLOAD, ENS_ADVB, ENS_ADVE, ENS_LOAD, ENS_GHOST, ENS_FOLD };

// A box of a batched ensemble as the push sees it.
struct EnsPushBox {
  float qdt_2mc, cdt_dx, cdt_dy, cdt_dz, qsp;
  GridDev G;
  int64_t poff, voff;
};

void ens_make_field_box(EnsFieldBox *b, const vpic_grid_t *g,
                        const vpic_field_params_t *m, int64_t voff);
int launch_ens_field(int phase, vpic_field_t *F, vpic_interpolator_t *FI,
                     const vpic_accumulator_t *A, const EnsFieldBox *boxes,
                     const EnsItem *items, int nitems, int *queue,
                     int nblocks, int bs, hipStream_t s);
int launch_ens_reset(int *queues, int n, hipStream_t s);
// The tiled push of a batched ensemble (push.hip): per-box tiles of
// tx x ty x tz cells (clamped to each box), their LDS windows and
// out-of-window sub-buckets, and the box-relative partitions (box b's at
// ens_tiled_epart() + voff_b + b, nv_b + 1 ints, written by the batched
// sort).  launch_ens_push_tiled enqueues the push and the merge of every
// box tile; its counters (4 ints: [1] movers, [3] particles off their box)
// accumulate until the caller clears them.
struct EnsTiled;
int ens_tiled_create(EnsTiled **out, int nbox, const vpic_grid_t *grids,
                     const int *np, const float *q, const float *m,
                     const int64_t *voff, const int64_t *poff,
                     vpic_particle_t *P, const vpic_interpolator_t *FI,
                     vpic_accumulator_t *A, int tx, int ty, int tz,
                     int halo);
// Largest window halo of an ensemble's tiles (its boxes are often 2D, where
// a wide halo costs little LDS).
constexpr int ENS_MAX_HALO = 3;
void ens_tiled_destroy(EnsTiled *t);
int *ens_tiled_epart(EnsTiled *t);
int *ens_tiled_counters(EnsTiled *t);
int launch_ens_push_tiled(EnsTiled *et, int *queue, int nblocks,
                          hipStream_t s);
// Sort of every box's particles by voxel (sort_p per box on the packed
// arrays), enqueued without a host synchronisation so that it can be
// captured: keys (np_total ints), count and start (nv_total + 1 ints each),
// sums (device_exclusive_scan's), aux (np_total particles).  Equal voxels
// keep no particular order.  With epart, box b's partition (box-relative,
// nv_b + 1 ints) is written at epart + voff_b + b.
int launch_ens_sort(vpic_particle_t *P, vpic_particle_t *aux,
                    const EnsPushBox *boxes, const EnsItem *items, int nitems,
                    int nbox, int64_t np_total, int64_t nv_total, int *keys,
                    int *count, int *start, int *sums, int *nbad, int *epart,
                    hipStream_t s);
int launch_clear_accumulator_n(vpic_accumulator_t *a, int64_t nv,
                               hipStream_t s);
int launch_ens_push(vpic_particle_t *P, const vpic_interpolator_t *FI,
                    vpic_accumulator_t *A, const EnsPushBox *boxes,
                    const EnsItem *items, int nitems, int *queue, int *nbad,
                    int nblocks, int bs, hipStream_t s);

// Per-device scratch kept by the library (grown on demand).
struct DeviceScratch {
  void *ptr = nullptr;
  size_t bytes = 0;
};
int scratch_reserve(DeviceScratch &s, size_t bytes);

// Launch-side kernels entry points implemented in the .hip files.
// advance_p of several species into one accumulator (one tiled launch when
// all have partitions).
int launch_advance_p_set(int nsp, const vpic_push_species_t *sp,
                         vpic_accumulator_t *a, const vpic_interpolator_t *f,
                         const vpic_grid_t *g, hipStream_t s);
// keys: when not null, keys[k] = p[k].i as the push leaves it, for a
// following launch_sort_p_keys (-1 for a mover returned to the caller).
int launch_advance_p(vpic_particle_t *p, int np, float q, float m,
                     vpic_particle_mover_t *pm, int max_nm, int *dev_nm,
                     vpic_accumulator_t *a, const vpic_interpolator_t *f,
                     const vpic_grid_t *g, const int *partition, int *keys,
                     hipStream_t s);
// The flat push on the library's fixed per-device scratch (no allocation):
// vpic_hip_advance_p_noalloc.  push_noalloc_setup looks up that scratch's
// device address on the current device (once per device).
int launch_advance_p_noalloc(vpic_particle_t *p, int np, float q, float m,
                             vpic_particle_mover_t *pm, int max_nm,
                             int *dev_nm, vpic_accumulator_t *a,
                             const vpic_interpolator_t *f,
                             const vpic_grid_t *g, hipStream_t s);
int push_noalloc_setup();
int push_set_tile(int tx, int ty, int tz, int halo);
void push_get_tile(int out[4]);
int push_set_direct_runs(int runs);
int push_set_group(int gs);
int push_set_bin_mode(int mode);
// Forget the tile lists of a particle array whose order or voxels changed
// wholesale (sort_p, loading).
void push_invalidate_bins(const void *p);
// Exclusive scan of n ints (in may equal out).  sums: ceil(n/2048)+1 ints.
int device_exclusive_scan(const int *in, int *out, int64_t n, int *sums,
                          hipStream_t s);
int launch_move_p_list(vpic_particle_t *p, vpic_particle_mover_t *pm, int nm,
                       vpic_accumulator_t *a, const vpic_grid_t *g, float q,
                       int *dev_ret, hipStream_t s);
int launch_boundary_p(vpic_particle_t *p, int *np_host,
                      const vpic_particle_mover_t *pm, int nm,
                      vpic_field_t *f, float q, const vpic_grid_t *g,
                      vpic_boundary_hit_t *hits, int *counts_host,
                      hipStream_t s);
int launch_takizuka_abe(vpic_particle_t *pi, const int *part_i, int np_i,
                        float qi, float mi, vpic_particle_t *pj,
                        const int *part_j, int np_j, float qj, float mj,
                        double cvar0, int interval, uint64_t seed,
                        uint64_t call, const vpic_grid_t *g, hipStream_t s);
int launch_accumulate_rhob(vpic_field_t *f, const vpic_particle_t *p, int n,
                           float qsp, const vpic_grid_t *g, int *dev_bad,
                           hipStream_t s);
int launch_inject_p(vpic_particle_t *p, int *np, int max_np,
                    vpic_particle_mover_t *pm, int *nm, int max_nm,
                    const vpic_particle_injector_t *pi, int n,
                    vpic_accumulator_t *a, float q, const vpic_grid_t *g,
                    int *counts, hipStream_t s);
int launch_sort_p(vpic_particle_t *p, vpic_particle_t *aux, int np,
                  int *partition, const vpic_grid_t *g, int in_place,
                  hipStream_t s);
// sort_p from the caller's copy of the voxel keys, checked against the
// particles (any difference: sorted from the particles instead, and
// counted in sort_keys_stale_count).
int launch_sort_p_keys(vpic_particle_t *p, vpic_particle_t *aux, int np,
                       int *partition, const int *keys, const vpic_grid_t *g,
                       int in_place, hipStream_t s);
long long sort_keys_stale_count();
int launch_energy_p(const vpic_particle_t *p, int np, float q, float m,
                    const vpic_interpolator_t *f, const vpic_grid_t *g,
                    double *host_en, hipStream_t s);
int launch_accumulate_rho_p(vpic_field_t *f, const vpic_particle_t *p, int np,
                            float q, const vpic_grid_t *g, hipStream_t s);
int launch_center_p(vpic_particle_t *p, int np, float q, float m,
                    const vpic_interpolator_t *f, const vpic_grid_t *g,
                    int uncenter, hipStream_t s);
int launch_load_interpolator(vpic_interpolator_t *fi, const vpic_field_t *f,
                             const vpic_grid_t *g, hipStream_t s);
int launch_clear_accumulator(vpic_accumulator_t *a, const vpic_grid_t *g,
                             hipStream_t s);
int launch_unload_accumulator(vpic_field_t *f, const vpic_accumulator_t *a,
                              const vpic_grid_t *g, hipStream_t s);
int launch_clear_jf(vpic_field_t *f, const vpic_grid_t *g, hipStream_t s);
int launch_clear_unload_accumulator(vpic_field_t *f,
                                    const vpic_accumulator_t *a,
                                    const vpic_grid_t *g, hipStream_t s);
int launch_clear_rhof(vpic_field_t *f, const vpic_grid_t *g, hipStream_t s);
int launch_synchronize_jf(vpic_field_t *f, const vpic_grid_t *g,
                          hipStream_t s);
int launch_synchronize_rho(vpic_field_t *f, const vpic_grid_t *g,
                           hipStream_t s);
int launch_advance_b(vpic_field_t *f, const vpic_grid_t *g, float frac,
                     hipStream_t s);
int launch_vacuum_advance_e(vpic_field_t *f, const vpic_grid_t *g,
                            const vpic_field_params_t *mp, hipStream_t s);
int launch_energy_f(double en[6], const vpic_field_t *f, const vpic_grid_t *g,
                    const vpic_field_params_t *mp, hipStream_t s);
int launch_compute_div_e_err(vpic_field_t *f, const vpic_grid_t *g,
                             const vpic_field_params_t *m, int rhob,
                             hipStream_t s);
int launch_compute_rms_div_err(double *out, const vpic_field_t *f,
                               const vpic_grid_t *g, int divb, hipStream_t s);
int launch_clean_div_e(vpic_field_t *f, const vpic_grid_t *g,
                       const vpic_field_params_t *m, hipStream_t s);
int launch_compute_div_b_err(vpic_field_t *f, const vpic_grid_t *g,
                             hipStream_t s);
int launch_clean_div_b(vpic_field_t *f, const vpic_grid_t *g, hipStream_t s);
int launch_compute_curl_b(vpic_field_t *f, const vpic_grid_t *g,
                          const vpic_field_params_t *m, hipStream_t s);
int launch_synchronize_tang_e_norm_b(double *out, vpic_field_t *f,
                                     const vpic_grid_t *g, hipStream_t s);
int launch_clear_hydro(vpic_hydro_t *h, const vpic_grid_t *g, hipStream_t s);
int launch_accumulate_hydro_p(vpic_hydro_t *h, const vpic_particle_t *p,
                              int np, float q, float m,
                              const vpic_interpolator_t *f,
                              const vpic_grid_t *g, const int *partition,
                              hipStream_t s);
int launch_synchronize_hydro(vpic_hydro_t *h, const vpic_grid_t *g,
                             hipStream_t s);
int launch_load_uniform_plasma(vpic_particle_t *p, const vpic_grid_t *g,
                               int ppc, float uthx, float uthy, float uthz,
                               float udx, float udy, float udz, float w,
                               uint64_t seed, hipStream_t s);

// Push-kernel timing hooks (api.hip).
bool push_timing_on();
void push_timing_record(hipStream_t s, int np, bool begin);
