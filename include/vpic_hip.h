/*
 * vpic_hip.h -- C-ABI of libvpichip, the MI355X (gfx950) implementation of
 * VPIC's per-step particle-in-cell hot path.
 *
 * Every entry point is `extern "C"`, takes plain pointers and sizes, and
 * returns an int status (0 = ok, < 0 = error; vpic_hip_last_error() gives
 * the message).  Array arguments are DEVICE pointers (HBM) unless the name
 * says `host_`.  `stream` is a hipStream_t passed as void* (NULL = the
 * default stream).  Calls are stream-ordered; a call returns a host-visible
 * scalar (nm, energies) only after synchronising its stream.
 *
 * The data types below are byte-for-byte the reference's, so a reference
 * array can be copied to HBM and handed over without repacking.
 * Citations are paths in lanl/vpic (legacy), relative to its root.
 */
#ifndef VPIC_HIP_H
#define VPIC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------ */
/* Data types (layouts fixed by the reference)                              */
/* ------------------------------------------------------------------------ */

/* src/species_advance/species_advance_aos.h:21-32 (32 B).  dx,dy,dz on
 * [-1,1] in cell coordinates, i = local voxel index (8*voxel+face while a
 * particle waits for boundary handling). */
typedef struct vpic_particle {
  float dx, dy, dz;
  int32_t i;
  float ux, uy, uz;
  float w;
} vpic_particle_t;

/* src/species_advance/species_advance_aos.h:37-40 (16 B).  Remaining
 * displacement in half-cell units and the index of the particle. */
typedef struct vpic_particle_mover {
  float dispx, dispy, dispz;
  int32_t i;
} vpic_particle_mover_t;

/* src/species_advance/species_advance_aos.h:45-52 (48 B): a particle to
 * add to species sp_id, with the displacement it still has to move. */
typedef struct vpic_particle_injector {
  float dx, dy, dz;
  int32_t i;
  float ux, uy, uz, w;
  float dispx, dispy, dispz;
  int32_t sp_id;
} vpic_particle_injector_t;

/* What boundary_p found for one mover (64 B): the particle as the boundary
 * handler sees it (voxel index restored, boundary_p.cc:257-266), the mover,
 * the face of the voxel it hit, and the boundary's code (the neighbor
 * table entry of that face: VPIC_BC_ABSORB_PARTICLES, a custom particle
 * boundary <= -3, or VPIC_BOUNDARY_UNKNOWN). */
typedef struct vpic_boundary_hit {
  vpic_particle_t p;
  vpic_particle_mover_t m;
  int32_t face;
  int32_t code;
  int32_t pad_[2];
} vpic_boundary_hit_t;

/* src/sf_interface/sf_interface.h:62-74, scalar build (PAD_SIZE 2, 80 B). */
typedef struct vpic_interpolator {
  float ex, dexdy, dexdz, d2exdydz;
  float ey, deydz, deydx, d2eydzdx;
  float ez, dezdx, dezdy, d2ezdxdy;
  float cbx, dcbxdx;
  float cby, dcbydy;
  float cbz, dcbzdz;
  float _pad[2];
} vpic_interpolator_t;

/* src/sf_interface/sf_interface.h:115-123 (48 B): 4x the charge through
 * each quarter face per step. */
typedef struct vpic_accumulator {
  float jx[4];
  float jy[4];
  float jz[4];
} vpic_accumulator_t;

/* src/field_advance/field_advance.h:152-160 (80 B). */
typedef struct vpic_field {
  float ex, ey, ez, div_e_err;
  float cbx, cby, cbz, div_b_err;
  float tcax, tcay, tcaz, rhob;
  float jfx, jfy, jfz, rhof;
  int16_t ematx, ematy, ematz, nmat;
  int16_t fmatx, fmaty, fmatz, cmat;
} vpic_field_t;

/* Boundary codes, src/grid/grid.h:14-38. */
enum {
  VPIC_BC_ANTI_SYMMETRIC_FIELDS = -1, /* pec / metal: E_tang = 0 */
  VPIC_BC_SYMMETRIC_FIELDS      = -2, /* B_tang = 0, B_norm = 0 */
  VPIC_BC_PMC_FIELDS            = -3, /* B_tang = 0 */
  VPIC_BC_ABSORB_FIELDS         = -4, /* not supported by the device path */
  VPIC_BC_REFLECT_PARTICLES     = -1,
  VPIC_BC_ABSORB_PARTICLES      = -2,
  /* Face partner is this same domain (one-rank periodic wrap). */
  VPIC_BC_PERIODIC              = 0,
  /* vpic_boundary_hit_t.code of a mover whose face leads to a voxel of
   * another domain (boundary_p.cc:277-309): no such face on one domain. */
  VPIC_BOUNDARY_UNKNOWN         = 1
};

/* The local grid, restating the fields of grid_t (src/grid/grid.h:73-131)
 * that the path reads.  `neighbor` is an optional DEVICE copy of the
 * reference's 6*nv int64 neighbor table (src/grid/ops.cc:45-76); when it
 * is NULL the kernels derive neighbours from the voxel coordinates and
 * `pbc` (particle BC per domain face, order -x,-y,-z,+x,+y,+z), which
 * reproduces size_grid + join_grid for one rank (src/grid/ops.cc:19-165,
 * src/grid/partition.cc:35-90).  `fbc` gives the field BC of each domain
 * face in the same order (VPIC_BC_PERIODIC or a local field code). */
typedef struct vpic_grid {
  float dt, cvac, eps0;
  float dx, dy, dz, dV;
  float rdx, rdy, rdz, r8V;
  int32_t nx, ny, nz;
  int32_t sx, sy, sz, nv;
  int64_t rangel, rangeh;
  const int64_t *neighbor;  /* device, 6*nv, or NULL */
  int32_t pbc[6];
  int32_t fbc[6];
} vpic_grid_t;

/* hydro_t (src/sf_interface/sf_interface.h:185-192): 16 floats, 64 B. */
typedef struct vpic_hydro {
  float jx, jy, jz, rho;
  float px, py, pz, ke;
  float txx, tyy, tzz;
  float tyz, tzx, txy;
  float pad_[2];
} vpic_hydro_t;

/* Field-solve material coefficients for the one vacuum material and the
 * TCA damping, src/field_advance/standard/sfa_private.h:14-30 and
 * src/field_advance/standard/sfa.cc:100-136. */
typedef struct vpic_field_params {
  float decayx, drivex, decayy, drivey, decayz, drivez;
  float rmux, rmuy, rmuz;
  float epsx, epsy, epsz;
  float damp;
  float nonconductive; /* 1 when sigma = 0 (divergence cleaning weight) */
} vpic_field_params_t;

/* ------------------------------------------------------------------------ */
/* Library state                                                           */
/* ------------------------------------------------------------------------ */

/* Select and initialise device `device` for this thread.  Idempotent. */
int vpic_hip_init(int device);

/* Message of the last failing call on this thread ("" if none). */
const char *vpic_hip_last_error(void);

/* Library version string. */
const char *vpic_hip_version(void);

/* ------------------------------------------------------------------------ */
/* Species advance (src/species_advance/species_advance.h:64-157)          */
/* ------------------------------------------------------------------------ */

/* advance_p: relativistic Boris push with energy-conserving interpolation,
 * in-cell Villasenor-Buneman deposit and move_p for cell crossers.
 * Replaces advance_p(species_t*, accumulator_array_t*, const
 * interpolator_array_t*) (species_advance.h:74-76; advance_p.cc:12-19 ->
 * advance_p_pipeline.cc:253-340).  q and m are the species charge and
 * mass; the coupling constants are formed in float exactly as
 * advance_p_pipeline.cc:279-283.  Currents are added into `a` (one
 * accumulator copy, nv entries).  Particles that hit a face the grid
 * cannot handle are returned as movers in `pm` in particle-index order
 * (up to max_nm, the rest rolled back as the reference does); `*host_nm`
 * receives their count and `*host_n_ignored` (may be NULL) the number
 * that did not fit.  `partition` (may be NULL) is the nv+1 voxel partition
 * written by the last vpic_hip_sort_p of these particles (species_t's
 * partition, species_advance_aos.h:69-89): when given, the particles are
 * pushed tile by tile with LDS-resident current accumulation (any
 * particles past partition[nv] are pushed too).  It changes only the
 * speed, never the result beyond the summation order of the currents.
 * Synchronous on return. */
int vpic_hip_advance_p(vpic_particle_t *p, int32_t np, float q, float m,
                       vpic_particle_mover_t *pm, int32_t max_nm,
                       int32_t *host_nm, int32_t *host_n_ignored,
                       vpic_accumulator_t *a, const vpic_interpolator_t *f,
                       const vpic_grid_t *g, const int32_t *partition,
                       void *stream);

/* Stream-ordered form of vpic_hip_advance_p: writes {nm, n_ignored,
 * n_bad} to the three device ints at `dev_nm` instead of the host and does
 * not synchronise.  n_bad counts particles whose voxel index lies outside
 * [0,nv); they are left untouched (the synchronous form reports them as an
 * error). */
int vpic_hip_advance_p_async(vpic_particle_t *p, int32_t np, float q,
                             float m, vpic_particle_mover_t *pm,
                             int32_t max_nm, int32_t *dev_nm,
                             vpic_accumulator_t *a,
                             const vpic_interpolator_t *f,
                             const vpic_grid_t *g, const int32_t *partition,
                             void *stream);

/* vpic_hip_advance_p_async that also writes keys[k] = p[k].i as the push
 * leaves it (k < np; -1 for a mover returned to the caller, whose index
 * the caller rewrites; keys NULL: no keys), for a vpic_hip_sort_p_keys of
 * these particles at the start of the next step (advance.cc:25-29 sorts
 * there): the sort then reads 4 B per particle instead of the 32-B
 * records.  No reference
 * counterpart (the reference's push and sort share no state). */
int vpic_hip_advance_p_keys(vpic_particle_t *p, int32_t np, float q, float m,
                            vpic_particle_mover_t *pm, int32_t max_nm,
                            int32_t *dev_nm, vpic_accumulator_t *a,
                            const vpic_interpolator_t *f,
                            const vpic_grid_t *g, const int32_t *partition,
                            int32_t *keys, void *stream);

/* One species' arguments of vpic_hip_advance_p_species (those of
 * vpic_hip_advance_p_keys; keys may be NULL). */
typedef struct vpic_push_species {
  vpic_particle_t *p;
  int32_t np;
  float q, m;
  vpic_particle_mover_t *pm;
  int32_t max_nm;
  int32_t *dev_nm;
  const int32_t *partition;
  int32_t *keys;
} vpic_push_species_t;

/* advance_p of nsp species (1 or 2) into one accumulator: the same as
 * vpic_hip_advance_p_keys for each in turn (advance.cc:56-62 loops
 * advance_p over the species list), except for the summation order of
 * the currents.  When every species has a partition, one tiled launch
 * pushes them all, tile by tile, into the same LDS windows with the same
 * staged coefficients, and one merge adds the windows.  No reference
 * counterpart. */
int vpic_hip_advance_p_species(int32_t nsp, const vpic_push_species_t *sp,
                               vpic_accumulator_t *a,
                               const vpic_interpolator_t *f,
                               const vpic_grid_t *g, void *stream);

/* vpic_hip_advance_p_async without a partition (the flat push), for
 * callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  No
 * counterpart in the reference.  Its crosser list, returned-mover list,
 * counters and ordering buffer live in one fixed array of the library
 * (2.5 MiB per GPU, held until the process ends; its device address is
 * looked up by vpic_hip_init, or by the first call on a device), so it
 * calls no hipMalloc, hipFree or allocating library routine.  Capacity:
 * 65536 crossers listed per call (the rest are moved inside the push, as
 * vpic_hip_advance_p_async does past its own list) and 65536 returned
 * movers; max_nm above 65536 is rejected.  For np <= 327680 its lists are
 * exactly as large as vpic_hip_advance_p_async's, and its results equal
 * that call's (bit for bit except the summation order of the currents).
 * Every call on one GPU shares that GPU's array: calls that overlap on one
 * GPU (on different streams or threads) race.  Writes {nm, n_ignored,
 * n_bad} to dev_nm; does not synchronise. */
int vpic_hip_advance_p_noalloc(vpic_particle_t *p, int32_t np, float q,
                               float m, vpic_particle_mover_t *pm,
                               int32_t max_nm, int32_t *dev_nm,
                               vpic_accumulator_t *a,
                               const vpic_interpolator_t *f,
                               const vpic_grid_t *g, void *stream);

/* Tile shape (cells) of the partition-driven push; its LDS window holds
 * (tx+2)(ty+2)(tz+2) accumulators (a halo of one cell).  Default 7x8x7.
 * No reference counterpart (the reference's pipelines split particles,
 * not cells). */
int vpic_hip_push_tile(int32_t tx, int32_t ty, int32_t tz);

/* The same with the window's halo: 1 or 2 cells, (tx+2h)(ty+2h)(tz+2h)
 * accumulators of 96 B, which with 72 B per tile cell must fit 140 KiB
 * of LDS.  Along an axis where the grid is one cell thick the tile is one
 * cell and the window has no halo (a 2D run's tiles can be 32x16x1); a
 * push on a grid where the window does not fit is an error.  A wider halo
 * keeps the deposits of particles that drifted further since the last
 * sort_p in LDS.  Results do not depend on it beyond the order of the
 * current sums.  No reference counterpart. */
int vpic_hip_push_tile_window(int32_t tx, int32_t ty, int32_t tz,
                              int32_t halo);

/* The tile shape and halo in use: out[0..3] = tx, ty, tz, halo.  No
 * reference counterpart. */
int vpic_hip_push_get_tile(int32_t *out);

/* Wavefronts whose 64 particles fall in at least `runs` runs of equal
 * voxels deposit lane by lane instead of summing runs first (default 65,
 * which disables it: on gfx950 the run-segmented scan was faster at every
 * threshold tried).  A tuning knob; no reference counterpart. */
int vpic_hip_push_direct_runs(int32_t runs);

/* Lanes that share one run of particles in the partition-driven push (1,
 * 2, 4, 8, 16, 32 or 64; default 16): a group reads its run's particles in
 * turn, so larger groups read wider contiguous spans.  A tuning knob; no
 * reference counterpart. */
int vpic_hip_push_group(int32_t gs);

/* Tile lists for the partition-driven push (default 0).  Each tiled push
 * can record which tile holds every particle it pushed; with lists on, the
 * next push of the same particle array gives each tile the particles now
 * in its cells, so particles that drifted since the sort still deposit
 * into their tile's LDS window.  0: off (always the sort-time partition;
 * out-of-window deposits go through the per-tile buckets, which on the
 * benchmark plasma costs less than recording and reading the lists); 1: on
 * once the last push made out-of-window deposits for at least 1/20 of its
 * particles, until the next sort_p of that array; 2: on whenever the last
 * push's records match.  Results do not depend on the mode.  A tuning
 * knob; no reference counterpart. */
int vpic_hip_push_bin_mode(int32_t mode);

/* move_p for one mover (species_advance.h:153-157; move_p.cc:215-378),
 * applied to the nm movers pm[0..nm) in order on the device (one lane, as
 * the reference's callers do, e.g. vpic/misc.cc:95).  `dev_ret` receives
 * one int per mover: 1 if the mover is still in use.  q = species charge. */
int vpic_hip_move_p(vpic_particle_t *p, vpic_particle_mover_t *pm,
                    int32_t nm, vpic_accumulator_t *a, const vpic_grid_t *g,
                    float q, int32_t *dev_ret, void *stream);

/* sort_p: stable counting sort of p[0..np) by voxel and the voxel
 * partition (species_advance.h:66; sort_p.cc:154-165 ->
 * sort_p_pipeline.cc:221-371).  `partition` receives nv+1 ints.  `aux`
 * is np particles of scratch; when `in_place` is 0 the sorted particles
 * are left in `aux` (the caller swaps the two arrays) and p is scratch. */
int vpic_hip_sort_p(vpic_particle_t *p, vpic_particle_t *aux, int32_t np,
                    int32_t *partition, const vpic_grid_t *g, int in_place,
                    void *stream);

/* vpic_hip_sort_p from the caller's copy of the voxel keys (keys[k] =
 * p[k].i, k < np, as vpic_hip_advance_p_keys writes them).  The result is
 * vpic_hip_sort_p's, whatever the keys hold: the sort compares every
 * particle's voxel index with its key, and when any differs (or a key lies
 * outside [0,nv)) it sorts again from the particles, leaving p as it was
 * until then.  Synchronises the stream once (twice on a fallback).  No
 * reference counterpart. */
int vpic_hip_sort_p_keys(vpic_particle_t *p, vpic_particle_t *aux,
                         int32_t np, int32_t *partition, const int32_t *keys,
                         const vpic_grid_t *g, int in_place, void *stream);

/* Number of vpic_hip_sort_p_keys calls in this process whose keys were not
 * the particles' own (each then sorted from the particles). */
int vpic_hip_sort_keys_stale(int64_t *count);

/* The Takizuka-Abe binary collision operator: replaces
 * apply_takizuka_abe_pipeline (src/collision/pipeline/
 * takizuka_abe_pipeline.cc:81-222) of takizuka_abe(name, spi, spj, rp,
 * cvar0, interval) (src/collision/takizuka_abe.cc:50-75).  Species i and j
 * must be sorted (their partitions from vpic_hip_sort_p), as
 * apply_takizuka_abe makes sure (takizuka_abe.cc:6-12); the caller applies
 * the step % interval test there.  pi == pj (with part_i == part_j) is the
 * intra-species operator.  Every voxel's particles of species i are
 * shuffled and paired with species j's as the reference does, with the
 * reference's collision arithmetic (its float sin and cos of the
 * scattering phase taken as the correctly rounded float values); the
 * random numbers come from a counter-based stream per voxel
 * (Philox4x32-10 with key `seed` and counter {draw/4, voxel, call}: pass a
 * new `call` each time, e.g. the step, and a different `seed` to each
 * operator applied at the same step, or their draws coincide) instead of
 * the reference's per-pipeline SFMT generator, so the result is
 * deterministic whatever the launch shape.  Synchronous. */
int vpic_hip_takizuka_abe(vpic_particle_t *pi, const int32_t *part_i,
                          int32_t np_i, float qi, float mi,
                          vpic_particle_t *pj, const int32_t *part_j,
                          int32_t np_j, float qj, float mj, double cvar0,
                          int32_t interval, uint64_t seed, uint64_t call,
                          const vpic_grid_t *g, void *stream);

/* accumulate_rhob (rho_p.cc:126-173; species_advance.h:133-136) of the n
 * particles p (device): each particle's charge q w on the 8 nodes of its
 * voxel into f[].rhob, a node weight doubled on each face of the local
 * domain the voxel touches -- what boundary_p does for an absorbed
 * particle, and vpic_simulation::advance for a mover it could not place
 * (advance.cc:78-101).  Synchronous (particles whose nodes are not on the
 * grid are skipped and reported). */
int vpic_hip_accumulate_rhob(vpic_field_t *f, const vpic_particle_t *p,
                             int32_t n, float qsp, const vpic_grid_t *g,
                             void *stream);

/* boundary_p for one species on a domain with no face shared with another
 * rank (boundary_p.cc:42-750; the loop of :257-375).  pm[0..nm) are the
 * movers advance_p left, in increasing particle order; each one's particle
 * has the index 8*voxel + face of the face it stopped on.  Taken last
 * first, as the reference does: the face's neighbor code decides
 *   VPIC_BC_ABSORB_PARTICLES  its charge q*w goes to f.rhob
 *                             (accumulate_rhob, rho_p.cc:126-173);
 *   a code <= -3              a custom particle boundary: the caller's
 *                             handler runs on hits[m] (host side);
 *   anything else             dropped (the reference warns, :348-352);
 * and every mover's particle is removed by moving the last particle into
 * its slot (:354-371), which this call does in parallel with the same
 * result.  hits[m] (nm entries, may be NULL) receives what mover m found.
 * *np is updated; counts[0] = custom hits, counts[1] = dropped (host
 * values).  Synchronous.  Particles are bit-identical to the reference;
 * rhob sums differ only in order. */
int vpic_hip_boundary_p(vpic_particle_t *p, int32_t *np,
                        const vpic_particle_mover_t *pm, int32_t nm,
                        vpic_field_t *f, float q, const vpic_grid_t *g,
                        vpic_boundary_hit_t *hits, int32_t *counts,
                        void *stream);

/* The injection loop of boundary_p (boundary_p.cc:595-711) for one
 * species: injectors pi[n-1] down to pi[0] are appended to p (p[np] =
 * particle, np += 1) and moved by their displacement with move_p, which
 * deposits into a; a particle that hits a boundary again leaves a mover
 * (pm[nm], nm += 1).  In order, on one lane.  An injector that finds p or
 * pm full is dropped and counted in counts[0] / counts[1] (the reference
 * built without dynamic resizing does the same, :652-698).  np, nm and
 * counts are host values.  Synchronous. */
int vpic_hip_inject_p(vpic_particle_t *p, int32_t *np, int32_t max_np,
                      vpic_particle_mover_t *pm, int32_t *nm, int32_t max_nm,
                      const vpic_particle_injector_t *pi, int32_t n,
                      vpic_accumulator_t *a, float q, const vpic_grid_t *g,
                      int32_t *counts, void *stream);

/* energy_p: kinetic energy of a species in FP64 (species_advance.h:119-120;
 * energy_p_pipeline.cc:18-115).  Result in *host_en.  Synchronous. */
int vpic_hip_energy_p(const vpic_particle_t *p, int32_t np, float q, float m,
                      const vpic_interpolator_t *f, const vpic_grid_t *g,
                      double *host_en, void *stream);

/* accumulate_rho_p: trilinear charge deposit into f.rhof
 * (species_advance.h; rho_p.cc:23-113). */
int vpic_hip_accumulate_rho_p(vpic_field_t *f, const vpic_particle_t *p,
                              int32_t np, float q, const vpic_grid_t *g,
                              void *stream);

/* center_p (uncenter = 0) / uncenter_p (uncenter = 1): half E kick and half
 * Boris rotation of the momenta, forward or backward in time; positions are
 * untouched.  Replaces center_p / uncenter_p(species_t*, const
 * interpolator_array_t*) (species_advance.h:83-110;
 * center_p_pipeline.cc:15-130, uncenter_p_pipeline.cc:15-130).
 * Synchronous (reports particles with a voxel index outside [0,nv)). */
int vpic_hip_center_p(vpic_particle_t *p, int32_t np, float q, float m,
                      const vpic_interpolator_t *f, const vpic_grid_t *g,
                      int32_t uncenter, void *stream);

/* ------------------------------------------------------------------------ */
/* Field/particle interface (src/sf_interface/sf_interface.h:82-176)        */
/* ------------------------------------------------------------------------ */

/* load_interpolator_array (sf_interface.h:100-101;
 * interpolator_array_pipeline.cc:21-135). */
int vpic_hip_load_interpolator_array(vpic_interpolator_t *fi,
                                     const vpic_field_t *f,
                                     const vpic_grid_t *g, void *stream);

/* clear_accumulator_array (sf_interface.h:148; clear_array_pipeline.cc:40-67):
 * zeroes all nv accumulators (ghosts included). */
int vpic_hip_clear_accumulator_array(vpic_accumulator_t *a,
                                     const vpic_grid_t *g, void *stream);

/* reduce_accumulator_array (sf_interface.h:159): the device keeps one
 * accumulator copy, so this is a checked no-op. */
int vpic_hip_reduce_accumulator_array(vpic_accumulator_t *a,
                                      const vpic_grid_t *g, void *stream);

/* unload_accumulator_array (sf_interface.h:173-174;
 * unload_accumulator_pipeline.cc:18-82,130-141): jf += c*(sum of quarter
 * face currents) for voxels 1..n+1. */
int vpic_hip_unload_accumulator_array(vpic_field_t *f,
                                      const vpic_accumulator_t *a,
                                      const vpic_grid_t *g, void *stream);

/* ------------------------------------------------------------------------ */
/* Field advance (src/field_advance/field_advance.h:170-218)               */
/* ------------------------------------------------------------------------ */

/* clear_jf (standard/sfa.cc:232-237). */
int vpic_hip_clear_jf(vpic_field_t *f, const vpic_grid_t *g, void *stream);

/* clear_jf followed by unload_accumulator_array, the pair advance.cc:107-109
 * calls, in one pass over the field records: the same jf bit for bit (the
 * unload's += on a cleared jf), rhof and every other field kept. */
int vpic_hip_clear_jf_unload_accumulator_array(vpic_field_t *f,
                                               const vpic_accumulator_t *a,
                                               const vpic_grid_t *g,
                                               void *stream);

/* clear_rhof. */
int vpic_hip_clear_rhof(vpic_field_t *f, const vpic_grid_t *g, void *stream);

/* synchronize_jf: local_adjust_jf then the shared-face fold
 * (standard/remote.cc:418-507, standard/local.cc:336-368). */
/* Waits for all the work queued on the current device (every stream):
 * for hosts that time or hand over results of the stream-ordered calls. */
int vpic_hip_device_synchronize(void);

int vpic_hip_synchronize_jf(vpic_field_t *f, const vpic_grid_t *g,
                            void *stream);

/* synchronize_rho: local_adjust_rhof/rhob then the shared-node fold of
 * rhof (sum) and rhob (average) (standard/remote.cc:534-625,
 * standard/local.cc:376-440).  Used with accumulate_rho_p for the charge
 * conservation diagnostic. */
int vpic_hip_synchronize_rho(vpic_field_t *f, const vpic_grid_t *g,
                             void *stream);

/* advance_b: cB -= frac*c*dt*curl E on all faces, then local_adjust_norm_b
 * (advance_b_pipeline.h:17-59; advance_b_pipeline.cc:20-125). */
int vpic_hip_advance_b(vpic_field_t *f, const vpic_grid_t *g, float frac,
                       void *stream);

/* vacuum_advance_e: tangential-B ghosts, E update on every edge, then
 * local_adjust_tang_e (vacuum_advance_e_pipeline.h:17-69;
 * vacuum_advance_e_pipeline.cc:20-332).  frac must be 1. */
int vpic_hip_vacuum_advance_e(vpic_field_t *f, const vpic_grid_t *g,
                              const vpic_field_params_t *mp, float frac,
                              void *stream);

/* energy_f: six FP64 field energies {ex,ey,ez,bx,by,bz}
 * (vacuum_energy_f_pipeline.h:17-75; vacuum_energy_f_pipeline.cc:13-97).
 * Synchronous. */
int vpic_hip_energy_f(double host_en[6], const vpic_field_t *f,
                      const vpic_grid_t *g, const vpic_field_params_t *mp,
                      void *stream);

/* Divergence cleaning and initialization kernels of the vacuum field table
 * (standard/sfa.cc:203-210; field_advance.h:199-216).  Field results are
 * bit-identical to the reference's; the FP64 sums (rms, err) are summed in
 * another order.  The rms and err entry points synchronise. */

/* vacuum_compute_div_e_err: normal-E ghosts, div_e_err on every node, then
 * local_adjust_div_e (vacuum_compute_div_e_err_pipeline.h:17-50,
 * .cc:61-174; standard/local.cc:129-179,299-334; remote.cc:137-207). */
int vpic_hip_compute_div_e_err(vpic_field_t *f, const vpic_grid_t *g,
                               const vpic_field_params_t *mp, void *stream);

/* compute_rms_div_e_err (compute_rms_div_e_err_pipeline.cc:16-185). */
int vpic_hip_compute_rms_div_e_err(double *host_rms, const vpic_field_t *f,
                                   const vpic_grid_t *g, void *stream);

/* vacuum_clean_div_e: Marder pass on every E edge, then
 * local_adjust_tang_e (vacuum_clean_div_e_pipeline.h:16-55, .cc:43-147). */
int vpic_hip_clean_div_e(vpic_field_t *f, const vpic_grid_t *g,
                         const vpic_field_params_t *mp, void *stream);

/* compute_div_b_err (compute_div_b_err_pipeline.cc:15-57,260-296). */
int vpic_hip_compute_div_b_err(vpic_field_t *f, const vpic_grid_t *g,
                               void *stream);

/* compute_rms_div_b_err (compute_rms_div_b_err_pipeline.cc:14-111). */
int vpic_hip_compute_rms_div_b_err(double *host_rms, const vpic_field_t *f,
                                   const vpic_grid_t *g, void *stream);

/* clean_div_b: div_b_err ghosts, Marder pass on every B face, then
 * local_adjust_norm_b (clean_div_b_pipeline.h:17-27, .cc:230-328;
 * standard/local.cc:182-221,266-296; remote.cc:209-281). */
int vpic_hip_clean_div_b(vpic_field_t *f, const vpic_grid_t *g, void *stream);

/* vacuum_compute_rhob: normal-E ghosts, rhob on every node, then
 * local_adjust_rhob (vacuum_compute_rhob_pipeline.h:17-52, .cc:41-175;
 * standard/local.cc:415-440). */
int vpic_hip_compute_rhob(vpic_field_t *f, const vpic_grid_t *g,
                          const vpic_field_params_t *mp, void *stream);

/* vacuum_compute_curl_b: tangential-B ghosts, tca = curl B on every edge,
 * then local_adjust_tang_e (vacuum_compute_curl_b_pipeline.h:17-66,
 * .cc:50-326). */
int vpic_hip_compute_curl_b(vpic_field_t *f, const vpic_grid_t *g,
                            const vpic_field_params_t *mp, void *stream);

/* synchronize_tang_e_norm_b for one domain: local adjusts, then each
 * periodic face pair averages cb normal, e and tca tangential; *host_err
 * receives the summed squared differences (standard/remote.cc:298-416). */
int vpic_hip_synchronize_tang_e_norm_b(double *host_err, vpic_field_t *f,
                                       const vpic_grid_t *g, void *stream);

/* ------------------------------------------------------------------------ */
/* Hydro moments (src/species_advance/standard/hydro_p.cc,                  */
/* src/sf_interface/hydro_array.cc)                                         */
/* ------------------------------------------------------------------------ */

/* clear_hydro_array (sf_interface/clear_array.cc:29-38): h[0..nv) = 0. */
int vpic_hip_clear_hydro_array(vpic_hydro_t *h, const vpic_grid_t *g,
                               void *stream);

/* accumulate_hydro_p (hydro_p.cc:12-19 -> hydro_p_pipeline.cc:17-242):
 * adds the 14 moments of p[0..np) to the nodes, trilinear weights.  With
 * `partition` (from the last vpic_hip_sort_p of these particles) the
 * tiles of the partition are summed in LDS, the particles past it one by
 * one; without it every particle one by one.  The per-particle arithmetic
 * is the reference's; the node sums differ only in order.  Synchronous;
 * particles whose voxel is not interior are skipped and reported as an
 * error.  reduce_hydro_array has nothing to do (one device copy). */
int vpic_hip_accumulate_hydro_p(vpic_hydro_t *h, const vpic_particle_t *p,
                                int32_t np, float q, float m,
                                const vpic_interpolator_t *f,
                                const vpic_grid_t *g, const int32_t *partition,
                                void *stream);

/* synchronize_hydro_array for one domain (hydro_array.cc:131-311): local
 * faces doubled, then the periodic node planes summed. */
int vpic_hip_synchronize_hydro_array(vpic_hydro_t *h, const vpic_grid_t *g,
                                     void *stream);

/* ------------------------------------------------------------------------ */
/* Output formats (src/vpic/dump.cc, src/vpic/dumpmacros.h), written from  */
/* host arrays: the caller downloads what it dumps                          */
/* ------------------------------------------------------------------------ */

/* What the version-0 header takes besides the grid (dumpmacros.h:7-43):
 * the step, the local domain's low corner (grid_t.x0, y0, z0) and the
 * rank / rank count. */
typedef struct vpic_dump_info {
  int64_t step;
  float x0, y0, z0;
  int32_t rank, nproc;
} vpic_dump_info_t;

/* dump_energies (dump.cc:38-77): one line "step ex ey ez bx by bz KE..."
 * (%li, then %e), the field energies en_f as energy_f returns them and one
 * energy_p per species; a new file (append = 0) first gets the layout
 * comment with the species names and the time step. */
int vpic_hip_write_energies(const char *fname, int append,
                            const vpic_dump_info_t *info, float dt,
                            const double en_f[6], const double *en_p,
                            const char *const *sp_names, int32_t nsp);

/* dump_fields (dump.cc:181-219): header (dump type 1), then the
 * (nx+2)(ny+2)(nz+2) field_t array f. */
int vpic_hip_write_fields(const char *fname, const vpic_field_t *f,
                          const vpic_grid_t *g, const vpic_dump_info_t *info);

/* dump_hydro (dump.cc:221-270): header (type 2, species id, q/m), then the
 * hydro array h (as synchronize_hydro_array leaves it). */
int vpic_hip_write_hydro(const char *fname, const vpic_hydro_t *h,
                         const vpic_grid_t *g, const vpic_dump_info_t *info,
                         int32_t sp_id, float q_m);

/* dump_particles (dump.cc:273-335): header (type 3, species id, q/m) and an
 * array header of np_total particles, then the n particles of p; the rest
 * (np_total - n, already time-centred with center_p as dump_particles
 * does) is appended to the file by the caller. */
int vpic_hip_write_particles(const char *fname, const vpic_particle_t *p,
                             int32_t n, int32_t np_total,
                             const vpic_grid_t *g,
                             const vpic_dump_info_t *info, int32_t sp_id,
                             float q_m);

/* dump_grid (dump.cc:138-179): header (type 0), then grid_t.bc[27],
 * range[nproc+1] and the 6*nv neighbor table. */
int vpic_hip_write_grid(const char *fname, const vpic_grid_t *g,
                        const vpic_dump_info_t *info, const int32_t bc[27],
                        const int64_t *range, const int64_t *neighbor);

/* The strided dumps and their global header (dump.cc:337-880, the "new
 * dump logic"), with the reference's DumpParameters (vpic.h:92-117):
 * output_vars is its BitField (bit i = variable i: field_t words 0..23,
 * hydro_t words 0..13; 0xffffffff = all), format band (one block per
 * variable) or band_interleave (whole records). */
enum { VPIC_DUMP_BAND = 0, VPIC_DUMP_BAND_INTERLEAVE = 1 };
typedef struct vpic_dump_params {
  const char *base_dir;        /* DumpParameters.baseDir */
  const char *base_file_name;  /* DumpParameters.baseFileName */
  int32_t stride[3];           /* stride_x, stride_y, stride_z */
  int32_t format;              /* VPIC_DUMP_BAND or _BAND_INTERLEAVE */
  uint32_t output_vars;
} vpic_dump_params_t;

/* What global_header takes besides the grid: the global domain's extents
 * (grid_t x0..z1) and the domain topology px, py, pz; only rank 0 writes. */
typedef struct vpic_global_info {
  float x0, x1, y0, y1, z0, z1;
  int32_t px, py, pz, rank;
} vpic_global_info_t;

/* global_header (dump.cc:405-532): writes "<base>.vpc", the text header of
 * the strided dumps: version, data-header size, time step, cvac, eps0,
 * extents, spacing, topology, then the field variables of params[0] and
 * the hydro variables of each species' params[1..nparams-1]
 * (dump.cc:337-371 tables).  Needs nparams >= 2, as the reference asserts. */
int vpic_hip_global_header(const char *base, const vpic_grid_t *g,
                           const vpic_global_info_t *gi,
                           const vpic_dump_params_t *params, int32_t nparams);

/* field_dump (dump.cc:533-706): mkdir "<base_dir>/T.<step>" and write
 * "<base_dir>/T.<step>/<base_file_name>.<step>.<rank>": the version-0
 * header with the strided extents and spacing (nx/stride_x ...,
 * dx*stride_x ...), an array header of (nx/sx+2)(ny/sy+2)(nz/sz+2)
 * elements, then the data: band = for each selected word v (ascending),
 * word v of every output voxel; band_interleave = whole field_t records.
 * Output voxels along each axis are the ghost 0, then k*stride - 1 for
 * k = 1..n/stride, then the ghost n+1.  step < 0 means info->step.  As in
 * the reference, words 20..23 (the last four material groups, counted as
 * one word each) lie in the next voxel's record; for the array's last
 * voxel, which has none, 0 is written.  A stride that does not divide the
 * grid, or an output_vars whose word count the reference would take past
 * its variable list, is an error (-1). */
int vpic_hip_field_dump(const vpic_field_t *f, const vpic_grid_t *g,
                        const vpic_dump_info_t *info,
                        const vpic_dump_params_t *p, int64_t step);

/* hydro_dump (dump.cc:708-880), the same for a hydro array (as
 * synchronize_hydro_array leaves it), header with the species id and q/m.
 * band as for the fields; band_interleave writes nx/sx * ny/sy * nz/sz
 * records without the ghost layers' extra count, taken as the reference
 * takes them (the first records of the array for unit strides). */
int vpic_hip_hydro_dump(const vpic_hydro_t *h, const vpic_grid_t *g,
                        const vpic_dump_info_t *info, int32_t sp_id,
                        float q_m, const vpic_dump_params_t *p, int64_t step);

/* ------------------------------------------------------------------------ */
/* Set-up helpers (no reference counterpart on the device; they fill the   */
/* same arrays the reference's set-up code fills on the host)               */
/* ------------------------------------------------------------------------ */

/* Fill g as partition_periodic_box does for one rank
 * (src/grid/partition.cc:35-90 + src/grid/ops.cc:19-165); neighbor = NULL,
 * pbc and fbc = periodic. */
int vpic_hip_grid_periodic(vpic_grid_t *g, double gx0, double gy0,
                           double gz0, double gx1, double gy1, double gz1,
                           int32_t gnx, int32_t gny, int32_t gnz, float dt,
                           float cvac, float eps0);

/* Vacuum material coefficients as create_sfa_params computes them for
 * eps = mu = 1, sigma = 0 (standard/sfa.cc:100-136). */
int vpic_hip_field_params_vacuum(vpic_field_params_t *mp, float damp);

/* Fill the host neighbor table (6*nv int64) that size_grid + join_grid
 * build for this one-rank grid (src/grid/ops.cc:45-165). */
int vpic_hip_grid_neighbor_table(const vpic_grid_t *g, int64_t *host_nb);

/* Copy `bytes` between any two host or device buffers (hipMemcpyDefault),
 * synchronous on `stream`: how a host caller fills or reads arrays it only
 * holds device pointers to (e.g. an ensemble box's). */
int vpic_hip_copy(void *dst, const void *src, int64_t bytes, void *stream);

/* Load a uniform thermal plasma: `ppc` particles per interior cell,
 * positions uniform in the cell, momenta N(drift, uth) per axis, weight w;
 * counter-based RNG keyed by `seed`.  Particle k belongs to cell
 * k / ppc in x-fastest order.  Writes np = ppc*nx*ny*nz particles. */
int vpic_hip_load_uniform_plasma(vpic_particle_t *p, const vpic_grid_t *g,
                                 int32_t ppc, float uth_x, float uth_y,
                                 float uth_z, float ud_x, float ud_y,
                                 float ud_z, float w, uint64_t seed,
                                 void *stream);

/* ------------------------------------------------------------------------ */
/* Batched ensemble of independent boxes (no counterpart in the reference) */
/* ------------------------------------------------------------------------ */

/* Many small independent periodic boxes advanced together.  Box b has its
 * own grid (any size; all faces periodic, no neighbor table) and one pushed
 * species of np[b] particles with charge q[b] and mass m[b] (immobile ions
 * are a neutralising background the step never sees).  The boxes' arrays
 * are packed one after another in HBM, each in the reference's single-grid
 * layout (particle_t with box-local voxel indices, field_t, interpolator_t,
 * accumulator_t); vpic_hip_ensemble_box returns a box's device pointers so
 * its particles and fields can be loaded and read with the single-grid
 * entry points.  A step is vpic_simulation::advance's hot path
 * (src/vpic/advance.cc:25-185) for every box: one persistent launch per
 * phase over a work queue of (box, range) items, the whole step captured in
 * a hipGraph and replayed.  Fields start at zero; after loading the
 * particles (and any initial fields) call vpic_hip_ensemble_prepare (the
 * interpolators).  `stream` carries every launch (NULL: the ensemble
 * creates a stream of its own, as a graph cannot be captured on the legacy
 * default stream). */
typedef struct vpic_ensemble vpic_ensemble_t;

int vpic_hip_ensemble_create(vpic_ensemble_t **e, int32_t nbox,
                             const vpic_grid_t *grids, const int32_t *np,
                             const float *q, const float *m, void *stream);
int vpic_hip_ensemble_destroy(vpic_ensemble_t *e);
/* Device pointers of box b's particles, fields, interpolators and
 * accumulators inside the packed arrays (any may be NULL). */
int vpic_hip_ensemble_box(vpic_ensemble_t *e, int32_t b, vpic_particle_t **p,
                          int32_t *np, vpic_field_t **f,
                          vpic_interpolator_t **fi, vpic_accumulator_t **a);
/* load_interpolator_array of every box (after loading particles/fields). */
int vpic_hip_ensemble_prepare(vpic_ensemble_t *e);
/* nstep steps of every box (with the sorts vpic_hip_ensemble_sort_interval
 * asks for); use_graph replays the captured step (captured at the first
 * call and after every shape change), else the phases are launched one by
 * one.  Synchronous; fails if a particle left its box. */
int vpic_hip_ensemble_step(vpic_ensemble_t *e, int32_t nstep,
                           int32_t use_graph);
/* Launch shapes: push block size (64..256) and particles per work item,
 * field block size (64..512) and voxels / face elements per work item, and
 * persistent blocks per compute unit. */
int vpic_hip_ensemble_shape(vpic_ensemble_t *e, int32_t push_bs,
                            int32_t push_chunk, int32_t field_bs,
                            int32_t field_chunk, int32_t blocks_per_cu);
int vpic_hip_ensemble_get_shape(const vpic_ensemble_t *e, int32_t out[5]);
/* sort_p of every box (src/species_advance/standard/sort_p.cc:154-165, the
 * voxel order; particles of one voxel keep no particular order) before
 * every step whose count of steps done is a multiple of `interval`, as
 * advance.cc:31-37 sorts a species every sort_interval steps; 0 never
 * sorts.  Default 20.  The sort is one counting sort over the packed
 * arrays, captured in a graph of its own. */
int vpic_hip_ensemble_sort_interval(vpic_ensemble_t *e, int32_t interval);
/* Tile of the ensemble's tiled push (cells per axis, clamped to each box;
 * default 16x8x4): once a sort has written the partitions, every box is
 * pushed tile by tile with LDS current windows, as vpic_hip_advance_p with a
 * partition pushes one grid.  0,0,0 keeps the flat push (run-summed global
 * atomics).  The tuner also picks the tile. */
int vpic_hip_ensemble_tile(vpic_ensemble_t *e, int32_t tx, int32_t ty,
                           int32_t tz);
int vpic_hip_ensemble_get_tile(const vpic_ensemble_t *e, int32_t out[3]);
/* The same with the tiles' LDS window halo (1 to 3 cells; none along an
 * axis of one cell, so a 2D box's window grows only in-plane):
 * vpic_hip_ensemble_tile sets halo 1.  out[3] of the getter is the halo.
 * The tuner tries halos 1 to 3 on flat tiles. */
int vpic_hip_ensemble_tile_window(vpic_ensemble_t *e, int32_t tx, int32_t ty,
                                  int32_t tz, int32_t halo);
int vpic_hip_ensemble_get_tile_window(const vpic_ensemble_t *e,
                                      int32_t out[4]);
/* The tuner: times each candidate shape over trial_steps graph replays from
 * the same saved state (restored afterwards), keeps the fastest and
 * returns its ms per step. */
int vpic_hip_ensemble_tune(vpic_ensemble_t *e, int32_t trial_steps,
                           double *host_best_ms);

/* ------------------------------------------------------------------------ */
/* Differentiable run (no counterpart in the reference)                    */
/* ------------------------------------------------------------------------ */

/* The magnetic field energy after n_step steps of the hot path and, when
 * u_bar is given, its gradient by a hand-written reverse-mode adjoint.
 * The step is vpic_simulation::advance (src/vpic/advance.cc:25-185) for one
 * pushed species on a periodic box (immobile ions are a neutralising
 * background the step never sees): clear_accumulator_array, advance_p
 * (advance_p_pipeline.cc:96-238, move_p.cc:233-375), unload_accumulator +
 * synchronize_jf, advance_b(0.5), vacuum_advance_e with no damping,
 * advance_b(0.5), load_interpolator_array; the energy is the magnetic part
 * of energy_f (vacuum_energy_f_pipeline.h:44-75), FP64 sums.
 *
 * precision 4: p0 is vpic_particle_t[np] and f0 vpic_field_t[nv] (the
 * product run, FP32, the reference's push arithmetic); precision 8: the
 * same members as doubles (64-B particles: dx,dy,dz, int32 i, pad, ux,uy,uz,
 * w; 144-B fields: 16 doubles then the 8 int16 material ids), the FP64 run
 * that finite-difference checks use.  Only E and cB of f0 are read; the
 * initial momenta are taken as given (uncentred already).  The forward
 * keeps a checkpoint every ckpt_interval steps; the reverse recomputes each
 * segment from its checkpoint and runs the transposed field operations and
 * the reverse push (the cell-crossing branch held fixed) backwards.
 * Outputs (device arrays in the run's precision, may be NULL except u_bar
 * when gradients are wanted): u_bar[np][3] = dW/d(ux,uy,uz), d_bar[np][3] =
 * dW/d(dx,dy,dz), f_bar[nv][6] = dW/d(ex,ey,ez,cbx,cby,cbz) of the initial
 * state.  *host_energy receives W.  All grid faces must be periodic.
 * Synchronous; allocates its checkpoints (about (n_step/ckpt_interval +
 * ckpt_interval) particle and field arrays). */
int vpic_hip_adjoint_bfield_energy(const void *p0, int32_t np, float q,
                                   float m, const void *f0,
                                   const vpic_grid_t *g, int32_t n_step,
                                   int32_t ckpt_interval, int32_t precision,
                                   double *host_energy, void *u_bar,
                                   void *d_bar, void *f_bar, void *stream);

/* ------------------------------------------------------------------------ */
/* Timing of the push kernel (hipEvents on the launch stream)               */
/* ------------------------------------------------------------------------ */

/* When enabled, every advance_p brackets its main push kernel with
 * hipEvents on its stream (k_advance_p_tiled when a partition is given,
 * else k_advance_p); vpic_hip_push_timing returns the launches, total ms and
 * particles since the last reset (synchronises the recorded events). */
int vpic_hip_push_timing_enable(int on);
int vpic_hip_push_timing(int64_t *host_launches, double *host_total_ms,
                         int64_t *host_particles);

#ifdef __cplusplus
}
#endif

#endif /* VPIC_HIP_H */
