/*
 * vpic_dropin.h -- the reference-named entry points of libvpichip: the
 * functions of lanl/vpic's per-step particle and field kernels, same names,
 * same signatures, same struct layouts, running on the MI355X.
 *
 * A VPIC build uses them by linking libvpichip.so in place of the
 * translation units that define these functions (advance_p.cc, sort_p.cc,
 * move_p.cc, energy_p.cc, rho_p.cc, hydro_p.cc, boundary_p.cc,
 * interpolator_array.cc's load_interpolator_array, hydro_array.cc's
 * synchronize_hydro_array, clear_array.cc, reduce_array.cc,
 * unload_accumulator.cc), and by calling vpic_hip_install_field_kernels()
 * on each field array after new_standard_field_array().  Such a build keeps
 * including its own headers; this header restates the same types for
 * callers that do not have them (the layouts are the reference's
 * byte for byte, and a check compiled against the reference headers
 * compares every offset with vpic_hip_dropin_layout()).
 *
 * Residency.  The caller's arrays stay ordinary host arrays.  The library
 * keeps a device (HBM) copy of each array it has seen, keyed by the host
 * pointer.  Two modes:
 *   coherent (default)  every call uploads what it reads and downloads
 *                       what it writes, so host arrays are always current:
 *                       any reference caller works unchanged, at the cost
 *                       of PCIe traffic per call;
 *   resident            the device copies are authoritative; the host
 *                       copies are refreshed only by the *_sync calls
 *                       below (to_device = 0), and host edits are pushed by
 *                       the same calls with to_device = 1.  Only transfers
 *                       the caller asks for happen.
 * The grid's neighbor table is an input like the others (uploaded once in
 * resident mode: call vpic_hip_dropin_release(g->neighbor) after changing
 * it there).  Device copies are keyed by host address: release them
 * before freeing a host array whose address may be reused.
 *
 * Errors follow the reference's convention: a message on stderr and
 * exit(1) (util_base.h:267-273), unless an error handler is installed;
 * overflowing the mover array prints the reference's warning
 * (advance_p_pipeline.cc:314-329).  Calls are synchronous on return and
 * not reentrant, as in the reference.
 *
 * Citations are paths in lanl/vpic (legacy), relative to its root.
 */
#ifndef VPIC_DROPIN_H
#define VPIC_DROPIN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#ifndef VPIC_DROPIN_NO_TYPES
/* ------------------------------------------------------------------------ */
/* Reference types (scalar build: ALIGNED() expands to nothing,             */
/* util_base.h:93-94; PAD_SIZE_INTERPOLATOR = 2; no accumulator padding)    */
/* ------------------------------------------------------------------------ */

typedef int32_t species_id;   /* species_advance_aos.h:15 */
typedef int16_t material_id;  /* material/material.h:10 */
struct mp;                    /* util/mp: opaque here */

/* src/species_advance/species_advance_aos.h:21-32 */
typedef struct particle {
  float dx, dy, dz;
  int32_t i;
  float ux, uy, uz;
  float w;
} particle_t;

/* src/species_advance/species_advance_aos.h:37-40 */
typedef struct particle_mover {
  float dispx, dispy, dispz;
  int32_t i;
} particle_mover_t;

/* src/species_advance/species_advance_aos.h:45-52 */
typedef struct particle_injector {
  float dx, dy, dz;
  int32_t i;
  float ux, uy, uz, w;
  float dispx, dispy, dispz;
  species_id sp_id;
} particle_injector_t;

/* src/grid/grid.h:73-131 */
typedef struct grid {
  float dt, cvac, eps0;
  int64_t step;
  double t0;
  float x0, y0, z0;
  float x1, y1, z1;
  int nx, ny, nz;
  float dx, dy, dz, dV;
  float rdx, rdy, rdz, r8V;
  int sx, sy, sz, nv;
  int bc[27];
  int64_t *range;
  int64_t *neighbor;
  int64_t rangel, rangeh;
  struct mp *mp;
} grid_t;

/* src/species_advance/species_advance_aos.h:54-94 */
typedef struct species {
  char *name;
  float q;
  float m;
  int np, max_np;
  particle_t *p;
  int nm, max_nm;
  particle_mover_t *pm;
  int64_t last_sorted;
  int sort_interval;
  int sort_out_of_place;
  int *partition;
  grid_t *g;
  species_id id;
  struct species *next;
} species_t;

/* src/sf_interface/sf_interface.h:62-80 */
typedef struct interpolator {
  float ex, dexdy, dexdz, d2exdydz;
  float ey, deydz, deydx, d2eydzdx;
  float ez, dezdx, dezdy, d2ezdxdy;
  float cbx, dcbxdx;
  float cby, dcbydy;
  float cbz, dcbzdz;
  float _pad1[2];
} interpolator_t;

typedef struct interpolator_array {
  interpolator_t *i;
  grid_t *g;
} interpolator_array_t;

/* src/sf_interface/sf_interface.h:115-131 */
typedef struct accumulator {
  float jx[4];
  float jy[4];
  float jz[4];
} accumulator_t;

typedef struct accumulator_array {
  accumulator_t *a;
  int n_pipeline;
  int stride;
  grid_t *g;
} accumulator_array_t;

/* src/sf_interface/sf_interface.h:185-200 */
typedef struct hydro {
  float jx, jy, jz, rho;
  float px, py, pz, ke;
  float txx, tyy, tzz;
  float tyz, tzx, txy;
  float _pad[2];
} hydro_t;

typedef struct hydro_array {
  hydro_t *h;
  int n_pipeline;
  int stride;
  grid_t *g;
} hydro_array_t;

/* src/boundary/boundary_private.h:17-42: a custom particle boundary; its
 * id is the (negative) code its faces carry in the neighbor table, and
 * boundary_p calls interact for each particle that hits one. */
struct particle_bc;
typedef struct particle_bc particle_bc_t;
typedef int (*particle_bc_func_t)(void *b, struct species *sp, particle_t *p,
                                  particle_mover_t *pm,
                                  particle_injector_t *pi, int max_pi,
                                  int face);
typedef void (*delete_particle_bc_func_t)(particle_bc_t *pbc);
struct particle_bc {
  void *params;
  particle_bc_func_t interact;
  delete_particle_bc_func_t delete_pbc;
  int64_t id;
  particle_bc_t *next;
};

/* src/field_advance/field_advance.h:152-229 */
typedef struct field {
  float ex, ey, ez, div_e_err;
  float cbx, cby, cbz, div_b_err;
  float tcax, tcay, tcaz, rhob;
  float jfx, jfy, jfz, rhof;
  material_id ematx, ematy, ematz, nmat;
  material_id fmatx, fmaty, fmatz, cmat;
} field_t;

struct field_array;

typedef struct field_advance_kernels {
  void (*delete_fa)(struct field_array *fa);
  void (*advance_b)(struct field_array *fa, float frac);
  void (*advance_e)(struct field_array *fa, float frac);
  void (*energy_f)(double *en, const struct field_array *fa);
  void (*clear_jf)(struct field_array *fa);
  void (*synchronize_jf)(struct field_array *fa);
  void (*clear_rhof)(struct field_array *fa);
  void (*synchronize_rho)(struct field_array *fa);
  void (*compute_rhob)(struct field_array *fa);
  void (*compute_curl_b)(struct field_array *fa);
  double (*synchronize_tang_e_norm_b)(struct field_array *fa);
  void (*compute_div_e_err)(struct field_array *fa);
  double (*compute_rms_div_e_err)(const struct field_array *fa);
  void (*clean_div_e)(struct field_array *fa);
  void (*compute_div_b_err)(struct field_array *fa);
  double (*compute_rms_div_b_err)(const struct field_array *fa);
  void (*clean_div_b)(struct field_array *fa);
} field_advance_kernels_t;

typedef struct field_array {
  field_t *f;
  struct grid *g;
  void *params;
  field_advance_kernels_t kernel[1];
} field_array_t;

/* src/field_advance/standard/sfa_private.h:14-30: what
 * new_standard_field_array stores in field_array_t.params. */
typedef struct material_coefficient {
  float decayx, drivex;
  float decayy, drivey;
  float decayz, drivez;
  float rmux, rmuy, rmuz;
  float nonconductive;
  float epsx, epsy, epsz;
  float pad[3];
} material_coefficient_t;

typedef struct sfa_params {
  material_coefficient_t *mc;
  int n_mc;
  float damp;
} sfa_params_t;

/* src/grid/grid.h:16 */
#ifndef BOUNDARY
#define BOUNDARY(i, j, k) (13 + (i) + 3 * (j) + 9 * (k))
#endif
#endif /* VPIC_DROPIN_NO_TYPES */

/* ------------------------------------------------------------------------ */
/* Species kernels (src/species_advance/species_advance.h)                  */
/* ------------------------------------------------------------------------ */

/* species_advance.h:74-76 (advance_p.cc:12-19, advance_p_pipeline.cc:
 * 253-340).  Pushes sp->p[0:np], adds currents into aa->a (copy 0), leaves
 * the movers the grid cannot handle in sp->pm[0:nm] and sets sp->nm.  After
 * a sort_p of this library the push runs cell-tiled. */
void advance_p(species_t *sp, accumulator_array_t *aa,
               const interpolator_array_t *ia);

/* species_advance.h:66 (sort_p.cc:154-165, sort_p_pipeline.cc:31-371):
 * stable sort of sp->p by voxel, fills sp->partition (when allocated) and
 * sets sp->last_sorted = g->step. */
void sort_p(species_t *sp);

/* species_advance.h:153-157 (move_p.cc:215-378): moves p0[m->i] by the
 * mover's displacement, depositing into a0; returns 1 when the particle
 * leaves the local domain (p->i = 8*voxel + face), else 0. */
int move_p(particle_t *p0, particle_mover_t *m, accumulator_t *a0,
           const grid_t *g, const float qsp);

/* species_advance.h:119-120 (energy_p_pipeline.cc:18-115). */
double energy_p(const species_t *sp, const interpolator_array_t *ia);

/* species_advance.h:128-130 (rho_p.cc:23-113): adds the species' charge
 * density into fa->f[].rhof. */
void accumulate_rho_p(field_array_t *fa, const species_t *sp);

/* species_advance.h:133-136 (rho_p.cc:126-173): the charge of one
 * particle (voxel index restored, not 8*voxel+face) on the 8 nodes of its
 * voxel into f[].rhob, doubled on the local domain's faces; called by
 * vpic_simulation::advance for movers boundary_p could not place
 * (advance.cc:78-101).  f is the field array's host array. */
void accumulate_rhob(field_t *f, const particle_t *p, const grid_t *g,
                     const float qsp);

/* species_advance.h:84-100 (center_p_pipeline.cc, uncenter_p_pipeline.cc):
 * half Boris step forward / backward in time of the momenta. */
void center_p(species_t *sp, const interpolator_array_t *ia);
void uncenter_p(species_t *sp, const interpolator_array_t *ia);

/* species_advance.h:140-143 (hydro_p.cc:12-19, hydro_p_pipeline.cc:
 * 218-244): adds the species' 14 moments into ha->h (copy 0).  After a
 * sort_p of this library the sum runs cell-tiled.  Particles whose voxel
 * is not interior are an error. */
void accumulate_hydro_p(hydro_array_t *ha, const species_t *sp,
                        const interpolator_array_t *ia);

/* boundary.h:35-39 (boundary_p.cc:42-750) on a domain with no face shared
 * with another rank: for every species, the movers advance_p left are
 * taken last first; a particle on an absorbing face puts its charge on
 * rhob and is removed, one on a custom boundary face (a particle_bc_t of
 * pbc_list, matched by id) is handed to that boundary's interact function
 * on the host and removed, anything else is dropped with the reference's
 * warning; the last particle fills each hole, as in the reference.  The
 * particles the handlers inject are then appended and moved with move_p,
 * last first.  Particle storage is not resized: an injection that does not
 * fit is dropped with the warning of the reference's fixed-size build.
 * Custom handlers see the host field array (downloaded before and uploaded
 * after them when any of them runs). */
void boundary_p(particle_bc_t *pbc_list, species_t *sp_list,
                field_array_t *fa, accumulator_array_t *aa);

/* ------------------------------------------------------------------------ */
/* Field-particle interface (src/sf_interface/sf_interface.h)               */
/* ------------------------------------------------------------------------ */

/* sf_interface.h:100-101 (interpolator_array.cc:52-63). */
void load_interpolator_array(interpolator_array_t *ia,
                             const field_array_t *fa);

/* sf_interface.h:148 (clear_array.cc:11-21): zeroes every copy. */
void clear_accumulator_array(accumulator_array_t *aa);

/* sf_interface.h:159 (reduce_array.cc:11-21).  The device keeps one
 * accumulator copy, so there is nothing to reduce: a checked no-op. */
void reduce_accumulator_array(accumulator_array_t *aa);

/* sf_interface.h:173-174 (unload_accumulator.cc:11-23). */
void unload_accumulator_array(field_array_t *fa,
                              const accumulator_array_t *aa);

/* sf_interface.h:218-222 (clear_array.cc:29-38): zeroes every copy. */
void clear_hydro_array(hydro_array_t *ha);

/* sf_interface.h:224-231 (reduce_array.cc:29-38).  One device copy: a
 * checked no-op, as for the accumulators. */
void reduce_hydro_array(hydro_array_t *ha);

/* sf_interface.h:233-240 (hydro_array.cc:131-311) for a domain that is
 * the whole problem: local faces doubled, periodic node planes summed. */
void synchronize_hydro_array(hydro_array_t *ha);

/* ------------------------------------------------------------------------ */
/* Field kernels.  The reference reaches them through fa->kernel            */
/* (field_advance.h:170-218), filled by new_standard_field_array            */
/* (sfa.cc:191-216).  These are the HIP versions of the vacuum kernels.     */
/* ------------------------------------------------------------------------ */

/* Replace the entries of fa->kernel with the functions below (every entry
 * but delete_fa: advance_b, advance_e, energy_f, clear_jf, synchronize_jf,
 * clear_rhof, synchronize_rho, compute_rhob, compute_curl_b,
 * synchronize_tang_e_norm_b, compute_div_e_err, compute_rms_div_e_err,
 * clean_div_e, compute_div_b_err, compute_rms_div_b_err, clean_div_b).  fa must come from
 * new_standard_field_array with ONE material (the vacuum kernels,
 * sfa.cc:205-214) and fa->params must be its sfa_params_t.  The other
 * entries are left as they are.  Returns 0, or < 0 with
 * vpic_hip_last_error() set (then fa is unchanged). */
int vpic_hip_install_field_kernels(field_array_t *fa);

/* advance_b.cc / advance_b_pipeline.cc:20-125 (sfa_private.h:51-56). */
void vpic_hip_fa_advance_b(field_array_t *fa, float frac);
/* vacuum_advance_e.cc:12-30 -> vacuum_advance_e_pipeline.cc:20-332. */
void vpic_hip_fa_advance_e(field_array_t *fa, float frac);
/* vacuum_energy_f.cc -> vacuum_energy_f_pipeline.cc:13-97. */
void vpic_hip_fa_energy_f(double *en, const field_array_t *fa);
/* sfa.cc:232-237, sfa.cc:239-245. */
void vpic_hip_fa_clear_jf(field_array_t *fa);
void vpic_hip_fa_clear_rhof(field_array_t *fa);
/* remote.cc:418-507 + local.cc:336-368; remote.cc:535-625 + local.cc:
 * 298-334. */
void vpic_hip_fa_synchronize_jf(field_array_t *fa);
void vpic_hip_fa_synchronize_rho(field_array_t *fa);
/* vacuum_compute_rhob.cc -> vacuum_compute_rhob_pipeline.cc:41-175. */
void vpic_hip_fa_vacuum_compute_rhob(field_array_t *fa);
/* vacuum_compute_curl_b.cc -> vacuum_compute_curl_b_pipeline.cc:50-326. */
void vpic_hip_fa_vacuum_compute_curl_b(field_array_t *fa);
/* remote.cc:298-416 (one domain). */
double vpic_hip_fa_synchronize_tang_e_norm_b(field_array_t *fa);
/* vacuum_compute_div_e_err.cc -> vacuum_compute_div_e_err_pipeline.cc:
 * 61-174. */
void vpic_hip_fa_vacuum_compute_div_e_err(field_array_t *fa);
/* compute_rms_div_e_err.cc -> compute_rms_div_e_err_pipeline.cc:64-185. */
double vpic_hip_fa_compute_rms_div_e_err(const field_array_t *fa);
/* vacuum_clean_div_e.cc -> vacuum_clean_div_e_pipeline.cc:72-147. */
void vpic_hip_fa_vacuum_clean_div_e(field_array_t *fa);
/* compute_div_b_err.cc -> compute_div_b_err_pipeline.cc:260-296. */
void vpic_hip_fa_compute_div_b_err(field_array_t *fa);
/* compute_rms_div_b_err.cc -> compute_rms_div_b_err_pipeline.cc:53-111. */
double vpic_hip_fa_compute_rms_div_b_err(const field_array_t *fa);
/* clean_div_b.cc -> clean_div_b_pipeline.cc:230-328. */
void vpic_hip_fa_clean_div_b(field_array_t *fa);

/* ------------------------------------------------------------------------ */
/* Residency control (no reference counterpart: the reference's arrays     */
/* live in host memory only)                                                */
/* ------------------------------------------------------------------------ */

/* 1 = resident mode, 0 = coherent mode (default).  Switching to coherent
 * first brings every host copy up to date. */
int vpic_hip_dropin_resident(int on);

/* Copy between a host array and its device copy: to_device = 1 uploads the
 * host data (after host edits), 0 downloads the device data.  Species:
 * p[0:np] and pm[0:nm]; field array: f[0:nv]; interpolator: i[0:nv];
 * accumulator and hydro arrays: copy 0. */
int vpic_hip_species_sync(species_t *sp, int to_device);
int vpic_hip_field_array_sync(field_array_t *fa, int to_device);
int vpic_hip_interpolator_array_sync(interpolator_array_t *ia, int to_device);
int vpic_hip_accumulator_array_sync(accumulator_array_t *aa, int to_device);
int vpic_hip_hydro_array_sync(hydro_array_t *ha, int to_device);

/* Free the device copy kept for a host array (call before freeing or
 * reusing the host array).  Unknown pointers are ignored. */
int vpic_hip_dropin_release(const void *host_array);
/* Free every device copy and per-species state. */
int vpic_hip_dropin_release_all(void);

/* Install a handler for errors of the reference-named functions (NULL
 * restores the reference behaviour: message on stderr, exit(1)).  The
 * handler may return; the failing call then returns without effect. */
typedef void (*vpic_hip_error_handler_t)(const char *message);
int vpic_hip_dropin_error_handler(vpic_hip_error_handler_t handler);

/* Layout table of the types above as this library sees them: pairs
 * (sizeof or offsetof) in a fixed order, for a build to compare against
 * its own headers.  Writes min(n, count) values to out and returns count. */
int vpic_hip_dropin_layout(int64_t *out, int n);

#ifdef __cplusplus
}
#endif

#endif /* VPIC_DROPIN_H */
