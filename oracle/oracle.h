/*
 * oracle.h -- CPU restatement of VPIC's per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the device path in
 * vpic_amd/ is compared against; only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it, and never as the thing that is
 * measured or shipped.  Arrays are host arrays with the reference's
 * layouts (include/vpic_hip.h).  Each function cites the reference
 * (lanl/vpic legacy) file:line whose arithmetic order it follows; FP32
 * expressions are evaluated exactly as written there (build with
 * -ffp-contract=off, no fast-math).
 */
#ifndef VPIC_ORACLE_H
#define VPIC_ORACLE_H

#include <stdint.h>
#include "../include/vpic_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* partition_periodic_box for one rank (src/grid/partition.cc:35-90) plus
 * size_grid strides (src/grid/ops.cc:25-30). */
void orc_grid_periodic(vpic_grid_t *g, double gx0, double gy0, double gz0,
                       double gx1, double gy1, double gz1, int gnx, int gny,
                       int gnz, float dt, float cvac, float eps0);

/* size_grid + join_grid / set_pbc neighbour table for one rank
 * (src/grid/ops.cc:45-76,119-165): 6*nv int64 into nb. */
void orc_neighbor_table(const vpic_grid_t *g, int64_t *nb);

/* move_p, scalar branch (src/species_advance/standard/move_p.cc:215-378).
 * Uses g->neighbor (host table). */
int orc_move_p(vpic_particle_t *p0, vpic_particle_mover_t *pm,
               vpic_accumulator_t *a0, const vpic_grid_t *g, float qsp);

/* advance_p with one pipeline plus the host straggler pipeline, as
 * EXEC_PIPELINES runs it with N_PIPELINE = 1
 * (advance_p_pipeline.cc:21-245,253-340; pipelines_exec_pth.h:56-65):
 * particles [0, np&~15) accumulate into copy 1, the tail into copy 0, then
 * reduce_accumulator_array adds copy 1 into copy 0
 * (reduce_array_pipeline.cc:21-60).  `a` is the single accumulator copy
 * the caller sees (nv entries, added to, not overwritten).  Movers are
 * compacted as advance_p_pipeline.cc:310-339 does.  Returns nm. */
int orc_advance_p(vpic_particle_t *p, int np, float q, float m,
                  vpic_particle_mover_t *pm, int max_nm, int *n_ignored,
                  vpic_accumulator_t *a, const vpic_interpolator_t *f,
                  const vpic_grid_t *g);

/* The per-particle coupling constants of advance_p_pipeline.cc:279-283. */
void orc_push_coefficients(float q, float m, const vpic_grid_t *g,
                           float out[4]);

/* load_interpolator_pipeline_scalar over voxels 1..n
 * (interpolator_array_pipeline.cc:21-135). */
void orc_load_interpolator(vpic_interpolator_t *fi, const vpic_field_t *f,
                           const vpic_grid_t *g);

/* unload_accumulator over voxels 1..n+1 with the coupling constants of
 * unload_accumulator_pipeline.cc:137-139 (unload_accumulator_pipeline.cc:
 * 18-82). */
void orc_unload_accumulator(vpic_field_t *f, const vpic_accumulator_t *a,
                            const vpic_grid_t *g);

/* clear_jf (standard/sfa.cc:232-237) and clear_rhof. */
void orc_clear_jf(vpic_field_t *f, const vpic_grid_t *g);
void orc_clear_rhof(vpic_field_t *f, const vpic_grid_t *g);

/* synchronize_jf: local_adjust_jf then the x, y, z face folds in that
 * order (standard/remote.cc:418-507, standard/local.cc:336-368). */
void orc_synchronize_jf(vpic_field_t *f, const vpic_grid_t *g);

/* synchronize_rho for one domain: local_adjust_rhof/rhob then the periodic
 * node-plane folds (standard/remote.cc:534-625, standard/local.cc:376-440). */
void orc_synchronize_rho(vpic_field_t *f, const vpic_grid_t *g);

/* advance_b: bulk + host leftovers + local_adjust_norm_b
 * (advance_b_pipeline.h:17-59, advance_b_pipeline.cc:20-125,
 * standard/local.cc:266-296). */
void orc_advance_b(vpic_field_t *f, const vpic_grid_t *g, float frac);

/* vacuum_advance_e: remote/local tangential-B ghosts, every E edge, then
 * local_adjust_tang_e (vacuum_advance_e_pipeline.h:17-69,
 * vacuum_advance_e_pipeline.cc:49-332, standard/local.cc:50-122,224-264,
 * standard/remote.cc:61-134). */
void orc_vacuum_advance_e(vpic_field_t *f, const vpic_grid_t *g,
                          const vpic_field_params_t *mp);

/* create_sfa_params for the vacuum material eps=mu=1, sigma=0
 * (standard/sfa.cc:100-136). */
void orc_field_params_vacuum(vpic_field_params_t *mp, float damp);

/* sort_p with one pipeline: stable counting sort by voxel and the
 * partition array (sort_p_pipeline.cc:143-214,221-371).  partition has
 * nv+1 entries; aux holds np particles of scratch. */
void orc_sort_p(vpic_particle_t *p, vpic_particle_t *aux, int np,
                int *partition, const vpic_grid_t *g);

/* energy_p with one pipeline + host tail (energy_p_pipeline.cc:18-115). */
double orc_energy_p(const vpic_particle_t *p, int np, float q, float m,
                    const vpic_interpolator_t *f, const vpic_grid_t *g);

/* vacuum energy_f with one pipeline + host tail
 * (vacuum_energy_f_pipeline.h:17-75, vacuum_energy_f_pipeline.cc:13-97). */
void orc_energy_f(double en[6], const vpic_field_t *f, const vpic_grid_t *g,
                  const vpic_field_params_t *mp);

/* center_p (uncenter = 0) and uncenter_p (uncenter = 1): half E kick and
 * half Boris rotation of the momenta, forward or backward in time
 * (center_p_pipeline.cc:15-97, uncenter_p_pipeline.cc:15-97). */
void orc_center_p(vpic_particle_t *p, int np, float q, float m,
                  const vpic_interpolator_t *f0, const vpic_grid_t *g,
                  int uncenter);

/* accumulate_rho_p (src/species_advance/standard/rho_p.cc:23-113). */
void orc_accumulate_rho_p(vpic_field_t *f, const vpic_particle_t *p, int np,
                          float q, const vpic_grid_t *g);

/* accumulate_rhob (rho_p.cc:126-173) of one particle. */
void orc_accumulate_rhob(vpic_field_t *f, const vpic_particle_t *p,
                         float qsp, const vpic_grid_t *g);

/* boundary_p's mover loop and injection loop for one species of a domain
 * with no shared face (boundary_p.cc:254-375, 595-711); the grid needs its
 * neighbor table.  orc_boundary_p returns the new particle count. */
int orc_boundary_p(vpic_particle_t *p0, int np,
                   const vpic_particle_mover_t *pm, int nm, vpic_field_t *f,
                   float q, const vpic_grid_t *g, vpic_boundary_hit_t *hits,
                   int32_t *counts);
void orc_inject_p(vpic_particle_t *p, int *np, int max_np,
                  vpic_particle_mover_t *pm, int *nm, int max_nm,
                  const vpic_particle_injector_t *pi, int n,
                  vpic_accumulator_t *a, float q, const vpic_grid_t *g,
                  int32_t *counts);

/* The divergence-cleaning and initialization kernels of the vacuum field
 * table (sfa.cc:203-210), each over the whole local domain with its
 * ghosts and local adjusts:
 *   compute_div_e_err  vacuum_compute_div_e_err_pipeline.{h:17-50,cc:61-174}
 *   compute_rhob       vacuum_compute_rhob_pipeline.{h:17-52,cc:41-175}
 *   compute_rms_div_e_err  compute_rms_div_e_err_pipeline.cc:16-180
 *   clean_div_e        vacuum_clean_div_e_pipeline.{h:16-55,cc:43-145}
 *   compute_div_b_err  compute_div_b_err_pipeline.cc:15-57
 *   compute_rms_div_b_err  compute_rms_div_b_err_pipeline.cc:14-111
 *   clean_div_b        clean_div_b_pipeline.{h:17-27,cc:230-328}
 *   compute_curl_b     vacuum_compute_curl_b_pipeline.{h:17-66,cc:50-326}
 *   synchronize_tang_e_norm_b  remote.cc:298-416 (one domain)
 * with the ghosts of local.cc:129-221 and remote.cc:137-281. */
void orc_compute_div_e_err(vpic_field_t *f, const vpic_grid_t *g,
                           const vpic_field_params_t *mp);
void orc_compute_rhob(vpic_field_t *f, const vpic_grid_t *g,
                      const vpic_field_params_t *mp);
double orc_compute_rms_div_e_err(const vpic_field_t *f, const vpic_grid_t *g);
void orc_clean_div_e(vpic_field_t *f, const vpic_grid_t *g,
                     const vpic_field_params_t *mp);
void orc_compute_div_b_err(vpic_field_t *f, const vpic_grid_t *g);
double orc_compute_rms_div_b_err(const vpic_field_t *f, const vpic_grid_t *g);
void orc_clean_div_b(vpic_field_t *f, const vpic_grid_t *g);
void orc_compute_curl_b(vpic_field_t *f, const vpic_grid_t *g,
                        const vpic_field_params_t *mp);
double orc_synchronize_tang_e_norm_b(vpic_field_t *f, const vpic_grid_t *g);

/* accumulate_hydro_p with one pipeline and its reduction
 * (hydro_p_pipeline.cc:17-242, reduce_array_pipeline.cc:209-231), clear,
 * and synchronize_hydro_array for one domain (hydro_array.cc:131-311). */
void orc_accumulate_hydro_p(vpic_hydro_t *h, const vpic_particle_t *p, int np,
                            float q, float m, const vpic_interpolator_t *f,
                            const vpic_grid_t *g);
void orc_clear_hydro(vpic_hydro_t *h, const vpic_grid_t *g);
void orc_synchronize_hydro(vpic_hydro_t *h, const vpic_grid_t *g);

/* The reference's CPU parallel scheme for advance_p, used only as the
 * timed CPU baseline: n_thread pthreads each run one DISTRIBUTE(np,16)
 * block into a private accumulator copy, the caller runs the np%16 tail,
 * then the copies are summed (pipelines_exec_pth.h:56-65,
 * advance_p_pipeline.cc:55-85, accumulator_array.cc:65-87).  `acc` must
 * hold (n_thread+1)*nv accumulators; copy 0 receives the reduced sum. */
int orc_advance_p_threaded(vpic_particle_t *p, int np, float q, float m,
                           vpic_particle_mover_t *pm, int max_nm,
                           vpic_accumulator_t *acc,
                           const vpic_interpolator_t *f, const vpic_grid_t *g,
                           int n_thread);

/* Philox4x32-10 (Salmon, Moraes, Dror and Shaw, "Parallel random numbers:
 * as easy as 1, 2, 3", SC'11): out = the 10-round bijection of ctr under
 * key.  The collision operator's random stream (the device draws the same
 * numbers). */
void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2],
                    uint32_t out[4]);

/* takizuka_abe_pipeline_scalar (src/collision/pipeline/
 * takizuka_abe_pipeline.cc:81-215) on one pipeline: for every voxel
 * 0..VOXEL(nx,ny,nz) the shuffle of species i's particles, then the
 * binary collisions of takizuka_abe_collision (:14-79).  pi == pj (with
 * part_i == part_j) is the intra-species operator.  The random numbers
 * are the collision stream of the device (Philox4x32-10, key = seed,
 * counter = {draw/4, voxel, lo(call), hi(call)}; uirand = a word,
 * frand_c0 = (word >> 8) 2^-24, frandn = Box-Muller in double from two
 * words) in place of the reference's per-pipeline SFMT generator. */
void orc_takizuka_abe(vpic_particle_t *pi, const int *part_i, float qi,
                      float mi, vpic_particle_t *pj, const int *part_j,
                      float qj, float mj, double cvar0, int interval,
                      uint64_t seed, uint64_t call, const vpic_grid_t *g);

/* ------------------------------------------------------------------------ */
/* FP64 oracle: the step functions of oracle_step.inc compiled over double */
/* ------------------------------------------------------------------------ */

/* The FP64 twins of the reference's types: the same members in the same
 * order, every float a double (ints unchanged). */
typedef struct {
  double dx, dy, dz;
  int32_t i, pad_;
  double ux, uy, uz, w;
} orc64_particle_t;

typedef struct {
  double dispx, dispy, dispz;
  int32_t i, pad_;
} orc64_particle_mover_t;

typedef struct {
  double ex, dexdy, dexdz, d2exdydz;
  double ey, deydz, deydx, d2eydzdx;
  double ez, dezdx, dezdy, d2ezdxdy;
  double cbx, dcbxdx;
  double cby, dcbydy;
  double cbz, dcbzdz;
  double _pad[2];
} orc64_interpolator_t;

typedef struct {
  double jx[4];
  double jy[4];
  double jz[4];
} orc64_accumulator_t;

typedef struct {
  double ex, ey, ez, div_e_err;
  double cbx, cby, cbz, div_b_err;
  double tcax, tcay, tcaz, rhob;
  double jfx, jfy, jfz, rhof;
  int16_t ematx, ematy, ematz, nmat;
  int16_t fmatx, fmaty, fmatz, cmat;
} orc64_field_t;

typedef struct {
  double dt, cvac, eps0;
  double dx, dy, dz, dV;
  double rdx, rdy, rdz, r8V;
  int32_t nx, ny, nz;
  int32_t sx, sy, sz, nv;
  int64_t rangel, rangeh;
  const int64_t *neighbor; /* host, 6*nv */
  int32_t pbc[6];
  int32_t fbc[6];
} orc64_grid_t;

typedef struct {
  double decayx, drivex, decayy, drivey, decayz, drivez;
  double rmux, rmuy, rmuz;
  double epsx, epsy, epsz;
  double damp;
  double nonconductive;
} orc64_field_params_t;

/* The grid's constants widened to double (the FP64 run then differs from
 * the FP32 one only in its arithmetic). */
void orc64_grid_from(orc64_grid_t *g64, const vpic_grid_t *g);

int orc64_move_p(orc64_particle_t *p0, orc64_particle_mover_t *pm,
                 orc64_accumulator_t *a0, const orc64_grid_t *g, double qsp);
void orc64_push_coefficients(double q, double m, const orc64_grid_t *g,
                             double out[4]);
int orc64_advance_p(orc64_particle_t *p, int np, double q, double m,
                    orc64_particle_mover_t *pm, int max_nm, int *n_ignored,
                    orc64_accumulator_t *a, const orc64_interpolator_t *f,
                    const orc64_grid_t *g);
void orc64_load_interpolator(orc64_interpolator_t *fi, const orc64_field_t *f,
                             const orc64_grid_t *g);
void orc64_unload_accumulator(orc64_field_t *f, const orc64_accumulator_t *a,
                              const orc64_grid_t *g);
void orc64_clear_jf(orc64_field_t *f, const orc64_grid_t *g);
void orc64_clear_rhof(orc64_field_t *f, const orc64_grid_t *g);
void orc64_synchronize_jf(orc64_field_t *f, const orc64_grid_t *g);
void orc64_synchronize_rho(orc64_field_t *f, const orc64_grid_t *g);
void orc64_advance_b(orc64_field_t *f, const orc64_grid_t *g, double frac);
void orc64_vacuum_advance_e(orc64_field_t *f, const orc64_grid_t *g,
                            const orc64_field_params_t *m);
void orc64_sort_p(orc64_particle_t *p, orc64_particle_t *aux, int np,
                  int *partition, const orc64_grid_t *g);
double orc64_energy_p(const orc64_particle_t *p, int np, double q, double m,
                      const orc64_interpolator_t *f, const orc64_grid_t *g);
void orc64_energy_f(double en[6], const orc64_field_t *f,
                    const orc64_grid_t *g, const orc64_field_params_t *m);
void orc64_accumulate_rho_p(orc64_field_t *f, const orc64_particle_t *p,
                            int np, double q, const orc64_grid_t *g);
void orc64_compute_div_e_err(orc64_field_t *f, const orc64_grid_t *g,
                             const orc64_field_params_t *m);
void orc64_compute_rhob(orc64_field_t *f, const orc64_grid_t *g,
                        const orc64_field_params_t *m);
double orc64_compute_rms_div_e_err(const orc64_field_t *f,
                                   const orc64_grid_t *g);
void orc64_center_p(orc64_particle_t *p, int np, double q, double m,
                    const orc64_interpolator_t *f0, const orc64_grid_t *g,
                    int uncenter);

#ifdef __cplusplus
}
#endif

#endif
