// ref_spa.cc -- harness for the reference's species-advance kernels:
// advance_p_pipeline_scalar, move_p, energy_p_pipeline_scalar,
// center_p/uncenter_p_pipeline_scalar and subsort_pipeline_scalar (built
// from lanl/vpic src/species_advance).
// TEST INFRASTRUCTURE ONLY: the checker for oracle/oracle.c and the timed
// "reference" CPU baseline.
#define IN_spa
#include "src/species_advance/standard/pipeline/spa_private.h"
#define IN_sf_interface
#include "src/sf_interface/pipeline/sf_interface_pipeline.h"

#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#include "ref_common.h"

static_assert(sizeof(particle_t) == sizeof(vpic_particle_t), "particle_t");
static_assert(sizeof(particle_mover_t) == sizeof(vpic_particle_mover_t), "pm");
static_assert(sizeof(interpolator_t) == sizeof(vpic_interpolator_t), "interp");
static_assert(sizeof(accumulator_t) == sizeof(vpic_accumulator_t), "acc");

void ref_fill_grid(grid_t *rg, const vpic_grid_t *g) {
  memset(rg, 0, sizeof(*rg));
  rg->dt = g->dt;
  rg->cvac = g->cvac;
  rg->eps0 = g->eps0;
  rg->nx = g->nx;
  rg->ny = g->ny;
  rg->nz = g->nz;
  rg->dx = g->dx;
  rg->dy = g->dy;
  rg->dz = g->dz;
  rg->dV = g->dV;
  rg->rdx = g->rdx;
  rg->rdy = g->rdy;
  rg->rdz = g->rdz;
  rg->r8V = g->r8V;
  rg->sx = g->sx;
  rg->sy = g->sy;
  rg->sz = g->sz;
  rg->nv = g->nv;
  rg->rangel = g->rangel;
  rg->rangeh = g->rangeh;
  rg->neighbor = (int64_t *)g->neighbor;
}

struct ref_push_job {
  advance_p_pipeline_args_t *args;
  int rank, n;
};

static void *ref_push_thread(void *v) {
  ref_push_job *j = (ref_push_job *)v;
  advance_p_pipeline_scalar(j->args, j->rank, j->n);
  return 0;
}

// advance_p with n_pipeline pipelines (advance_p_pipeline.cc:253-340); the
// accumulator copies and their reduction follow accumulator_array.cc:65-87
// and reduce_accumulator_array_pipeline (reduce_array_pipeline.cc:178-204).
extern "C" int ref_advance_p(vpic_particle_t *p, int np, float q, float m,
                             vpic_particle_mover_t *pm, int max_nm,
                             int *n_ignored, vpic_accumulator_t *a,
                             const vpic_interpolator_t *f,
                             const vpic_grid_t *g, int n_pipeline) {
  grid_t rg;
  ref_fill_grid(&rg, g);
  if (n_pipeline < 1) n_pipeline = 1;
  if (n_pipeline > MAX_PIPELINE) n_pipeline = MAX_PIPELINE;
  int stride = POW2_CEIL(rg.nv, 2);
  accumulator_t *acc = (accumulator_t *)aligned_alloc(
      128, sizeof(accumulator_t) * (size_t)stride * (size_t)(n_pipeline + 1) +
               128);
  memset(acc, 0, sizeof(accumulator_t) * (size_t)stride * (n_pipeline + 1));
  memcpy(acc, a, sizeof(accumulator_t) * (size_t)rg.nv);

  static particle_mover_seg_t seg[MAX_PIPELINE + 1];
  advance_p_pipeline_args_t args[1];
  args->p0 = (particle_t *)p;
  args->pm = (particle_mover_t *)pm;
  args->a0 = acc;
  args->f0 = (const interpolator_t *)f;
  args->seg = seg;
  args->g = &rg;
  args->qdt_2mc = (q * rg.dt) / (2 * m * rg.cvac);
  args->cdt_dx = rg.cvac * rg.dt * rg.rdx;
  args->cdt_dy = rg.cvac * rg.dt * rg.rdy;
  args->cdt_dz = rg.cvac * rg.dt * rg.rdz;
  args->qsp = q;
  args->np = np;
  args->max_nm = max_nm;
  args->nx = rg.nx;
  args->ny = rg.ny;
  args->nz = rg.nz;

  if (n_pipeline == 1) {
    advance_p_pipeline_scalar(args, 0, 1);
  } else {
    pthread_t th[MAX_PIPELINE];
    ref_push_job job[MAX_PIPELINE];
    for (int r = 0; r < n_pipeline; r++) {
      job[r].args = args;
      job[r].rank = r;
      job[r].n = n_pipeline;
      pthread_create(&th[r], 0, ref_push_thread, &job[r]);
    }
    advance_p_pipeline_scalar(args, n_pipeline, n_pipeline);
    for (int r = 0; r < n_pipeline; r++) pthread_join(th[r], 0);
  }
  if (n_pipeline == 1) advance_p_pipeline_scalar(args, 1, 1);

  // Mover compaction, advance_p_pipeline.cc:310-339.
  int nm = 0, ign = 0;
  for (int r = 0; r <= n_pipeline; r++) {
    ign += seg[r].n_ignored;
    if ((particle_mover_t *)pm + nm != seg[r].pm)
      memmove((particle_mover_t *)pm + nm, seg[r].pm,
              sizeof(particle_mover_t) * (size_t)seg[r].nm);
    nm += seg[r].nm;
  }
  if (n_ignored) *n_ignored = ign;

  // reduce_accumulator_array_pipeline, reduce_array_pipeline.cc:178-204.
  {
#define VOX_(x, y, z) VOXEL(x, y, z, rg.nx, rg.ny, rg.nz)
    reduce_pipeline_args_t ra[1];
    int i0 = (VOX_(1, 1, 1) / 2) * 2;
    int na = (((VOX_(rg.nx, rg.ny, rg.nz) - i0 + 1) + 1) / 2) * 2;
    int nfloats = sizeof(accumulator_t) / sizeof(float);
#undef VOX_
    ra->a = (float *)(acc + i0);
    ra->n = na * nfloats;
    ra->n_array = n_pipeline + 1;
    ra->s_array = stride * nfloats;
    ra->n_block = accumulators_n_block;
    reduce_array_pipeline_scalar(ra, 0, 1);
    reduce_array_pipeline_scalar(ra, 1, 1);
  }
  memcpy(a, acc, sizeof(accumulator_t) * (size_t)rg.nv);
  free(acc);
  return nm;
}

// The timed part of advance_p only (EXEC_PIPELINES + WAIT_PIPELINES + mover
// compaction, advance_p_pipeline.cc:272-339) on n_pipeline pthreads, with
// caller-owned accumulator copies ((n_pipeline+1) * POW2_CEIL(nv,2)
// entries, as new_accumulator_array allocates them).  Used as the CPU
// baseline; the accumulator reduction is a separate step in the reference
// (reduce_accumulator_array) and is not included.
extern "C" int ref_advance_p_pipelines(vpic_particle_t *p, int np, float q,
                                       float m, vpic_particle_mover_t *pm,
                                       int max_nm, vpic_accumulator_t *acc,
                                       const vpic_interpolator_t *f,
                                       const vpic_grid_t *g, int n_pipeline) {
  grid_t rg;
  ref_fill_grid(&rg, g);
  if (n_pipeline < 1) n_pipeline = 1;
  if (n_pipeline > MAX_PIPELINE) n_pipeline = MAX_PIPELINE;
  static particle_mover_seg_t seg[MAX_PIPELINE + 1];
  advance_p_pipeline_args_t args[1];
  args->p0 = (particle_t *)p;
  args->pm = (particle_mover_t *)pm;
  args->a0 = (accumulator_t *)acc;
  args->f0 = (const interpolator_t *)f;
  args->seg = seg;
  args->g = &rg;
  args->qdt_2mc = (q * rg.dt) / (2 * m * rg.cvac);
  args->cdt_dx = rg.cvac * rg.dt * rg.rdx;
  args->cdt_dy = rg.cvac * rg.dt * rg.rdy;
  args->cdt_dz = rg.cvac * rg.dt * rg.rdz;
  args->qsp = q;
  args->np = np;
  args->max_nm = max_nm;
  args->nx = rg.nx;
  args->ny = rg.ny;
  args->nz = rg.nz;
  pthread_t th[MAX_PIPELINE];
  ref_push_job job[MAX_PIPELINE];
  for (int r = 0; r < n_pipeline; r++) {
    job[r].args = args;
    job[r].rank = r;
    job[r].n = n_pipeline;
    pthread_create(&th[r], 0, ref_push_thread, &job[r]);
  }
  advance_p_pipeline_scalar(args, n_pipeline, n_pipeline);
  for (int r = 0; r < n_pipeline; r++) pthread_join(th[r], 0);
  int nm = 0;
  for (int r = 0; r <= n_pipeline; r++) {
    if ((particle_mover_t *)pm + nm != seg[r].pm)
      memmove((particle_mover_t *)pm + nm, seg[r].pm,
              sizeof(particle_mover_t) * (size_t)seg[r].nm);
    nm += seg[r].nm;
  }
  return nm;
}

extern "C" int ref_move_p(vpic_particle_t *p0, vpic_particle_mover_t *pm,
                          vpic_accumulator_t *a0, const vpic_grid_t *g,
                          float qsp) {
  grid_t rg;
  ref_fill_grid(&rg, g);
  return move_p((particle_t *)p0, (particle_mover_t *)pm,
                (accumulator_t *)a0, &rg, qsp);
}

// energy_p with one pipeline + host straggler (energy_p_pipeline.cc:76-115).
extern "C" double ref_energy_p(const vpic_particle_t *p, int np, float q,
                               float m, const vpic_interpolator_t *f,
                               const vpic_grid_t *g) {
  energy_p_pipeline_args_t args[1];
  double en[2];
  args->p = (const particle_t *)p;
  args->f = (const interpolator_t *)f;
  args->en = en;
  args->qdt_2mc = (q * g->dt) / (2 * m * g->cvac);
  args->msp = m;
  args->np = np;
  energy_p_pipeline_scalar(args, 0, 1);
  energy_p_pipeline_scalar(args, 1, 1);
  double local = 0.0;
  for (int r = 0; r <= 1; r++) local += en[r];
  return local * ((double)g->cvac * (double)g->cvac);
}

// center_p / uncenter_p with one pipeline + host straggler
// (center_p_pipeline.cc:105-130, uncenter_p_pipeline.cc:105-130).
extern "C" void ref_center_p(vpic_particle_t *p, int np, float q, float m,
                             const vpic_interpolator_t *f,
                             const vpic_grid_t *g, int uncenter) {
  center_p_pipeline_args_t args[1];
  args->p0 = (particle_t *)p;
  args->f0 = (const interpolator_t *)f;
  args->qdt_2mc = (q * g->dt) / (2 * m * g->cvac);
  args->np = np;
  for (int rank = 0; rank <= 1; rank++) {
    if (uncenter)
      uncenter_p_pipeline_scalar(args, rank, 1);
    else
      center_p_pipeline_scalar(args, rank, 1);
  }
}

// sort_p with one pipeline (sort_p_pipeline.cc:345-370).
extern "C" void ref_sort_p(vpic_particle_t *p, int np, int *partition,
                           const vpic_grid_t *g) {
  int nx = g->nx, ny = g->ny, nz = g->nz;
  int vl = VOXEL(1, 1, 1, nx, ny, nz);
  int vh = VOXEL(nx, ny, nz, nx, ny, nz);
  particle_t *aux = (particle_t *)malloc(sizeof(particle_t) * (size_t)(np + 1));
  int *next = (int *)calloc((size_t)g->nv, sizeof(int));
  int cp[2];
  sort_p_pipeline_args_t args[1];
  cp[0] = 0;
  cp[1] = np;
  args->p = aux;
  args->aux_p = (particle_t *)p;
  args->coarse_partition = cp;
  args->partition = partition;
  args->next = next;
  args->n = np;
  args->n_subsort = 1;
  args->vl = vl;
  args->vh = vh;
  args->n_voxel = g->nv;
  subsort_pipeline_scalar(args, 0, 1);
  memset(partition, 0, sizeof(int) * (size_t)vl);
  for (int i = vh + 1; i < g->nv; i++) partition[i] = np;
  memcpy(p, aux, sizeof(particle_t) * (size_t)np);
  free(aux);
  free(next);
}
