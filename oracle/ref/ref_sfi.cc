// ref_sfi.cc -- harness for the reference's sf_interface kernels:
// load_interpolator_pipeline_scalar and unload_accumulator_pipeline_scalar
// (built from lanl/vpic src/sf_interface).  TEST INFRASTRUCTURE ONLY.
#define IN_sf_interface
#include "src/sf_interface/pipeline/sf_interface_pipeline.h"

#include "ref_common.h"

static_assert(sizeof(field_t) == sizeof(vpic_field_t), "field_t");

// load_interpolator_array_pipeline's single-pipeline case: rank 0 of 1
// covers voxels 1..n (interpolator_array_pipeline.cc:21-50,256-354).
extern "C" void ref_load_interpolator(vpic_interpolator_t *fi,
                                      const vpic_field_t *f,
                                      const vpic_grid_t *g) {
  load_interpolator_pipeline_args_t args[1];
  args->fi = (interpolator_t *)fi;
  args->f = (const field_t *)f;
  args->nb = g->neighbor;
  args->nx = g->nx;
  args->ny = g->ny;
  args->nz = g->nz;
  load_interpolator_pipeline_scalar(args, 0, 1);
}

// unload_accumulator_array_pipeline (unload_accumulator_pipeline.cc:90-144).
extern "C" void ref_unload_accumulator(vpic_field_t *f,
                                       const vpic_accumulator_t *a,
                                       const vpic_grid_t *g) {
  unload_accumulator_pipeline_args_t args[1];
  args->f = (field_t *)f;
  args->a = (const accumulator_t *)a;
  args->nx = g->nx;
  args->ny = g->ny;
  args->nz = g->nz;
  args->cx = 0.25 * g->rdy * g->rdz / g->dt;
  args->cy = 0.25 * g->rdz * g->rdx / g->dt;
  args->cz = 0.25 * g->rdx * g->rdy / g->dt;
  unload_accumulator_pipeline_scalar(args, 0, 1);
}
