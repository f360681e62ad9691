// ref_fb.cc -- harness for the reference's advance_b_pipeline_scalar (bulk
// voxels 1..n of every component).  TEST INFRASTRUCTURE ONLY.
#define IN_sfa
#define IN_advance_b_pipeline
#include "src/field_advance/standard/pipeline/advance_b_pipeline.h"

#include "ref_common.h"

// The pipelined part of advance_b_pipeline (advance_b_pipeline.cc:63-69):
// pipeline 0 of 1 plus the straggler block.  The host's face leftovers
// and local_adjust_norm_b live in the driver and are not called.
extern "C" void ref_advance_b_bulk(vpic_field_t *f, const vpic_grid_t *g,
                                   float frac) {
  grid_t rg;
  ref_fill_grid(&rg, g);
  pipeline_args_t args[1];
  args->f = (field_t *)f;
  args->g = &rg;
  args->frac = frac;
  advance_b_pipeline_scalar(args, 0, 1);
  advance_b_pipeline_scalar(args, 1, 1);
}
