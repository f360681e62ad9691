// ref_dccb.cc -- harness for the reference's clean_div_b_pipeline.cc (compiled here from
// the lanl/vpic source where it lies).  TEST INFRASTRUCTURE ONLY.
#include "src/field_advance/standard/pipeline/clean_div_b_pipeline.cc"

#include <math.h>

#include "ref_common.h"
#include "ref_field_args.h"

// The pipelined interior faces 2..n of clean_div_b
// (clean_div_b_pipeline.cc:257-268).
extern "C" void ref_clean_div_b_bulk(vpic_field_t *f, const vpic_grid_t *g) {
  grid_t rg;
  ref_fill_grid(&rg, g);
  pipeline_args_t args[1];
  args->f = (field_t *)f;
  args->g = &rg;
  clean_div_b_pipeline_scalar(args, 0, 1);
  clean_div_b_pipeline_scalar(args, 1, 1);
}
