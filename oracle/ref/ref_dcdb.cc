// ref_dcdb.cc -- harness for the reference's compute_div_b_err_pipeline.cc (compiled here from
// the lanl/vpic source where it lies).  TEST INFRASTRUCTURE ONLY.
#include "src/field_advance/standard/pipeline/compute_div_b_err_pipeline.cc"

#include <math.h>

#include "ref_common.h"
#include "ref_field_args.h"

// compute_div_b_err: the pipelines cover every cell 1..n
// (compute_div_b_err_pipeline.cc:287-293).
extern "C" void ref_div_b_err(vpic_field_t *f, const vpic_grid_t *g) {
  grid_t rg;
  ref_fill_grid(&rg, g);
  pipeline_args_t args[1];
  args->f = (field_t *)f;
  args->g = &rg;
  compute_div_b_err_pipeline_scalar(args, 0, 1);
  compute_div_b_err_pipeline_scalar(args, 1, 1);
}
