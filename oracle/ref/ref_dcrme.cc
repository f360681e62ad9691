// ref_dcrme.cc -- harness for the reference's compute_rms_div_e_err_pipeline.cc (compiled here from
// the lanl/vpic source where it lies).  TEST INFRASTRUCTURE ONLY.
#include "src/field_advance/standard/pipeline/compute_rms_div_e_err_pipeline.cc"

#include <math.h>

#include "ref_common.h"
#include "ref_field_args.h"

// The pipelines' FP64 partial sum of div_e_err^2 over the interior nodes
// 2..n of compute_rms_div_e_err (compute_rms_div_e_err_pipeline.cc:
// 90-96, 171-176).
extern "C" double ref_rms_div_e_err_bulk(const vpic_field_t *f,
                                         const vpic_grid_t *g) {
  grid_t rg;
  ref_fill_grid(&rg, g);
  pipeline_args_t args[1];
  args->f = (const field_t *)f;
  args->g = &rg;
  compute_rms_div_e_err_pipeline_scalar(args, 0, 1);
  compute_rms_div_e_err_pipeline_scalar(args, 1, 1);
  return args->err[0] + args->err[1];
}
