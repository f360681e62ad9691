// ref_dcrmb.cc -- harness for the reference's compute_rms_div_b_err_pipeline.cc (compiled here from
// the lanl/vpic source where it lies).  TEST INFRASTRUCTURE ONLY.
#include "src/field_advance/standard/pipeline/compute_rms_div_b_err_pipeline.cc"

#include <math.h>

#include "ref_common.h"
#include "ref_field_args.h"

// compute_rms_div_b_err for one domain: the pipelines' FP64 partial sums
// over every cell, then the reduction of compute_rms_div_b_err_pipeline.cc:
// 93-110 (mp_allsum_d over one rank is the identity).
extern "C" double ref_rms_div_b_err(const vpic_field_t *f,
                                    const vpic_grid_t *g) {
  grid_t rg;
  ref_fill_grid(&rg, g);
  pipeline_args_t args[1];
  args->f = (const field_t *)f;
  args->g = &rg;
  compute_rms_div_b_err_pipeline_scalar(args, 0, 1);
  compute_rms_div_b_err_pipeline_scalar(args, 1, 1);
  double err = 0, local[2];
  for (int p = 0; p <= 1; p++) err += args->err[p];
  local[0] = err * rg.dV;
  local[1] = (rg.nx * rg.ny * rg.nz) * rg.dV;
  return rg.eps0 * sqrt(local[0] / local[1]);
}
