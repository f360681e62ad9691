// ref_dcde.cc -- harness for the reference's vacuum_compute_div_e_err_pipeline.cc (compiled here from
// the lanl/vpic source where it lies).  TEST INFRASTRUCTURE ONLY.
#include "src/field_advance/standard/pipeline/vacuum_compute_div_e_err_pipeline.cc"

#include <math.h>

#include "ref_common.h"
#include "ref_field_args.h"

// The pipelined interior nodes 2..n of vacuum_compute_div_e_err
// (vacuum_compute_div_e_err_pipeline.cc:87-97).
extern "C" void ref_div_e_err_bulk(vpic_field_t *f, const vpic_grid_t *g,
                                   const vpic_field_params_t *mp) {
  REF_FIELD_ARGS(f, g, mp);
  vacuum_compute_div_e_err_pipeline_scalar(args, 0, 1);
  vacuum_compute_div_e_err_pipeline_scalar(args, 1, 1);
}
