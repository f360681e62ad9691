// ref_dcce.cc -- harness for the reference's vacuum_clean_div_e_pipeline.cc (compiled here from
// the lanl/vpic source where it lies).  TEST INFRASTRUCTURE ONLY.
#include "src/field_advance/standard/pipeline/vacuum_clean_div_e_pipeline.cc"

#include <math.h>

#include "ref_common.h"
#include "ref_field_args.h"

// The pipelined edges of voxels 1..n of vacuum_clean_div_e
// (vacuum_clean_div_e_pipeline.cc:60-71).
extern "C" void ref_clean_div_e_bulk(vpic_field_t *f, const vpic_grid_t *g,
                                     const vpic_field_params_t *mp) {
  REF_FIELD_ARGS(f, g, mp);
  vacuum_clean_div_e_pipeline_scalar(args, 0, 1);
  vacuum_clean_div_e_pipeline_scalar(args, 1, 1);
}
