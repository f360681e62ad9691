// ref_dccu.cc -- harness for the reference's vacuum_compute_curl_b_pipeline.cc (compiled here from
// the lanl/vpic source where it lies).  TEST INFRASTRUCTURE ONLY.
#include "src/field_advance/standard/pipeline/vacuum_compute_curl_b_pipeline.cc"

#include <math.h>

#include "ref_common.h"
#include "ref_field_args.h"

// The pipelined interior edges of vacuum_compute_curl_b
// (vacuum_compute_curl_b_pipeline.cc:66-74).
extern "C" void ref_curl_b_bulk(vpic_field_t *f, const vpic_grid_t *g,
                                const vpic_field_params_t *mp) {
  REF_FIELD_ARGS(f, g, mp);
  vacuum_compute_curl_b_pipeline_scalar(args, 0, 1);
  vacuum_compute_curl_b_pipeline_scalar(args, 1, 1);
}
