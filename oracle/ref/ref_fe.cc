// ref_fe.cc -- harness for the reference's vacuum_advance_e_pipeline_scalar
// (interior voxels 2..n).  TEST INFRASTRUCTURE ONLY.
#define IN_sfa
#define IN_vacuum_advance_e_pipeline
#include "src/field_advance/standard/pipeline/vacuum_advance_e_pipeline.h"

#include "ref_common.h"

// The pipelined interior of vacuum_advance_e_pipeline
// (vacuum_advance_e_pipeline.cc:79-87).
extern "C" void ref_vacuum_advance_e_bulk(vpic_field_t *f,
                                          const vpic_grid_t *g,
                                          const vpic_field_params_t *mp) {
  grid_t rg;
  ref_fill_grid(&rg, g);
  material_coefficient_t mc[1];
  sfa_params_t sp[1];
  mc->decayx = mp->decayx;
  mc->drivex = mp->drivex;
  mc->decayy = mp->decayy;
  mc->drivey = mp->drivey;
  mc->decayz = mp->decayz;
  mc->drivez = mp->drivez;
  mc->rmux = mp->rmux;
  mc->rmuy = mp->rmuy;
  mc->rmuz = mp->rmuz;
  mc->nonconductive = 1;
  mc->epsx = mp->epsx;
  mc->epsy = mp->epsy;
  mc->epsz = mp->epsz;
  sp->mc = mc;
  sp->n_mc = 1;
  sp->damp = mp->damp;
  pipeline_args_t args[1];
  args->f = (field_t *)f;
  args->p = sp;
  args->g = &rg;
  vacuum_advance_e_pipeline_scalar(args, 0, 1);
  vacuum_advance_e_pipeline_scalar(args, 1, 1);
}
