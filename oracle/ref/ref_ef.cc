// ref_ef.cc -- harness for the reference's vacuum_energy_f_pipeline_scalar.
// TEST INFRASTRUCTURE ONLY.
#define IN_sfa
#define IN_vacuum_energy_f_pipeline
#include "src/field_advance/standard/pipeline/vacuum_energy_f_pipeline.h"

#include "ref_common.h"

// vacuum_energy_f_pipeline with one pipeline + host straggler, then the
// reduction and unit conversion of vacuum_energy_f_pipeline.cc:67-96.
extern "C" void ref_energy_f(double en[6], const vpic_field_t *f,
                             const vpic_grid_t *g,
                             const vpic_field_params_t *mp) {
  grid_t rg;
  ref_fill_grid(&rg, g);
  material_coefficient_t mc[1];
  sfa_params_t sp[1];
  mc->rmux = mp->rmux;
  mc->rmuy = mp->rmuy;
  mc->rmuz = mp->rmuz;
  mc->epsx = mp->epsx;
  mc->epsy = mp->epsy;
  mc->epsz = mp->epsz;
  sp->mc = mc;
  sp->n_mc = 1;
  sp->damp = mp->damp;
  static pipeline_args_t args[1];
  args->f = (const field_t *)f;
  args->p = sp;
  args->g = &rg;
  vacuum_energy_f_pipeline_scalar(args, 0, 1);
  vacuum_energy_f_pipeline_scalar(args, 1, 1);
  for (int k = 0; k < 6; k++) args->en[0][k] += args->en[1][k];
  double v0 = 0.5 * rg.eps0 * rg.dV;
  for (int k = 0; k < 6; k++) en[k] = args->en[0][k] * v0;
}
