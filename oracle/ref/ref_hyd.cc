// ref_hyd.cc -- harness for the reference's accumulate_hydro_p_pipeline_scalar
// (hydro_p_pipeline.cc, compiled here from the lanl/vpic source where it
// lies).  TEST INFRASTRUCTURE ONLY.
#include "src/species_advance/standard/pipeline/hydro_p_pipeline.cc"

#include <string.h>

#include "ref_common.h"

// accumulate_hydro_p_pipeline with one pipeline: rank 0 into h, the
// straggler rank into a cleared second copy, then the O1 reduction of
// reduce_array_pipeline.cc (h = h + copy 1) (hydro_p_pipeline.cc:214-242).
extern "C" void ref_accumulate_hydro_p(vpic_hydro_t *h,
                                       const vpic_particle_t *p, int np,
                                       float q, float m,
                                       const vpic_interpolator_t *f,
                                       const vpic_grid_t *g) {
  grid_t rg;
  ref_fill_grid(&rg, g);
  species_t sp;
  memset(&sp, 0, sizeof(sp));
  sp.q = q;
  sp.m = m;
  sp.p = (particle_t *)p;
  sp.np = np;
  sp.g = &rg;
  const int nv = g->nv;
  hydro_t *buf = (hydro_t *)calloc(2 * (size_t)nv, sizeof(hydro_t));
  memcpy(buf, h, sizeof(hydro_t) * (size_t)nv);
  accumulate_hydro_p_pipeline_args_t args[1];
  args->sp = &sp;
  args->f = (const interpolator_t *)f;
  args->h = buf;
  args->h_size = nv;
  args->qdt_2mc = (sp.q * rg.dt) / (2 * sp.m * rg.cvac);
  args->msp = sp.m;
  args->np = np;
  accumulate_hydro_p_pipeline_scalar(args, 0, 1);
  accumulate_hydro_p_pipeline_scalar(args, 1, 1);
  for (int v = 0; v < nv; v++) {
    float *a = (float *)(buf + v);
    const float *b = (const float *)(buf + nv + v);
    for (int k = 0; k < 14; k++) a[k] = a[k] + b[k];
  }
  memcpy(h, buf, sizeof(hydro_t) * (size_t)nv);
  free(buf);
}
