// ref_field_args.h -- the sfa_params_t set-up shared by the field harnesses
// (sfa.cc:100-148 holds one material coefficient per material; the vacuum
// kernels read mc[0]).  TEST INFRASTRUCTURE ONLY.
#ifndef VPIC_REF_FIELD_ARGS_H
#define VPIC_REF_FIELD_ARGS_H

static inline void ref_fill_mc(material_coefficient_t *mc,
                               const vpic_field_params_t *mp) {
  mc->decayx = mp->decayx;
  mc->drivex = mp->drivex;
  mc->decayy = mp->decayy;
  mc->drivey = mp->drivey;
  mc->decayz = mp->decayz;
  mc->drivez = mp->drivez;
  mc->rmux = mp->rmux;
  mc->rmuy = mp->rmuy;
  mc->rmuz = mp->rmuz;
  mc->nonconductive = mp->nonconductive;
  mc->epsx = mp->epsx;
  mc->epsy = mp->epsy;
  mc->epsz = mp->epsz;
}

#define REF_FIELD_ARGS(F, G, MP)           \
  grid_t rg;                               \
  ref_fill_grid(&rg, G);                   \
  material_coefficient_t mc[1];            \
  sfa_params_t sp[1];                      \
  ref_fill_mc(mc, MP);                     \
  sp->mc = mc;                             \
  sp->n_mc = 1;                            \
  sp->damp = (MP)->damp;                   \
  pipeline_args_t args[1];                 \
  args->f = (field_t *)(F);                \
  args->p = sp;                            \
  args->g = &rg

#endif
