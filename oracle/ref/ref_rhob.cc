// ref_rhob.cc -- harness for the reference's accumulate_rhob (rho_p.cc,
// compiled here from the lanl/vpic source where it lies; the file's other
// function, accumulate_rho_p, is dropped by --gc-sections).  TEST
// INFRASTRUCTURE ONLY.
#include "src/species_advance/standard/rho_p.cc"

#include <string.h>

#include "ref_common.h"

// accumulate_rhob (rho_p.cc:126-173) of each particle in turn, in index
// order, as boundary_p calls it for the absorbed particles
// (boundary_p.cc:270-277, there in reverse mover order).
extern "C" void ref_accumulate_rhob(vpic_field_t *f, const vpic_particle_t *p,
                                    int np, float qsp, const vpic_grid_t *g) {
  grid_t rg;
  ref_fill_grid(&rg, g);
  for (int n = 0; n < np; n++)
    accumulate_rhob((field_t *)f, (const particle_t *)(p + n), &rg, qsp);
}
