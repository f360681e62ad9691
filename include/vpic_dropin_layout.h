/*
 * vpic_dropin_layout.h -- the layout contract of the reference types that
 * the reference-named entry points take (include/vpic_dropin.h), as an
 * X-macro list: S(type) stands for sizeof(type), O(type, member) for
 * offsetof(type, member), in a fixed order.  A program compiled against
 * either set of headers expands the same list over its own definitions, so
 * two such tables compare member by member (vpic_hip_dropin_layout()).
 */
#ifndef VPIC_DROPIN_LAYOUT_H
#define VPIC_DROPIN_LAYOUT_H

#include <stddef.h>
#include <stdint.h>

#define VPIC_DROPIN_SIZE(T) (int64_t) sizeof(T),
#define VPIC_DROPIN_OFFSET(T, M) (int64_t) offsetof(T, M),

#define VPIC_DROPIN_LAYOUT(S, O)                                            \
  S(particle_t) O(particle_t, dx) O(particle_t, dy) O(particle_t, dz)       \
  O(particle_t, i) O(particle_t, ux) O(particle_t, uy) O(particle_t, uz)    \
  O(particle_t, w)                                                          \
  S(particle_mover_t) O(particle_mover_t, dispx) O(particle_mover_t, dispy) \
  O(particle_mover_t, dispz) O(particle_mover_t, i)                         \
  S(grid_t) O(grid_t, dt) O(grid_t, cvac) O(grid_t, eps0) O(grid_t, step)   \
  O(grid_t, t0) O(grid_t, x0) O(grid_t, y0) O(grid_t, z0) O(grid_t, x1)     \
  O(grid_t, y1) O(grid_t, z1) O(grid_t, nx) O(grid_t, ny) O(grid_t, nz)     \
  O(grid_t, dx) O(grid_t, dy) O(grid_t, dz) O(grid_t, dV) O(grid_t, rdx)    \
  O(grid_t, rdy) O(grid_t, rdz) O(grid_t, r8V) O(grid_t, sx) O(grid_t, sy)  \
  O(grid_t, sz) O(grid_t, nv) O(grid_t, bc) O(grid_t, range)                \
  O(grid_t, neighbor) O(grid_t, rangel) O(grid_t, rangeh) O(grid_t, mp)     \
  S(species_t) O(species_t, name) O(species_t, q) O(species_t, m)           \
  O(species_t, np) O(species_t, max_np) O(species_t, p) O(species_t, nm)    \
  O(species_t, max_nm) O(species_t, pm) O(species_t, last_sorted)           \
  O(species_t, sort_interval) O(species_t, sort_out_of_place)               \
  O(species_t, partition) O(species_t, g) O(species_t, id)                  \
  O(species_t, next)                                                        \
  S(interpolator_t) O(interpolator_t, ex) O(interpolator_t, dexdy)          \
  O(interpolator_t, dexdz) O(interpolator_t, d2exdydz)                      \
  O(interpolator_t, ey) O(interpolator_t, deydz) O(interpolator_t, deydx)   \
  O(interpolator_t, d2eydzdx) O(interpolator_t, ez)                         \
  O(interpolator_t, dezdx) O(interpolator_t, dezdy)                         \
  O(interpolator_t, d2ezdxdy) O(interpolator_t, cbx)                        \
  O(interpolator_t, dcbxdx) O(interpolator_t, cby) O(interpolator_t, dcbydy)\
  O(interpolator_t, cbz) O(interpolator_t, dcbzdz)                          \
  S(interpolator_array_t) O(interpolator_array_t, i)                        \
  O(interpolator_array_t, g)                                                \
  S(accumulator_t) O(accumulator_t, jx) O(accumulator_t, jy)                \
  O(accumulator_t, jz)                                                      \
  S(accumulator_array_t) O(accumulator_array_t, a)                          \
  O(accumulator_array_t, n_pipeline) O(accumulator_array_t, stride)         \
  O(accumulator_array_t, g)                                                 \
  S(particle_injector_t) O(particle_injector_t, i)                          \
  O(particle_injector_t, w) O(particle_injector_t, dispx)                   \
  O(particle_injector_t, sp_id)                                             \
  S(particle_bc_t) O(particle_bc_t, params) O(particle_bc_t, interact)      \
  O(particle_bc_t, delete_pbc) O(particle_bc_t, id) O(particle_bc_t, next)  \
  S(hydro_t) O(hydro_t, jx) O(hydro_t, rho) O(hydro_t, px) O(hydro_t, ke)   \
  O(hydro_t, txx) O(hydro_t, tyz) O(hydro_t, txy)                           \
  S(hydro_array_t) O(hydro_array_t, h) O(hydro_array_t, n_pipeline)         \
  O(hydro_array_t, stride) O(hydro_array_t, g)                              \
  S(field_t) O(field_t, ex) O(field_t, ey) O(field_t, ez)                   \
  O(field_t, div_e_err) O(field_t, cbx) O(field_t, cby) O(field_t, cbz)     \
  O(field_t, div_b_err) O(field_t, tcax) O(field_t, tcay) O(field_t, tcaz)  \
  O(field_t, rhob) O(field_t, jfx) O(field_t, jfy) O(field_t, jfz)          \
  O(field_t, rhof) O(field_t, ematx) O(field_t, ematy) O(field_t, ematz)    \
  O(field_t, nmat) O(field_t, fmatx) O(field_t, fmaty) O(field_t, fmatz)    \
  O(field_t, cmat)                                                          \
  S(field_advance_kernels_t) O(field_advance_kernels_t, delete_fa)          \
  O(field_advance_kernels_t, advance_b)                                     \
  O(field_advance_kernels_t, advance_e)                                     \
  O(field_advance_kernels_t, energy_f)                                      \
  O(field_advance_kernels_t, clear_jf)                                      \
  O(field_advance_kernels_t, synchronize_jf)                                \
  O(field_advance_kernels_t, clear_rhof)                                    \
  O(field_advance_kernels_t, synchronize_rho)                               \
  O(field_advance_kernels_t, compute_rhob)                                  \
  O(field_advance_kernels_t, compute_curl_b)                                \
  O(field_advance_kernels_t, synchronize_tang_e_norm_b)                     \
  O(field_advance_kernels_t, compute_div_e_err)                             \
  O(field_advance_kernels_t, compute_rms_div_e_err)                         \
  O(field_advance_kernels_t, clean_div_e)                                   \
  O(field_advance_kernels_t, compute_div_b_err)                             \
  O(field_advance_kernels_t, compute_rms_div_b_err)                         \
  O(field_advance_kernels_t, clean_div_b)                                   \
  S(field_array_t) O(field_array_t, f) O(field_array_t, g)                  \
  O(field_array_t, params) O(field_array_t, kernel)                         \
  S(material_coefficient_t) O(material_coefficient_t, decayx)               \
  O(material_coefficient_t, drivex) O(material_coefficient_t, decayy)       \
  O(material_coefficient_t, drivey) O(material_coefficient_t, decayz)       \
  O(material_coefficient_t, drivez) O(material_coefficient_t, rmux)         \
  O(material_coefficient_t, rmuy) O(material_coefficient_t, rmuz)           \
  O(material_coefficient_t, nonconductive) O(material_coefficient_t, epsx)  \
  O(material_coefficient_t, epsy) O(material_coefficient_t, epsz)           \
  S(sfa_params_t) O(sfa_params_t, mc) O(sfa_params_t, n_mc)                 \
  O(sfa_params_t, damp)

#endif /* VPIC_DROPIN_LAYOUT_H */
