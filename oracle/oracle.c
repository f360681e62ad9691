/*
 * oracle.c -- CPU restatement of VPIC's per-step hot path (see oracle.h).
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the device path, never part of
 * it.  Build with -O2 -ffp-contract=off (no fast-math) so every FP32
 * expression rounds exactly as the reference's scalar build does.
 *
 * Parity of each function is pinned in tests/ by (a) the reference's own
 * scalar kernels compiled from source into oracle/_ref (bit-exact, where
 * they build), (b) the reference's known-answer decks (accel, cyclo,
 * interpe, inbndj/outbndj charge conservation), (c) closed forms and
 * brute force and (d) for the field walls (local.cc's ghost and adjust
 * branches), the method of images: a walled box against the periodic box
 * of twice its length with mirrored fields; the pin of each function is
 * listed in tests/test_oracle_*.  absorb_fields walls are unpinned (the
 * device path does not support them).  orc_takizuka_abe draws from the
 * device's Philox stream, not the reference's SFMT generator: its
 * arithmetic is pinned by conservation and scattering statistics, its
 * parity with the reference's outputs is unpinned.
 */
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

#define VOX(x, y, z, nx, ny) ((x) + ((nx) + 2) * ((y) + ((ny) + 2) * (z)))

/* DISTRIBUTE, src/util/util_base.h:191-197. */
static void distribute(int N, int b, int p, int P, int *i, int *n) {
  double t = (double)(N / b) / (double)P;
  int i0 = b * (int)(t * (double)p + 0.5);
  *n = (p == P) ? (N % b) : (b * (int)(t * (double)(p + 1) + 0.5) - i0);
  *i = i0;
}

/* ------------------------------------------------------------------------ */
/* Grid                                                                     */
/* ------------------------------------------------------------------------ */

void orc_grid_periodic(vpic_grid_t *g, double gx0, double gy0, double gz0,
                       double gx1, double gy1, double gz1, int gnx, int gny,
                       int gnz, float dt, float cvac, float eps0) {
  memset(g, 0, sizeof(*g));
  g->dt = dt;
  g->cvac = cvac;
  g->eps0 = eps0;
  /* partition.cc:58-70: computed in double, stored in float. */
  g->dx = (float)((gx1 - gx0) / (double)gnx);
  g->dy = (float)((gy1 - gy0) / (double)gny);
  g->dz = (float)((gz1 - gz0) / (double)gnz);
  g->dV = (float)(((gx1 - gx0) / (double)gnx) * ((gy1 - gy0) / (double)gny) *
                  ((gz1 - gz0) / (double)gnz));
  g->rdx = (float)((double)gnx / (gx1 - gx0));
  g->rdy = (float)((double)gny / (gy1 - gy0));
  g->rdz = (float)((double)gnz / (gz1 - gz0));
  g->r8V = (float)(((double)gnx / (gx1 - gx0)) * ((double)gny / (gy1 - gy0)) *
                   ((double)gnz / (gz1 - gz0)) * 0.125);
  /* ops.cc:25-30 */
  g->nx = gnx;
  g->ny = gny;
  g->nz = gnz;
  g->sx = 1;
  g->sy = (gnx + 2);
  g->sz = (gny + 2) * g->sy;
  g->nv = (gnz + 2) * g->sz;
  g->rangel = 0;
  g->rangeh = g->nv - 1;
  g->neighbor = 0;
  for (int k = 0; k < 6; k++) {
    g->pbc[k] = VPIC_BC_PERIODIC;
    g->fbc[k] = VPIC_BC_PERIODIC;
  }
}

void orc_neighbor_table(const vpic_grid_t *g, int64_t *nb) {
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const int64_t rl = g->rangel;
  /* ops.cc:45-76: interior links, domain faces reflect, ghosts reflect. */
  for (int z = 0; z <= nz + 1; z++)
    for (int y = 0; y <= ny + 1; y++)
      for (int x = 0; x <= nx + 1; x++) {
        int64_t *n = nb + 6 * (int64_t)VOX(x, y, z, nx, ny);
        n[0] = rl + VOX(x - 1, y, z, nx, ny);
        n[1] = rl + VOX(x, y - 1, z, nx, ny);
        n[2] = rl + VOX(x, y, z - 1, nx, ny);
        n[3] = rl + VOX(x + 1, y, z, nx, ny);
        n[4] = rl + VOX(x, y + 1, z, nx, ny);
        n[5] = rl + VOX(x, y, z + 1, nx, ny);
        if (x == 1) n[0] = VPIC_BC_REFLECT_PARTICLES;
        if (y == 1) n[1] = VPIC_BC_REFLECT_PARTICLES;
        if (z == 1) n[2] = VPIC_BC_REFLECT_PARTICLES;
        if (x == nx) n[3] = VPIC_BC_REFLECT_PARTICLES;
        if (y == ny) n[4] = VPIC_BC_REFLECT_PARTICLES;
        if (z == nz) n[5] = VPIC_BC_REFLECT_PARTICLES;
        if (x == 0 || x == nx + 1 || y == 0 || y == ny + 1 || z == 0 ||
            z == nz + 1)
          for (int k = 0; k < 6; k++) n[k] = VPIC_BC_REFLECT_PARTICLES;
      }
  /* join_grid to this same rank (ops.cc:119-165; periodic wrap to the
   * opposite face) or set_pbc (ops.cc:184-211) for local particle BCs. */
  for (int z = 1; z <= nz; z++)
    for (int y = 1; y <= ny; y++)
      for (int x = 1; x <= nx; x++) {
        int64_t *n = nb + 6 * (int64_t)VOX(x, y, z, nx, ny);
        if (x == 1)
          n[0] = g->pbc[0] == VPIC_BC_PERIODIC ? rl + VOX(nx, y, z, nx, ny)
                                               : g->pbc[0];
        if (y == 1)
          n[1] = g->pbc[1] == VPIC_BC_PERIODIC ? rl + VOX(x, ny, z, nx, ny)
                                               : g->pbc[1];
        if (z == 1)
          n[2] = g->pbc[2] == VPIC_BC_PERIODIC ? rl + VOX(x, y, nz, nx, ny)
                                               : g->pbc[2];
        if (x == nx)
          n[3] = g->pbc[3] == VPIC_BC_PERIODIC ? rl + VOX(1, y, z, nx, ny)
                                               : g->pbc[3];
        if (y == ny)
          n[4] = g->pbc[4] == VPIC_BC_PERIODIC ? rl + VOX(x, 1, z, nx, ny)
                                               : g->pbc[4];
        if (z == nz)
          n[5] = g->pbc[5] == VPIC_BC_PERIODIC ? rl + VOX(x, y, 1, nx, ny)
                                               : g->pbc[5];
      }
}

/* ------------------------------------------------------------------------ */
/* move_p (move_p.cc:215-378)                                               */
/* ------------------------------------------------------------------------ */

/* Deposit of one streak; the twelve updates of move_p.cc:281-303 and
 * advance_p_pipeline.cc:186-208 (identical arithmetic). */
#define ORC_ACC(a4, qd, mY, mZ, v5)     \
  do {                                  \
    float v0_, v1_, v2_, v3_, v4_;      \
    v4_ = (qd);                         \
    v1_ = v4_ * (mY);                   \
    v0_ = v4_ - v1_;                    \
    v1_ += v4_;                         \
    v4_ = 1 + (mZ);                     \
    v2_ = v0_ * v4_;                    \
    v3_ = v1_ * v4_;                    \
    v4_ = 1 - (mZ);                     \
    v0_ *= v4_;                         \
    v1_ *= v4_;                         \
    v0_ += (v5);                        \
    v1_ -= (v5);                        \
    v2_ -= (v5);                        \
    v3_ += (v5);                        \
    (a4)[0] += v0_;                     \
    (a4)[1] += v1_;                     \
    (a4)[2] += v2_;                     \
    (a4)[3] += v3_;                     \
  } while (0)

int orc_move_p(vpic_particle_t *p0, vpic_particle_mover_t *pm,
               vpic_accumulator_t *a0, const vpic_grid_t *g, float qsp) {
  float s_midx, s_midy, s_midz, s_dispx, s_dispy, s_dispz, s_dir[3];
  float v0, v1, v2, v3, v5, q;
  int axis, face;
  int64_t neighbor;
  float *a;
  vpic_particle_t *p = p0 + pm->i;

  q = qsp * p->w;
  for (;;) {
    s_midx = p->dx;
    s_midy = p->dy;
    s_midz = p->dz;
    s_dispx = pm->dispx;
    s_dispy = pm->dispy;
    s_dispz = pm->dispz;
    s_dir[0] = (s_dispx > 0.0f) ? 1.0f : -1.0f;
    s_dir[1] = (s_dispy > 0.0f) ? 1.0f : -1.0f;
    s_dir[2] = (s_dispz > 0.0f) ? 1.0f : -1.0f;
    v0 = (s_dispx == 0.0f) ? 3.4e38f : (s_dir[0] - s_midx) / s_dispx;
    v1 = (s_dispy == 0.0f) ? 3.4e38f : (s_dir[1] - s_midy) / s_dispy;
    v2 = (s_dispz == 0.0f) ? 3.4e38f : (s_dir[2] - s_midz) / s_dispz;
    v3 = 2.0f, axis = 3;
    if (v0 < v3) v3 = v0, axis = 0;
    if (v1 < v3) v3 = v1, axis = 1;
    if (v2 < v3) v3 = v2, axis = 2;
    v3 *= 0.5f;
    s_dispx *= v3;
    s_dispy *= v3;
    s_dispz *= v3;
    s_midx += s_dispx;
    s_midy += s_dispy;
    s_midz += s_dispz;
    /* move_p.cc:277: the product is promoted to double by 1.0/3.0. */
    v5 = (float)((double)(q * s_dispx * s_dispy * s_dispz) * (1.0 / 3.0));
    a = (float *)(a0 + p->i);
    ORC_ACC(a + 0, q * s_dispx, s_midy, s_midz, v5);
    ORC_ACC(a + 4, q * s_dispy, s_midz, s_midx, v5);
    ORC_ACC(a + 8, q * s_dispz, s_midx, s_midy, v5);
    pm->dispx -= s_dispx;
    pm->dispy -= s_dispy;
    pm->dispz -= s_dispz;
    p->dx += s_dispx + s_dispx;
    p->dy += s_dispy + s_dispy;
    p->dz += s_dispz + s_dispz;
    if (axis == 3) break;
    v0 = s_dir[axis];
    (&(p->dx))[axis] = v0;
    face = axis;
    if (v0 > 0.0f) face += 3;
    neighbor = g->neighbor[6 * (int64_t)p->i + face];
    if (neighbor == VPIC_BC_REFLECT_PARTICLES) {
      (&(p->ux))[axis] = -(&(p->ux))[axis];
      (&(pm->dispx))[axis] = -(&(pm->dispx))[axis];
      continue;
    }
    if (neighbor < g->rangel || neighbor > g->rangeh) {
      p->i = 8 * p->i + face;
      return 1;
    }
    p->i = (int32_t)(neighbor - g->rangel);
    (&(p->dx))[axis] = -v0;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* advance_p (advance_p_pipeline.cc:21-245,253-340)                         */
/* ------------------------------------------------------------------------ */

void orc_push_coefficients(float q, float m, const vpic_grid_t *g,
                           float out[4]) {
  out[0] = (q * g->dt) / (2 * m * g->cvac);
  out[1] = g->cvac * g->dt * g->rdx;
  out[2] = g->cvac * g->dt * g->rdy;
  out[3] = g->cvac * g->dt * g->rdz;
}

typedef struct {
  vpic_particle_t *p0;
  vpic_particle_mover_t *pm;
  vpic_accumulator_t *a0;
  const vpic_interpolator_t *f0;
  const vpic_grid_t *g;
  float qdt_2mc, cdt_dx, cdt_dy, cdt_dz, qsp;
  int np, max_nm;
  int64_t stride;
  /* per-pipeline results (index = pipeline rank) */
  vpic_particle_mover_t *seg_pm[257];
  int seg_nm[257], seg_ignored[257];
} orc_push_args_t;

static void orc_push_pipeline(orc_push_args_t *args, int rank, int P) {
  const float qdt_2mc = args->qdt_2mc, cdt_dx = args->cdt_dx,
              cdt_dy = args->cdt_dy, cdt_dz = args->cdt_dz, qsp = args->qsp;
  const float one = 1.0, one_third = 1.0 / 3.0, two_fifteenths = 2.0 / 15.0;
  vpic_particle_t *p0 = args->p0, *p;
  vpic_accumulator_t *a0 = args->a0;
  vpic_particle_mover_t *pm, local_pm;
  float dx, dy, dz, ux, uy, uz, q, hax, hay, haz, cbx, cby, cbz;
  float v0, v1, v2, v3, v4, v5;
  int ii, itmp, n, nm, max_nm;

  distribute(args->np, 16, rank, P, &itmp, &n);
  p = p0 + itmp;
  max_nm = args->max_nm - (args->np & 15);
  if (max_nm < 0) max_nm = 0;
  distribute(max_nm, 8, rank, P, &itmp, &max_nm);
  if (rank == P) max_nm = args->max_nm - itmp;
  pm = args->pm + itmp;
  nm = 0;
  itmp = 0;
  if (rank != P) a0 += (int64_t)(1 + rank) * args->stride;

  for (; n; n--, p++) {
    const vpic_interpolator_t *f;
    dx = p->dx;
    dy = p->dy;
    dz = p->dz;
    ii = p->i;
    f = args->f0 + ii;
    hax = qdt_2mc * ((f->ex + dy * f->dexdy) + dz * (f->dexdz + dy * f->d2exdydz));
    hay = qdt_2mc * ((f->ey + dz * f->deydz) + dx * (f->deydx + dz * f->d2eydzdx));
    haz = qdt_2mc * ((f->ez + dx * f->dezdx) + dy * (f->dezdy + dx * f->d2ezdxdy));
    cbx = f->cbx + dx * f->dcbxdx;
    cby = f->cby + dy * f->dcbydy;
    cbz = f->cbz + dz * f->dcbzdz;
    ux = p->ux;
    uy = p->uy;
    uz = p->uz;
    q = p->w;
    ux += hax;
    uy += hay;
    uz += haz;
    v0 = qdt_2mc / sqrtf(one + (ux * ux + (uy * uy + uz * uz)));
    v1 = cbx * cbx + (cby * cby + cbz * cbz);
    v2 = (v0 * v0) * v1;
    v3 = v0 * (one + v2 * (one_third + v2 * two_fifteenths));
    v4 = v3 / (one + v1 * (v3 * v3));
    v4 += v4;
    v0 = ux + v3 * (uy * cbz - uz * cby);
    v1 = uy + v3 * (uz * cbx - ux * cbz);
    v2 = uz + v3 * (ux * cby - uy * cbx);
    ux += v4 * (v1 * cbz - v2 * cby);
    uy += v4 * (v2 * cbx - v0 * cbz);
    uz += v4 * (v0 * cby - v1 * cbx);
    ux += hax;
    uy += hay;
    uz += haz;
    p->ux = ux;
    p->uy = uy;
    p->uz = uz;
    v0 = one / sqrtf(one + (ux * ux + (uy * uy + uz * uz)));
    ux *= cdt_dx;
    uy *= cdt_dy;
    uz *= cdt_dz;
    ux *= v0;
    uy *= v0;
    uz *= v0;
    v0 = dx + ux;
    v1 = dy + uy;
    v2 = dz + uz;
    v3 = v0 + ux;
    v4 = v1 + uy;
    v5 = v2 + uz;
    if (v3 <= one && v4 <= one && v5 <= one && -v3 <= one && -v4 <= one &&
        -v5 <= one) {
      float *a;
      q *= qsp;
      p->dx = v3;
      p->dy = v4;
      p->dz = v5;
      dx = v0;
      dy = v1;
      dz = v2;
      v5 = q * ux * uy * uz * one_third;
      a = (float *)(a0 + ii);
      ORC_ACC(a + 0, q * ux, dy, dz, v5);
      ORC_ACC(a + 4, q * uy, dz, dx, v5);
      ORC_ACC(a + 8, q * uz, dx, dy, v5);
    } else {
      local_pm.dispx = ux;
      local_pm.dispy = uy;
      local_pm.dispz = uz;
      local_pm.i = (int32_t)(p - p0);
      if (orc_move_p(p0, &local_pm, a0, args->g, qsp)) {
        if (nm < max_nm)
          pm[nm++] = local_pm;
        else {
          itmp++;
          p->i = p->i >> 3;
        }
      }
    }
  }
  args->seg_pm[rank] = pm;
  args->seg_nm[rank] = nm;
  args->seg_ignored[rank] = itmp;
}

/* Mover compaction of advance_p_pipeline.cc:310-339. */
static int orc_push_compact(orc_push_args_t *args, int P, int *n_ignored) {
  int nm = 0, ign = 0;
  for (int r = 0; r <= P; r++) {
    ign += args->seg_ignored[r];
    if (args->pm + nm != args->seg_pm[r])
      memmove(args->pm + nm, args->seg_pm[r],
              sizeof(vpic_particle_mover_t) * (size_t)args->seg_nm[r]);
    nm += args->seg_nm[r];
  }
  if (n_ignored) *n_ignored = ign;
  return nm;
}

int orc_advance_p(vpic_particle_t *p, int np, float q, float m,
                  vpic_particle_mover_t *pm, int max_nm, int *n_ignored,
                  vpic_accumulator_t *a, const vpic_interpolator_t *f,
                  const vpic_grid_t *g) {
  float c[4];
  orc_push_args_t args;
  int64_t nv = g->nv;
  vpic_accumulator_t *acc =
      (vpic_accumulator_t *)calloc((size_t)(2 * nv), sizeof(*acc));
  float *a0, *a1;
  int nm;

  orc_push_coefficients(q, m, g, c);
  memset(&args, 0, sizeof(args));
  args.p0 = p;
  args.pm = pm;
  args.a0 = acc;
  args.f0 = f;
  args.g = g;
  args.qdt_2mc = c[0];
  args.cdt_dx = c[1];
  args.cdt_dy = c[2];
  args.cdt_dz = c[3];
  args.qsp = q;
  args.np = np;
  args.max_nm = max_nm;
  args.stride = nv;
  /* Copy 0 starts from the caller's accumulator (the host pipeline's
   * copy is a->a[0] in the reference). */
  memcpy(acc, a, sizeof(*acc) * (size_t)nv);
  orc_push_pipeline(&args, 0, 1); /* pipeline 0 -> copy 1 */
  orc_push_pipeline(&args, 1, 1); /* host tail   -> copy 0 */
  nm = orc_push_compact(&args, 1, n_ignored);
  /* reduce_array_pipeline.cc O1: a0 = a0 + a1. */
  a0 = (float *)acc;
  a1 = (float *)(acc + nv);
  for (int64_t k = 0; k < 12 * nv; k++) a0[k] = a0[k] + a1[k];
  memcpy(a, acc, sizeof(*acc) * (size_t)nv);
  free(acc);
  return nm;
}

/* Threaded CPU baseline ------------------------------------------------- */

typedef struct {
  orc_push_args_t *args;
  int rank, P;
} orc_thread_job_t;

static void *orc_thread_main(void *v) {
  orc_thread_job_t *j = (orc_thread_job_t *)v;
  orc_push_pipeline(j->args, j->rank, j->P);
  return 0;
}

int orc_advance_p_threaded(vpic_particle_t *p, int np, float q, float m,
                           vpic_particle_mover_t *pm, int max_nm,
                           vpic_accumulator_t *acc,
                           const vpic_interpolator_t *f, const vpic_grid_t *g,
                           int n_thread) {
  float c[4];
  static orc_push_args_t args;
  pthread_t th[256];
  orc_thread_job_t job[256];
  int64_t nv = g->nv;
  if (n_thread < 1) n_thread = 1;
  if (n_thread > 256) n_thread = 256;
  orc_push_coefficients(q, m, g, c);
  memset(&args, 0, sizeof(args));
  args.p0 = p;
  args.pm = pm;
  args.a0 = acc;
  args.f0 = f;
  args.g = g;
  args.qdt_2mc = c[0];
  args.cdt_dx = c[1];
  args.cdt_dy = c[2];
  args.cdt_dz = c[3];
  args.qsp = q;
  args.np = np;
  args.max_nm = max_nm;
  args.stride = nv;
  for (int r = 0; r < n_thread; r++) {
    job[r].args = &args;
    job[r].rank = r;
    job[r].P = n_thread;
    pthread_create(&th[r], 0, orc_thread_main, &job[r]);
  }
  orc_push_pipeline(&args, n_thread, n_thread);
  for (int r = 0; r < n_thread; r++) pthread_join(th[r], 0);
  {
    float *a0 = (float *)acc;
    for (int r = 1; r <= n_thread; r++) {
      const float *ar = (const float *)(acc + (int64_t)r * nv);
      for (int64_t k = 0; k < 12 * nv; k++) a0[k] += ar[k];
    }
  }
  return orc_push_compact(&args, n_thread, 0);
}

/* ------------------------------------------------------------------------ */
/* Interpolator / accumulator (sf_interface)                                */
/* ------------------------------------------------------------------------ */

void orc_load_interpolator(vpic_interpolator_t *fi, const vpic_field_t *f,
                           const vpic_grid_t *g) {
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const float fourth = 0.25, half = 0.50;
  float w0, w1, w2, w3;
  for (int z = 1; z <= nz; z++)
    for (int y = 1; y <= ny; y++)
      for (int x = 1; x <= nx; x++) {
        vpic_interpolator_t *pi = fi + VOX(x, y, z, nx, ny);
        const vpic_field_t *pf0 = f + VOX(x, y, z, nx, ny);
        const vpic_field_t *pfx = f + VOX(x + 1, y, z, nx, ny);
        const vpic_field_t *pfy = f + VOX(x, y + 1, z, nx, ny);
        const vpic_field_t *pfz = f + VOX(x, y, z + 1, nx, ny);
        const vpic_field_t *pfyz = f + VOX(x, y + 1, z + 1, nx, ny);
        const vpic_field_t *pfzx = f + VOX(x + 1, y, z + 1, nx, ny);
        const vpic_field_t *pfxy = f + VOX(x + 1, y + 1, z, nx, ny);
        w0 = pf0->ex; w1 = pfy->ex; w2 = pfz->ex; w3 = pfyz->ex;
        pi->ex = fourth * ((w3 + w0) + (w1 + w2));
        pi->dexdy = fourth * ((w3 - w0) + (w1 - w2));
        pi->dexdz = fourth * ((w3 - w0) - (w1 - w2));
        pi->d2exdydz = fourth * ((w3 + w0) - (w1 + w2));
        w0 = pf0->ey; w1 = pfz->ey; w2 = pfx->ey; w3 = pfzx->ey;
        pi->ey = fourth * ((w3 + w0) + (w1 + w2));
        pi->deydz = fourth * ((w3 - w0) + (w1 - w2));
        pi->deydx = fourth * ((w3 - w0) - (w1 - w2));
        pi->d2eydzdx = fourth * ((w3 + w0) - (w1 + w2));
        w0 = pf0->ez; w1 = pfx->ez; w2 = pfy->ez; w3 = pfxy->ez;
        pi->ez = fourth * ((w3 + w0) + (w1 + w2));
        pi->dezdx = fourth * ((w3 - w0) + (w1 - w2));
        pi->dezdy = fourth * ((w3 - w0) - (w1 - w2));
        pi->d2ezdxdy = fourth * ((w3 + w0) - (w1 + w2));
        w0 = pf0->cbx; w1 = pfx->cbx;
        pi->cbx = half * (w1 + w0);
        pi->dcbxdx = half * (w1 - w0);
        w0 = pf0->cby; w1 = pfy->cby;
        pi->cby = half * (w1 + w0);
        pi->dcbydy = half * (w1 - w0);
        w0 = pf0->cbz; w1 = pfz->cbz;
        pi->cbz = half * (w1 + w0);
        pi->dcbzdz = half * (w1 - w0);
      }
}

void orc_unload_accumulator(vpic_field_t *f, const vpic_accumulator_t *a,
                            const vpic_grid_t *g) {
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const float cx = 0.25 * g->rdy * g->rdz / g->dt;
  const float cy = 0.25 * g->rdz * g->rdx / g->dt;
  const float cz = 0.25 * g->rdx * g->rdy / g->dt;
  for (int z = 1; z <= nz + 1; z++)
    for (int y = 1; y <= ny + 1; y++)
      for (int x = 1; x <= nx + 1; x++) {
        vpic_field_t *f0 = f + VOX(x, y, z, nx, ny);
        const vpic_accumulator_t *a0 = a + VOX(x, y, z, nx, ny);
        const vpic_accumulator_t *ax = a + VOX(x - 1, y, z, nx, ny);
        const vpic_accumulator_t *ay = a + VOX(x, y - 1, z, nx, ny);
        const vpic_accumulator_t *az = a + VOX(x, y, z - 1, nx, ny);
        const vpic_accumulator_t *ayz = a + VOX(x, y - 1, z - 1, nx, ny);
        const vpic_accumulator_t *azx = a + VOX(x - 1, y, z - 1, nx, ny);
        const vpic_accumulator_t *axy = a + VOX(x - 1, y - 1, z, nx, ny);
        f0->jfx += cx * (a0->jx[0] + ay->jx[1] + az->jx[2] + ayz->jx[3]);
        f0->jfy += cy * (a0->jy[0] + az->jy[1] + ax->jy[2] + azx->jy[3]);
        f0->jfz += cz * (a0->jz[0] + ax->jz[1] + ay->jz[2] + axy->jz[3]);
      }
}

void orc_clear_jf(vpic_field_t *f, const vpic_grid_t *g) {
  for (int v = 0; v < g->nv; v++) f[v].jfx = f[v].jfy = f[v].jfz = 0;
}

void orc_clear_rhof(vpic_field_t *f, const vpic_grid_t *g) {
  for (int v = 0; v < g->nv; v++) f[v].rhof = 0;
}

/* ------------------------------------------------------------------------ */
/* Field boundary helpers (standard/local.cc, standard/remote.cc)           */
/* ------------------------------------------------------------------------ */

/* Face k in -x,-y,-z,+x,+y,+z order -> axis and side. */
#define FACE_AXIS(k) ((k) % 3)
#define FACE_HIGH(k) ((k) >= 3)

/* Coordinates (x,y,z) addressed through an axis permutation: c[X], c[Y],
 * c[Z] with (X,Y,Z) the cyclic rotation starting at `ax`. */
static int fvox(const vpic_grid_t *g, const int c[3]) {
  return VOX(c[0], c[1], c[2], g->nx, g->ny);
}

static int nd(const vpic_grid_t *g, int axis) {
  return axis == 0 ? g->nx : axis == 1 ? g->ny : g->nz;
}

static float *fcomp(vpic_field_t *f, int which, int axis) {
  /* which: 0 = e, 1 = cb, 2 = tca, 3 = jf */
  float *base = which == 0 ? &f->ex : which == 1 ? &f->cbx
                : which == 2 ? &f->tcax : &f->jfx;
  return base + axis;
}

/* Iterate the "Y-oriented edges" of plane `pl` on the face normal to X:
 * X=plane, Y in 1..nY, Z in 1..nZ+1 (yz_EDGE_LOOP, local.cc:30). */
typedef void (*edge_fn)(vpic_field_t *f, const vpic_grid_t *g, int v,
                        int vsrc, void *ctx);

static void edge_loop(vpic_field_t *f, const vpic_grid_t *g, int X, int Y,
                      int pl, int yhi_extra, int zhi_extra, int src_off,
                      edge_fn fn, void *ctx) {
  int Z = 3 - X - Y;
  int c[3], s[3];
  for (int zz = 1; zz <= nd(g, Z) + zhi_extra; zz++)
    for (int yy = 1; yy <= nd(g, Y) + yhi_extra; yy++) {
      c[X] = pl;
      c[Y] = yy;
      c[Z] = zz;
      s[0] = c[0];
      s[1] = c[1];
      s[2] = c[2];
      s[X] += src_off;
      fn(f, g, fvox(g, c), fvox(g, s), ctx);
    }
}

typedef struct {
  int comp;
  float scale;
} scale_ctx_t;

static void fn_scale_jf(vpic_field_t *f, const vpic_grid_t *g, int v, int s,
                        void *ctx) {
  scale_ctx_t *c = (scale_ctx_t *)ctx;
  (void)g;
  (void)s;
  float *x = fcomp(f + v, 3, c->comp);
  if (c->scale == 0)
    *x = 0;
  else
    *x *= 2.;
}

static void fn_zero_e_tca(vpic_field_t *f, const vpic_grid_t *g, int v, int s,
                          void *ctx) {
  int comp = *(int *)ctx;
  (void)g;
  (void)s;
  *fcomp(f + v, 0, comp) = 0;
  *fcomp(f + v, 2, comp) = 0;
}

typedef struct {
  int comp;
  float sign;
} copy_ctx_t;

static void fn_ghost_cb(vpic_field_t *f, const vpic_grid_t *g, int v, int s,
                        void *ctx) {
  copy_ctx_t *c = (copy_ctx_t *)ctx;
  (void)g;
  float src = *fcomp(f + s, 1, c->comp);
  *fcomp(f + v, 1, c->comp) = c->sign > 0 ? src : -src;
}

/* For axis X with Y=(X+1)%3, Z=(X+2)%3 the reference's macros visit:
 *   tangential comp Y on "ZY" edges: Y in 1..nY+1? -- see below.
 * zy_EDGE_LOOP(x) = (x, 1..ny+1, 1..nz) carries cb_y / e_z / jf_z;
 * yz_EDGE_LOOP(x) = (x, 1..ny, 1..nz+1) carries cb_z / e_y / jf_y
 * (local.cc:30-36).  In rotated form for normal axis X: the loop over
 * (Y in 1..nY+1, Z in 1..nZ) holds cb_Y, e_Z, jf_Z and the loop over
 * (Y in 1..nY, Z in 1..nZ+1) holds cb_Z, e_Y, jf_Y. */

static void local_ghost_tang_b(vpic_field_t *f, const vpic_grid_t *g) {
  /* local.cc:116-121, order -x,-y,-z,+x,+y,+z. */
  for (int k = 0; k < 6; k++) {
    int bc = g->fbc[k];
    if (bc == VPIC_BC_PERIODIC) continue;
    int X = FACE_AXIS(k), Y = (X + 1) % 3, Z = (X + 2) % 3;
    int hi = FACE_HIGH(k);
    int ghost = hi ? nd(g, X) + 1 : 0;
    int off = hi ? -1 : 1; /* source = ghost - (i,j,k) */
    copy_ctx_t c;
    if (bc == VPIC_BC_ANTI_SYMMETRIC_FIELDS)
      c.sign = 1;
    else /* symmetric / pmc */
      c.sign = -1;
    c.comp = Y;
    edge_loop(f, g, X, Y, ghost, 1, 0, off, fn_ghost_cb, &c);
    c.comp = Z;
    edge_loop(f, g, X, Y, ghost, 0, 1, off, fn_ghost_cb, &c);
  }
}

/* Periodic ghost exchange with this same domain (remote.cc:61-134 with the
 * self-loop port of grid_comm.cc): every periodic face pair first sends
 * (face 1 towards -X, face nX towards +X), then receives in the order
 * -x,-y,-z,+x,+y,+z with rw = 1, lw = 0 (equal cell sizes). */
static void remote_ghost_tang_b(vpic_field_t *f, const vpic_grid_t *g) {
  float *buf[6] = {0};
  for (int k = 0; k < 6; k++) {
    if (g->fbc[k] != VPIC_BC_PERIODIC) continue;
    int X = FACE_AXIS(k), Y = (X + 1) % 3, Z = (X + 2) % 3;
    int face = FACE_HIGH(k) ? nd(g, X) : 1;
    int n = (nd(g, Y) + 1) * nd(g, Z) + nd(g, Y) * (nd(g, Z) + 1), m = 0;
    int c[3];
    buf[k] = (float *)malloc(sizeof(float) * (size_t)n);
    for (int zz = 1; zz <= nd(g, Z); zz++)
      for (int yy = 1; yy <= nd(g, Y) + 1; yy++) {
        c[X] = face; c[Y] = yy; c[Z] = zz;
        buf[k][m++] = *fcomp(f + fvox(g, c), 1, Y);
      }
    for (int zz = 1; zz <= nd(g, Z) + 1; zz++)
      for (int yy = 1; yy <= nd(g, Y); yy++) {
        c[X] = face; c[Y] = yy; c[Z] = zz;
        buf[k][m++] = *fcomp(f + fvox(g, c), 1, Z);
      }
  }
  for (int k = 0; k < 6; k++) {
    if (g->fbc[k] != VPIC_BC_PERIODIC) continue;
    int X = FACE_AXIS(k), Y = (X + 1) % 3, Z = (X + 2) % 3;
    /* Port k receives what the partner sent through the same-signed port,
     * i.e. this domain's face on the opposite side. */
    const float *p = buf[k];
    const float rw = 1.0f, lw = 0.0f;
    int lo = !FACE_HIGH(k);
    int face = lo ? nd(g, X) + 1 : 0;
    int inner = lo ? -1 : 1; /* field(x+i,...) */
    int c[3], s[3], m = 0;
    for (int zz = 1; zz <= nd(g, Z); zz++)
      for (int yy = 1; yy <= nd(g, Y) + 1; yy++) {
        c[X] = face; c[Y] = yy; c[Z] = zz;
        s[0] = c[0]; s[1] = c[1]; s[2] = c[2];
        s[X] += inner;
        *fcomp(f + fvox(g, c), 1, Y) =
            rw * p[m++] + lw * *fcomp(f + fvox(g, s), 1, Y);
      }
    for (int zz = 1; zz <= nd(g, Z) + 1; zz++)
      for (int yy = 1; yy <= nd(g, Y); yy++) {
        c[X] = face; c[Y] = yy; c[Z] = zz;
        s[0] = c[0]; s[1] = c[1]; s[2] = c[2];
        s[X] += inner;
        *fcomp(f + fvox(g, c), 1, Z) =
            rw * p[m++] + lw * *fcomp(f + fvox(g, s), 1, Z);
      }
  }
  for (int k = 0; k < 6; k++) free(buf[k]);
}

static void local_adjust_tang_e(vpic_field_t *f, const vpic_grid_t *g) {
  for (int k = 0; k < 6; k++) {
    if (g->fbc[k] != VPIC_BC_ANTI_SYMMETRIC_FIELDS) continue;
    int X = FACE_AXIS(k), Y = (X + 1) % 3, Z = (X + 2) % 3;
    int face = FACE_HIGH(k) ? nd(g, X) + 1 : 1;
    int comp = Y;
    edge_loop(f, g, X, Y, face, 0, 1, 0, fn_zero_e_tca, &comp);
    comp = Z;
    edge_loop(f, g, X, Y, face, 1, 0, 0, fn_zero_e_tca, &comp);
  }
}

static void local_adjust_norm_b(vpic_field_t *f, const vpic_grid_t *g) {
  for (int k = 0; k < 6; k++) {
    if (g->fbc[k] != VPIC_BC_SYMMETRIC_FIELDS) continue;
    int X = FACE_AXIS(k), Y = (X + 1) % 3, Z = (X + 2) % 3;
    int face = FACE_HIGH(k) ? nd(g, X) + 1 : 1;
    int c[3];
    for (int zz = 1; zz <= nd(g, Z); zz++)
      for (int yy = 1; yy <= nd(g, Y); yy++) {
        c[X] = face; c[Y] = yy; c[Z] = zz;
        *fcomp(f + fvox(g, c), 1, X) = 0;
      }
  }
}

static void local_adjust_jf(vpic_field_t *f, const vpic_grid_t *g) {
  for (int k = 0; k < 6; k++) {
    int bc = g->fbc[k];
    if (bc == VPIC_BC_PERIODIC) continue;
    int X = FACE_AXIS(k), Y = (X + 1) % 3, Z = (X + 2) % 3;
    int face = FACE_HIGH(k) ? nd(g, X) + 1 : 1;
    scale_ctx_t c;
    c.scale = bc == VPIC_BC_ANTI_SYMMETRIC_FIELDS ? 0.0f : 2.0f;
    c.comp = Y;
    edge_loop(f, g, X, Y, face, 0, 1, 0, fn_scale_jf, &c);
    c.comp = Z;
    edge_loop(f, g, X, Y, face, 1, 0, 0, fn_scale_jf, &c);
  }
}

void orc_synchronize_jf(vpic_field_t *f, const vpic_grid_t *g) {
  local_adjust_jf(f, g);
  /* remote.cc:474-502: x faces, then y, then z; both sends of an axis
   * are packed before either receive. */
  for (int X = 0; X < 3; X++) {
    if (g->fbc[X] != VPIC_BC_PERIODIC || g->fbc[X + 3] != VPIC_BC_PERIODIC)
      continue;
    int Y = (X + 1) % 3, Z = (X + 2) % 3;
    int n = nd(g, Y) * (nd(g, Z) + 1) + nd(g, Z) * (nd(g, Y) + 1);
    float *snd[2];
    int c[3];
    for (int side = 0; side < 2; side++) {
      int face = side == 0 ? 1 : nd(g, X) + 1, m = 0;
      snd[side] = (float *)malloc(sizeof(float) * (size_t)n);
      for (int zz = 1; zz <= nd(g, Z) + 1; zz++)
        for (int yy = 1; yy <= nd(g, Y); yy++) {
          c[X] = face; c[Y] = yy; c[Z] = zz;
          snd[side][m++] = *fcomp(f + fvox(g, c), 3, Y);
        }
      for (int zz = 1; zz <= nd(g, Z); zz++)
        for (int yy = 1; yy <= nd(g, Y) + 1; yy++) {
          c[X] = face; c[Y] = yy; c[Z] = zz;
          snd[side][m++] = *fcomp(f + fvox(g, c), 3, Z);
        }
    }
    /* END_RECV(-1): face nX+1 gets the face-1 send; END_RECV(+1): face 1
     * gets the face nX+1 send.  lw = rw = 1 for equal cell sizes. */
    for (int side = 0; side < 2; side++) {
      const float *p = side == 0 ? snd[0] : snd[1];
      int face = side == 0 ? nd(g, X) + 1 : 1, m = 0;
      const float lw = 1.0f, rw = 1.0f;
      for (int zz = 1; zz <= nd(g, Z) + 1; zz++)
        for (int yy = 1; yy <= nd(g, Y); yy++) {
          c[X] = face; c[Y] = yy; c[Z] = zz;
          float *x = fcomp(f + fvox(g, c), 3, Y);
          *x = lw * *x + rw * p[m++];
        }
      for (int zz = 1; zz <= nd(g, Z); zz++)
        for (int yy = 1; yy <= nd(g, Y) + 1; yy++) {
          c[X] = face; c[Y] = yy; c[Z] = zz;
          float *x = fcomp(f + fvox(g, c), 3, Z);
          *x = lw * *x + rw * p[m++];
        }
    }
    free(snd[0]);
    free(snd[1]);
  }
}

/* synchronize_rho (remote.cc:534-625): local_adjust_rhof / rhob
 * (local.cc:376-440) on local faces, then the x, y, z node-plane folds with
 * lw = rw = 1 for rhof and hlw = hrw = 1/2 for rhob (equal cell sizes). */
void orc_synchronize_rho(vpic_field_t *f, const vpic_grid_t *g) {
  for (int k = 0; k < 6; k++) {
    int bc = g->fbc[k];
    if (bc == VPIC_BC_PERIODIC) continue;
    int X = FACE_AXIS(k), Y = (X + 1) % 3, Z = (X + 2) % 3;
    int face = FACE_HIGH(k) ? nd(g, X) + 1 : 1;
    int c[3];
    for (int zz = 1; zz <= nd(g, Z) + 1; zz++)
      for (int yy = 1; yy <= nd(g, Y) + 1; yy++) {
        c[X] = face; c[Y] = yy; c[Z] = zz;
        vpic_field_t *fp = f + fvox(g, c);
        if (bc == VPIC_BC_ANTI_SYMMETRIC_FIELDS) fp->rhof = 0;
        else fp->rhof *= 2;
      }
  }
  for (int k = 0; k < 6; k++) {
    if (g->fbc[k] != VPIC_BC_ANTI_SYMMETRIC_FIELDS) continue;
    int X = FACE_AXIS(k), Y = (X + 1) % 3, Z = (X + 2) % 3;
    int face = FACE_HIGH(k) ? nd(g, X) + 1 : 1;
    int c[3];
    for (int zz = 1; zz <= nd(g, Z) + 1; zz++)
      for (int yy = 1; yy <= nd(g, Y) + 1; yy++) {
        c[X] = face; c[Y] = yy; c[Z] = zz;
        f[fvox(g, c)].rhob = 0;
      }
  }
  for (int X = 0; X < 3; X++) {
    if (g->fbc[X] != VPIC_BC_PERIODIC || g->fbc[X + 3] != VPIC_BC_PERIODIC)
      continue;
    int Y = (X + 1) % 3, Z = (X + 2) % 3;
    int n = (nd(g, Y) + 1) * (nd(g, Z) + 1);
    float *snd[2];
    int c[3];
    for (int side = 0; side < 2; side++) {
      int face = side == 0 ? 1 : nd(g, X) + 1, m = 0;
      snd[side] = (float *)malloc(sizeof(float) * 2 * (size_t)n);
      for (int zz = 1; zz <= nd(g, Z) + 1; zz++)
        for (int yy = 1; yy <= nd(g, Y) + 1; yy++) {
          c[X] = face; c[Y] = yy; c[Z] = zz;
          snd[side][m++] = f[fvox(g, c)].rhof;
          snd[side][m++] = f[fvox(g, c)].rhob;
        }
    }
    for (int side = 0; side < 2; side++) {
      const float *p = snd[side];
      int face = side == 0 ? nd(g, X) + 1 : 1, m = 0;
      const float hlw = 0.5f, hrw = 0.5f, lw = hlw + hlw, rw = hrw + hrw;
      for (int zz = 1; zz <= nd(g, Z) + 1; zz++)
        for (int yy = 1; yy <= nd(g, Y) + 1; yy++) {
          c[X] = face; c[Y] = yy; c[Z] = zz;
          vpic_field_t *fp = f + fvox(g, c);
          fp->rhof = lw * fp->rhof + rw * p[m++];
          fp->rhob = hlw * fp->rhob + hrw * p[m++];
        }
    }
    free(snd[0]);
    free(snd[1]);
  }
}

/* ------------------------------------------------------------------------ */
/* FDTD                                                                     */
/* ------------------------------------------------------------------------ */

void orc_advance_b(vpic_field_t *f, const vpic_grid_t *g, float frac) {
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const float px = (nx > 1) ? frac * g->cvac * g->dt * g->rdx : 0;
  const float py = (ny > 1) ? frac * g->cvac * g->dt * g->rdy : 0;
  const float pz = (nz > 1) ? frac * g->cvac * g->dt * g->rdz : 0;
  /* Every cb update reads only E, so the bulk/leftover split of
   * advance_b_pipeline.cc:69-122 has no effect on the values. */
  for (int z = 1; z <= nz + 1; z++)
    for (int y = 1; y <= ny + 1; y++)
      for (int x = 1; x <= nx + 1; x++) {
        vpic_field_t *f0 = f + VOX(x, y, z, nx, ny);
        const vpic_field_t *fx = f + VOX(x + 1, y, z, nx, ny);
        const vpic_field_t *fy = f + VOX(x, y + 1, z, nx, ny);
        const vpic_field_t *fz = f + VOX(x, y, z + 1, nx, ny);
        if (y <= ny && z <= nz)
          f0->cbx -= (py * (fy->ez - f0->ez) - pz * (fz->ey - f0->ey));
        if (x <= nx && z <= nz)
          f0->cby -= (pz * (fz->ex - f0->ex) - px * (fx->ez - f0->ez));
        if (x <= nx && y <= ny)
          f0->cbz -= (px * (fx->ey - f0->ey) - py * (fy->ex - f0->ex));
      }
  local_adjust_norm_b(f, g);
}

void orc_field_params_vacuum(vpic_field_params_t *mp, float damp) {
  /* sfa.cc:111-126 with sigma = 0, eps = mu = 1. */
  mp->decayx = mp->decayy = mp->decayz = (float)exp(-0.0);
  mp->drivex = mp->drivey = mp->drivez = (float)(1. / 1.0);
  mp->rmux = mp->rmuy = mp->rmuz = (float)(1. / 1.0);
  mp->epsx = mp->epsy = mp->epsz = 1.0f;
  mp->damp = damp;
  mp->nonconductive = 1.0f; /* sfa.cc:144 with ax = ay = az = 0 */
}

void orc_vacuum_advance_e(vpic_field_t *f, const vpic_grid_t *g,
                          const vpic_field_params_t *m) {
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const float decayx = m->decayx, drivex = m->drivex;
  const float decayy = m->decayy, drivey = m->drivey;
  const float decayz = m->decayz, drivez = m->drivez;
  const float damp = m->damp;
  const float px_muz = ((nx > 1) ? (1 + damp) * g->cvac * g->dt * g->rdx : 0) * m->rmuz;
  const float px_muy = ((nx > 1) ? (1 + damp) * g->cvac * g->dt * g->rdx : 0) * m->rmuy;
  const float py_mux = ((ny > 1) ? (1 + damp) * g->cvac * g->dt * g->rdy : 0) * m->rmux;
  const float py_muz = ((ny > 1) ? (1 + damp) * g->cvac * g->dt * g->rdy : 0) * m->rmuz;
  const float pz_muy = ((nz > 1) ? (1 + damp) * g->cvac * g->dt * g->rdz : 0) * m->rmuy;
  const float pz_mux = ((nz > 1) ? (1 + damp) * g->cvac * g->dt * g->rdz : 0) * m->rmux;
  const float cj = g->dt / g->eps0;

  remote_ghost_tang_b(f, g);
  local_ghost_tang_b(f, g);
  /* Each E update reads cb (not written here) and its own voxel, so the
   * pipeline/host split of vacuum_advance_e_pipeline.cc:87-329 does not
   * change any value. */
  for (int z = 1; z <= nz + 1; z++)
    for (int y = 1; y <= ny + 1; y++)
      for (int x = 1; x <= nx + 1; x++) {
        vpic_field_t *f0 = f + VOX(x, y, z, nx, ny);
        const vpic_field_t *fx = f + VOX(x - 1, y, z, nx, ny);
        const vpic_field_t *fy = f + VOX(x, y - 1, z, nx, ny);
        const vpic_field_t *fz = f + VOX(x, y, z - 1, nx, ny);
        if (x <= nx) {
          f0->tcax = (py_muz * (f0->cbz - fy->cbz) - pz_muy * (f0->cby - fz->cby)) -
                     damp * f0->tcax;
          f0->ex = decayx * f0->ex + drivex * (f0->tcax - cj * f0->jfx);
        }
        if (y <= ny) {
          f0->tcay = (pz_mux * (f0->cbx - fz->cbx) - px_muz * (f0->cbz - fx->cbz)) -
                     damp * f0->tcay;
          f0->ey = decayy * f0->ey + drivey * (f0->tcay - cj * f0->jfy);
        }
        if (z <= nz) {
          f0->tcaz = (px_muy * (f0->cby - fx->cby) - py_mux * (f0->cbx - fy->cbx)) -
                     damp * f0->tcaz;
          f0->ez = decayz * f0->ez + drivez * (f0->tcaz - cj * f0->jfz);
        }
      }
  local_adjust_tang_e(f, g);
}

/* ------------------------------------------------------------------------ */
/* sort_p (sort_p_pipeline.cc:143-214, 345-370)                            */
/* ------------------------------------------------------------------------ */

void orc_sort_p(vpic_particle_t *p, vpic_particle_t *aux, int np,
                int *partition, const vpic_grid_t *g) {
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const int vl = VOX(1, 1, 1, nx, ny), vh = VOX(nx, ny, nz, nx, ny);
  const int nv = g->nv;
  int *next = (int *)calloc((size_t)nv, sizeof(int));
  int sum = 0;
  for (int i = 0; i < np; i++) next[p[i].i]++;
  for (int v = vl; v <= vh; v++) {
    int count = next[v];
    next[v] = sum;
    partition[v] = sum;
    sum += count;
  }
  partition[vh + 1] = sum;
  for (int i = 0; i < np; i++) aux[next[p[i].i]++] = p[i];
  for (int v = 0; v < vl; v++) partition[v] = 0;
  for (int v = vh + 1; v <= nv; v++) partition[v] = np;
  memcpy(p, aux, sizeof(*p) * (size_t)np);
  free(next);
}

/* ------------------------------------------------------------------------ */
/* Diagnostics                                                              */
/* ------------------------------------------------------------------------ */

static double energy_p_range(const vpic_particle_t *p, int n0, int n1,
                             float qdt_2mc, float msp,
                             const vpic_interpolator_t *f) {
  const float one = 1.0;
  double en = 0.0;
  for (int n = n0; n < n1; n++) {
    float dx = p[n].dx, dy = p[n].dy, dz = p[n].dz, v0, v1, v2;
    int i = p[n].i;
    v0 = p[n].ux + qdt_2mc * ((f[i].ex + dy * f[i].dexdy) + dz * (f[i].dexdz + dy * f[i].d2exdydz));
    v1 = p[n].uy + qdt_2mc * ((f[i].ey + dz * f[i].deydz) + dx * (f[i].deydx + dz * f[i].d2eydzdx));
    v2 = p[n].uz + qdt_2mc * ((f[i].ez + dx * f[i].dezdx) + dy * (f[i].dezdy + dx * f[i].d2ezdxdy));
    v0 = v0 * v0 + v1 * v1 + v2 * v2;
    v0 = (msp * p[n].w) * (v0 / (one + sqrtf(one + v0)));
    en += (double)v0;
  }
  return en;
}

double orc_energy_p(const vpic_particle_t *p, int np, float q, float m,
                    const vpic_interpolator_t *f, const vpic_grid_t *g) {
  const float qdt_2mc = (q * g->dt) / (2 * m * g->cvac);
  int i0, n0, i1, n1;
  double en[2], local;
  distribute(np, 16, 0, 1, &i0, &n0);
  distribute(np, 16, 1, 1, &i1, &n1);
  en[0] = energy_p_range(p, i0, i0 + n0, qdt_2mc, m, f);
  en[1] = energy_p_range(p, i1, i1 + n1, qdt_2mc, m, f);
  local = 0.0;
  local += en[0];
  local += en[1];
  return local * ((double)g->cvac * (double)g->cvac);
}

/* center_p / uncenter_p, per particle (center_p_pipeline.cc:15-97,
 * uncenter_p_pipeline.cc:15-97).  Particles are independent, so the
 * DISTRIBUTE order does not matter. */
void orc_center_p(vpic_particle_t *p, int np, float q, float m,
                  const vpic_interpolator_t *f0, const vpic_grid_t *g,
                  int uncenter) {
  const float qdt_2mc_sp = (q * g->dt) / (2 * m * g->cvac);
  const float qdt_2mc = uncenter ? -qdt_2mc_sp : qdt_2mc_sp;
  const float qdt_4mc = uncenter ? -0.5 * qdt_2mc_sp : 0.5 * qdt_2mc_sp;
  const float one = 1.0, one_third = 1.0 / 3.0, two_fifteenths = 2.0 / 15.0;
  for (int n = 0; n < np; n++) {
    vpic_particle_t *pp = p + n;
    const float dx = pp->dx, dy = pp->dy, dz = pp->dz;
    const vpic_interpolator_t *f = f0 + pp->i;
    const float hax = qdt_2mc * ((f->ex + dy * f->dexdy) +
                                 dz * (f->dexdz + dy * f->d2exdydz));
    const float hay = qdt_2mc * ((f->ey + dz * f->deydz) +
                                 dx * (f->deydx + dz * f->d2eydzdx));
    const float haz = qdt_2mc * ((f->ez + dx * f->dezdx) +
                                 dy * (f->dezdy + dx * f->d2ezdxdy));
    const float cbx = f->cbx + dx * f->dcbxdx;
    const float cby = f->cby + dy * f->dcbydy;
    const float cbz = f->cbz + dz * f->dcbzdz;
    float ux = pp->ux, uy = pp->uy, uz = pp->uz;
    float v0, v1, v2, v3, v4;
    if (!uncenter) {
      ux += hax;
      uy += hay;
      uz += haz;
    }
    v0 = qdt_4mc / (float)sqrt(one + (ux * ux + (uy * uy + uz * uz)));
    v1 = cbx * cbx + (cby * cby + cbz * cbz);
    v2 = (v0 * v0) * v1;
    v3 = v0 * (one + v2 * (one_third + v2 * two_fifteenths));
    v4 = v3 / (one + v1 * (v3 * v3));
    v4 += v4;
    v0 = ux + v3 * (uy * cbz - uz * cby);
    v1 = uy + v3 * (uz * cbx - ux * cbz);
    v2 = uz + v3 * (ux * cby - uy * cbx);
    ux += v4 * (v1 * cbz - v2 * cby);
    uy += v4 * (v2 * cbx - v0 * cbz);
    uz += v4 * (v0 * cby - v1 * cbx);
    if (uncenter) {
      ux += hax;
      uy += hay;
      uz += haz;
    }
    pp->ux = ux;
    pp->uy = uy;
    pp->uz = uz;
  }
}

static void energy_f_range(double en[6], const vpic_field_t *f,
                           const vpic_grid_t *g,
                           const vpic_field_params_t *m, int l0, int cnt) {
  const int nx = g->nx, ny = g->ny;
  const float qepsx = 0.25 * m->epsx, qepsy = 0.25 * m->epsy,
              qepsz = 0.25 * m->epsz;
  const float hrmux = 0.50 * m->rmux, hrmuy = 0.50 * m->rmuy,
              hrmuz = 0.50 * m->rmuz;
  for (int k = 0; k < 6; k++) en[k] = 0;
  for (int l = l0; l < l0 + cnt; l++) {
    int x = 1 + l % nx, y = 1 + (l / nx) % ny, z = 1 + l / (nx * ny);
    const vpic_field_t *f0 = f + VOX(x, y, z, nx, ny);
    const vpic_field_t *fx = f + VOX(x + 1, y, z, nx, ny);
    const vpic_field_t *fy = f + VOX(x, y + 1, z, nx, ny);
    const vpic_field_t *fz = f + VOX(x, y, z + 1, nx, ny);
    const vpic_field_t *fyz = f + VOX(x, y + 1, z + 1, nx, ny);
    const vpic_field_t *fzx = f + VOX(x + 1, y, z + 1, nx, ny);
    const vpic_field_t *fxy = f + VOX(x + 1, y + 1, z, nx, ny);
    en[0] += qepsx * (f0->ex * f0->ex + fy->ex * fy->ex + fz->ex * fz->ex + fyz->ex * fyz->ex);
    en[1] += qepsy * (f0->ey * f0->ey + fz->ey * fz->ey + fx->ey * fx->ey + fzx->ey * fzx->ey);
    en[2] += qepsz * (f0->ez * f0->ez + fx->ez * fx->ez + fy->ez * fy->ez + fxy->ez * fxy->ez);
    en[3] += hrmux * (f0->cbx * f0->cbx + fx->cbx * fx->cbx);
    en[4] += hrmuy * (f0->cby * f0->cby + fy->cby * fy->cby);
    en[5] += hrmuz * (f0->cbz * f0->cbz + fz->cbz * fz->cbz);
  }
}

void orc_energy_f(double en[6], const vpic_field_t *f, const vpic_grid_t *g,
                  const vpic_field_params_t *m) {
  int ncell = g->nx * g->ny * g->nz;
  int blk = 16 * (ncell / 16);
  double e0[6], e1[6];
  energy_f_range(e0, f, g, m, 0, blk);
  energy_f_range(e1, f, g, m, blk, ncell - blk);
  double v0 = 0.5 * g->eps0 * g->dV;
  for (int k = 0; k < 6; k++) en[k] = (e0[k] + e1[k]) * v0;
}

void orc_accumulate_rho_p(vpic_field_t *f, const vpic_particle_t *p, int np,
                          float q, const vpic_grid_t *g) {
  const float q_8V = q * g->r8V;
  const int sy = g->sy, sz = g->sz;
  for (int n = 0; n < np; n++) {
    float w0 = p[n].dx, w1 = p[n].dy, dz = p[n].dz;
    int v = p[n].i;
    float w7 = p[n].w * q_8V, w2, w3, w4, w5, w6;
    /* rho_p.cc:64-74: FMA(x,y,z) = z + x*y, FNMS(x,y,z) = z - x*y. */
    w6 = w7 - w0 * w7;
    w7 = w7 + w0 * w7;
    w4 = w6 - w1 * w6;
    w5 = w7 - w1 * w7;
    w6 = w6 + w1 * w6;
    w7 = w7 + w1 * w7;
    w0 = w4 - dz * w4;
    w1 = w5 - dz * w5;
    w2 = w6 - dz * w6;
    w3 = w7 - dz * w7;
    w4 = w4 + dz * w4;
    w5 = w5 + dz * w5;
    w6 = w6 + dz * w6;
    w7 = w7 + dz * w7;
    f[v].rhof += w0;
    f[v + 1].rhof += w1;
    f[v + sy].rhof += w2;
    f[v + sy + 1].rhof += w3;
    f[v + sz].rhof += w4;
    f[v + sz + 1].rhof += w5;
    f[v + sz + sy].rhof += w6;
    f[v + sz + sy + 1].rhof += w7;
  }
}

/* accumulate_rhob (rho_p.cc:126-173): the particle's charge on the 8 nodes
 * of its voxel, FMA/FNMS written out as z + x*y and z - x*y as the file's
 * macros define them; a node weight is doubled on each face of the local
 * domain the voxel touches (rho_p.cc:153-165). */
void orc_accumulate_rhob(vpic_field_t *f, const vpic_particle_t *p,
                         float qsp, const vpic_grid_t *g) {
  float w0 = p->dx, w1 = p->dy, w2, w3, w4, w5, w6, w7, dz = p->dz;
  const int v = p->i, sy = g->sy, sz = g->sz;
  int x, y, z;
  w7 = (qsp * g->r8V) * p->w;
  w6 = w7 - w0 * w7;
  w7 = w7 + w0 * w7;
  w4 = w6 - w1 * w6;
  w5 = w7 - w1 * w7;
  w6 = w6 + w1 * w6;
  w7 = w7 + w1 * w7;
  w0 = w4 - dz * w4;
  w1 = w5 - dz * w5;
  w2 = w6 - dz * w6;
  w3 = w7 - dz * w7;
  w4 = w4 + dz * w4;
  w5 = w5 + dz * w5;
  w6 = w6 + dz * w6;
  w7 = w7 + dz * w7;
  x = v;
  z = x / sz;
  if (z == 1) w0 += w0, w1 += w1, w2 += w2, w3 += w3;
  if (z == g->nz) w4 += w4, w5 += w5, w6 += w6, w7 += w7;
  x -= sz * z;
  y = x / sy;
  if (y == 1) w0 += w0, w1 += w1, w4 += w4, w5 += w5;
  if (y == g->ny) w2 += w2, w3 += w3, w6 += w6, w7 += w7;
  x -= sy * y;
  if (x == 1) w0 += w0, w2 += w2, w4 += w4, w6 += w6;
  if (x == g->nx) w1 += w1, w3 += w3, w5 += w5, w7 += w7;
  f[v].rhob += w0;
  f[v + 1].rhob += w1;
  f[v + sy].rhob += w2;
  f[v + sy + 1].rhob += w3;
  f[v + sz].rhob += w4;
  f[v + sz + 1].rhob += w5;
  f[v + sz + sy].rhob += w6;
  f[v + sz + sy + 1].rhob += w7;
}

/* boundary_p's mover loop for one species of a domain with no shared face
 * (boundary_p.cc:254-375): the movers last first; the particle's voxel and
 * face come from its index 8*voxel + face (:257-266); an absorbing face
 * puts its charge on rhob (:268-275), a custom boundary (code <= -3,
 * :332-346) and anything else (:348-352) are recorded in hits[m]; then the
 * last particle moves into the mover's slot (:354-371).  counts[0] = custom
 * hits, counts[1] = dropped.  Returns the new particle count. */
int orc_boundary_p(vpic_particle_t *p0, int np,
                   const vpic_particle_mover_t *pm, int nm, vpic_field_t *f,
                   float q, const vpic_grid_t *g, vpic_boundary_hit_t *hits,
                   int32_t *counts) {
  counts[0] = counts[1] = 0;
  for (int m = nm - 1; m >= 0; m--) {
    const int i = pm[m].i;
    int voxel = p0[i].i;
    const int face = voxel & 7;
    voxel >>= 3;
    p0[i].i = voxel;
    const int64_t nn = g->neighbor[6 * (int64_t)voxel + face];
    int32_t code = VPIC_BOUNDARY_UNKNOWN;
    if (nn == VPIC_BC_ABSORB_PARTICLES) {
      orc_accumulate_rhob(f, p0 + i, q, g);
      code = VPIC_BC_ABSORB_PARTICLES;
    } else if (nn <= -3) {
      code = (int32_t)nn;
      counts[0]++;
    } else {
      counts[1]++;
    }
    if (hits) {
      memset(hits + m, 0, sizeof(*hits));
      hits[m].p = p0[i];
      hits[m].m = pm[m];
      hits[m].face = face;
      hits[m].code = code;
    }
    np--;
    p0[i] = p0[np];
  }
  return np;
}

/* The injection loop of boundary_p (boundary_p.cc:595-711) for one species:
 * the injectors last first, each appended to p with a mover for its
 * remaining displacement and moved with move_p; a full array drops it
 * (the fixed-size branch, :652-698).  counts[0] / counts[1] = particles /
 * movers dropped. */
void orc_inject_p(vpic_particle_t *p, int *np, int max_np,
                  vpic_particle_mover_t *pm, int *nm, int max_nm,
                  const vpic_particle_injector_t *pi, int n,
                  vpic_accumulator_t *a, float q, const vpic_grid_t *g,
                  int32_t *counts) {
  counts[0] = counts[1] = 0;
  for (int k = n - 1; k >= 0; k--) {
    const vpic_particle_injector_t *c = pi + k;
    if (*np >= max_np) {
      counts[0]++;
      continue;
    }
    vpic_particle_t *d = p + *np;
    d->dx = c->dx;
    d->dy = c->dy;
    d->dz = c->dz;
    d->i = c->i;
    d->ux = c->ux;
    d->uy = c->uy;
    d->uz = c->uz;
    d->w = c->w;
    const int at = (*np)++;
    if (*nm >= max_nm) {
      counts[1]++;
      continue;
    }
    vpic_particle_mover_t *m = pm + *nm;
    m->dispx = c->dispx;
    m->dispy = c->dispy;
    m->dispz = c->dispz;
    m->i = at;
    *nm += orc_move_p(p, m, a, g, q);
  }
}

/* ------------------------------------------------------------------------ */
/* Divergence cleaning and initialization kernels of the vacuum field table */
/* (sfa.cc:203-210)                                                         */
/* ------------------------------------------------------------------------ */

static float *fslot(vpic_field_t *f, int k) { return (float *)f + k; }

/* Normal-E ghosts (and tca for local faces) of every face: periodic faces
 * take the opposite interior node plane with rw = 1, lw = 0
 * (remote.cc:137-207); local faces copy (anti-symmetric) or negate
 * (symmetric, pmc) the adjacent plane (local.cc:129-179).  Ghost node
 * planes only read interior planes, so the order does not matter. */
static void ghost_norm_e(vpic_field_t *f, const vpic_grid_t *g) {
  for (int k = 0; k < 6; k++) {
    const int X = FACE_AXIS(k), Y = (X + 1) % 3, Z = (X + 2) % 3;
    const int hi = FACE_HIGH(k), bc = g->fbc[k];
    const int ghost = hi ? nd(g, X) + 1 : 0;
    int c[3], s[3];
    for (int zz = 1; zz <= nd(g, Z) + 1; zz++)
      for (int yy = 1; yy <= nd(g, Y) + 1; yy++) {
        c[X] = ghost; c[Y] = yy; c[Z] = zz;
        s[0] = c[0]; s[1] = c[1]; s[2] = c[2];
        vpic_field_t *f0 = f + fvox(g, c);
        if (bc == VPIC_BC_PERIODIC) {
          /* the opposite interior plane; lw multiplies the adjacent one */
          s[X] = hi ? 1 : nd(g, X);
          int a[3] = {c[0], c[1], c[2]};
          a[X] = hi ? nd(g, X) : 1;
          const float rw = 1.0f, lw = 0.0f;
          *fcomp(f0, 0, X) = rw * *fcomp(f + fvox(g, s), 0, X) +
                             lw * *fcomp(f + fvox(g, a), 0, X);
        } else {
          s[X] = hi ? nd(g, X) : 1;
          vpic_field_t *f1 = f + fvox(g, s);
          const float e = *fcomp(f1, 0, X), t = *fcomp(f1, 2, X);
          if (bc == VPIC_BC_ANTI_SYMMETRIC_FIELDS) {
            *fcomp(f0, 0, X) = e;
            *fcomp(f0, 2, X) = t;
          } else if (bc == VPIC_BC_ABSORB_FIELDS) {
            s[X] = hi ? nd(g, X) - 1 : 2;
            vpic_field_t *f2 = f + fvox(g, s);
            *fcomp(f0, 0, X) = 2 * e - *fcomp(f2, 0, X);
            *fcomp(f0, 2, X) = 2 * t - *fcomp(f2, 2, X);
          } else {
            *fcomp(f0, 0, X) = -e;
            *fcomp(f0, 2, X) = -t;
          }
        }
      }
  }
}

/* Zero field slot `slot` (floats into field_t) on the nodes of every
 * anti-symmetric face, and of every absorbing face when `absorb`
 * (local_adjust_div_e, local.cc:299-334; local_adjust_rhob,
 * local.cc:415-440). */
static void zero_node_faces(vpic_field_t *f, const vpic_grid_t *g, int slot,
                            int absorb) {
  for (int k = 0; k < 6; k++) {
    if (g->fbc[k] != VPIC_BC_ANTI_SYMMETRIC_FIELDS &&
        !(absorb && g->fbc[k] == VPIC_BC_ABSORB_FIELDS))
      continue;
    const int X = FACE_AXIS(k), Y = (X + 1) % 3, Z = (X + 2) % 3;
    const int face = FACE_HIGH(k) ? nd(g, X) + 1 : 1;
    int c[3];
    for (int zz = 1; zz <= nd(g, Z) + 1; zz++)
      for (int yy = 1; yy <= nd(g, Y) + 1; yy++) {
        c[X] = face; c[Y] = yy; c[Z] = zz;
        *fslot(f + fvox(g, c), slot) = 0;
      }
  }
}

void orc_compute_div_e_err(vpic_field_t *f, const vpic_grid_t *g,
                           const vpic_field_params_t *m) {
  /* vacuum_compute_div_e_err_pipeline.h:17-50, .cc:61-174 */
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const float nc = m->nonconductive;
  const float px = ((nx > 1) ? g->rdx : 0) * m->epsx;
  const float py = ((ny > 1) ? g->rdy : 0) * m->epsy;
  const float pz = ((nz > 1) ? g->rdz : 0) * m->epsz;
  const float cj = 1. / g->eps0;
  ghost_norm_e(f, g);
  for (int z = 1; z <= nz + 1; z++)
    for (int y = 1; y <= ny + 1; y++)
      for (int x = 1; x <= nx + 1; x++) {
        vpic_field_t *f0 = f + VOX(x, y, z, nx, ny);
        const vpic_field_t *fx = f + VOX(x - 1, y, z, nx, ny);
        const vpic_field_t *fy = f + VOX(x, y - 1, z, nx, ny);
        const vpic_field_t *fz = f + VOX(x, y, z - 1, nx, ny);
        f0->div_e_err = nc * (px * (f0->ex - fx->ex) + py * (f0->ey - fy->ey) +
                              pz * (f0->ez - fz->ez) - cj * (f0->rhof + f0->rhob));
      }
  zero_node_faces(f, g, 3, 1);
}

void orc_compute_rhob(vpic_field_t *f, const vpic_grid_t *g,
                      const vpic_field_params_t *m) {
  /* vacuum_compute_rhob_pipeline.h:17-52, .cc:41-175 */
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const float nc = m->nonconductive;
  const float px = (nx > 1) ? g->eps0 * m->epsx * g->rdx : 0;
  const float py = (ny > 1) ? g->eps0 * m->epsy * g->rdy : 0;
  const float pz = (nz > 1) ? g->eps0 * m->epsz * g->rdz : 0;
  ghost_norm_e(f, g);
  for (int z = 1; z <= nz + 1; z++)
    for (int y = 1; y <= ny + 1; y++)
      for (int x = 1; x <= nx + 1; x++) {
        vpic_field_t *f0 = f + VOX(x, y, z, nx, ny);
        const vpic_field_t *fx = f + VOX(x - 1, y, z, nx, ny);
        const vpic_field_t *fy = f + VOX(x, y - 1, z, nx, ny);
        const vpic_field_t *fz = f + VOX(x, y, z - 1, nx, ny);
        f0->rhob = nc * (px * (f0->ex - fx->ex) + py * (f0->ey - fy->ey) +
                         pz * (f0->ez - fz->ez) - f0->rhof);
      }
  zero_node_faces(f, g, 11, 0);
}

/* Weighted sum of div_e_err^2 over the nodes 1..n+1: interior nodes in
 * FP32 products, boundary nodes in FP64 with 1/2, 1/4, 1/8 on faces,
 * edges and corners (compute_rms_div_e_err_pipeline.cc:16-180). */
double orc_compute_rms_div_e_err(const vpic_field_t *f, const vpic_grid_t *g) {
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  double err = 0;
  for (int z = 2; z <= nz; z++)
    for (int y = 2; y <= ny; y++)
      for (int x = 2; x <= nx; x++) {
        const float d = f[VOX(x, y, z, nx, ny)].div_e_err;
        err += d * d;
      }
  for (int z = 1; z <= nz + 1; z++)
    for (int y = 1; y <= ny + 1; y++)
      for (int x = 1; x <= nx + 1; x++) {
        const int bx = x == 1 || x == nx + 1, by = y == 1 || y == ny + 1,
                  bz = z == 1 || z == nz + 1;
        const int nb = bx + by + bz;
        if (!nb) continue;
        const double w = nb == 1 ? 0.5 : nb == 2 ? 0.25 : 0.125;
        const double d = f[VOX(x, y, z, nx, ny)].div_e_err;
        err += w * d * d;
      }
  double local0 = err * g->dV;
  double local1 = (g->nx * g->ny * g->nz) * g->dV;
  return g->eps0 * sqrt(local0 / local1);
}

void orc_clean_div_e(vpic_field_t *f, const vpic_grid_t *g,
                     const vpic_field_params_t *m) {
  /* vacuum_clean_div_e_pipeline.h:16-55, .cc:43-145 */
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const float _rdx = (nx > 1) ? g->rdx : 0;
  const float _rdy = (ny > 1) ? g->rdy : 0;
  const float _rdz = (nz > 1) ? g->rdz : 0;
  const float alphadt = 0.3888889 / (_rdx * _rdx + _rdy * _rdy + _rdz * _rdz);
  const float px = (alphadt * _rdx) * m->drivex;
  const float py = (alphadt * _rdy) * m->drivey;
  const float pz = (alphadt * _rdz) * m->drivez;
  for (int z = 1; z <= nz + 1; z++)
    for (int y = 1; y <= ny + 1; y++)
      for (int x = 1; x <= nx + 1; x++) {
        vpic_field_t *f0 = f + VOX(x, y, z, nx, ny);
        if (x <= nx)
          f0->ex += px * (f[VOX(x + 1, y, z, nx, ny)].div_e_err - f0->div_e_err);
        if (y <= ny)
          f0->ey += py * (f[VOX(x, y + 1, z, nx, ny)].div_e_err - f0->div_e_err);
        if (z <= nz)
          f0->ez += pz * (f[VOX(x, y, z + 1, nx, ny)].div_e_err - f0->div_e_err);
      }
  local_adjust_tang_e(f, g);
}

void orc_compute_div_b_err(vpic_field_t *f, const vpic_grid_t *g) {
  /* compute_div_b_err_pipeline.cc:15-57 */
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const float px = (nx > 1) ? g->rdx : 0;
  const float py = (ny > 1) ? g->rdy : 0;
  const float pz = (nz > 1) ? g->rdz : 0;
  for (int z = 1; z <= nz; z++)
    for (int y = 1; y <= ny; y++)
      for (int x = 1; x <= nx; x++) {
        vpic_field_t *f0 = f + VOX(x, y, z, nx, ny);
        f0->div_b_err = px * (f[VOX(x + 1, y, z, nx, ny)].cbx - f0->cbx) +
                        py * (f[VOX(x, y + 1, z, nx, ny)].cby - f0->cby) +
                        pz * (f[VOX(x, y, z + 1, nx, ny)].cbz - f0->cbz);
      }
}

double orc_compute_rms_div_b_err(const vpic_field_t *f, const vpic_grid_t *g) {
  /* compute_rms_div_b_err_pipeline.cc:14-111 */
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  double err = 0;
  for (int z = 1; z <= nz; z++)
    for (int y = 1; y <= ny; y++)
      for (int x = 1; x <= nx; x++) {
        const float d = f[VOX(x, y, z, nx, ny)].div_b_err;
        err += d * d;
      }
  double local0 = err * g->dV;
  double local1 = (g->nx * g->ny * g->nz) * g->dV;
  return g->eps0 * sqrt(local0 / local1);
}

/* div_b_err ghosts (cell centred): periodic faces take the opposite
 * interior cell plane (remote.cc:209-281); local faces copy (anti-
 * symmetric), negate (symmetric, pmc) or zero (absorbing) the adjacent
 * plane (local.cc:182-221). */
static void ghost_div_b(vpic_field_t *f, const vpic_grid_t *g) {
  for (int k = 0; k < 6; k++) {
    const int X = FACE_AXIS(k), Y = (X + 1) % 3, Z = (X + 2) % 3;
    const int hi = FACE_HIGH(k), bc = g->fbc[k];
    const int ghost = hi ? nd(g, X) + 1 : 0;
    int c[3], s[3], a[3];
    for (int zz = 1; zz <= nd(g, Z); zz++)
      for (int yy = 1; yy <= nd(g, Y); yy++) {
        c[X] = ghost; c[Y] = yy; c[Z] = zz;
        s[0] = a[0] = c[0]; s[1] = a[1] = c[1]; s[2] = a[2] = c[2];
        vpic_field_t *f0 = f + fvox(g, c);
        if (bc == VPIC_BC_PERIODIC) {
          s[X] = hi ? 1 : nd(g, X);
          a[X] = hi ? nd(g, X) : 1;
          const float rw = 1.0f, lw = 0.0f;
          f0->div_b_err = rw * f[fvox(g, s)].div_b_err + lw * f[fvox(g, a)].div_b_err;
        } else {
          s[X] = hi ? nd(g, X) : 1;
          const float d = f[fvox(g, s)].div_b_err;
          f0->div_b_err = bc == VPIC_BC_ANTI_SYMMETRIC_FIELDS ? d
                          : bc == VPIC_BC_ABSORB_FIELDS      ? 0.0f
                                                             : -d;
        }
      }
  }
}

void orc_clean_div_b(vpic_field_t *f, const vpic_grid_t *g) {
  /* clean_div_b_pipeline.h:17-27, clean_div_b_pipeline.cc:230-328 */
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  float px = (nx > 1) ? g->rdx : 0;
  float py = (ny > 1) ? g->rdy : 0;
  float pz = (nz > 1) ? g->rdz : 0;
  const float alphadt = 0.3888889 / (px * px + py * py + pz * pz);
  px *= alphadt;
  py *= alphadt;
  pz *= alphadt;
  ghost_div_b(f, g);
  for (int z = 1; z <= nz + 1; z++)
    for (int y = 1; y <= ny + 1; y++)
      for (int x = 1; x <= nx + 1; x++) {
        vpic_field_t *f0 = f + VOX(x, y, z, nx, ny);
        if (y <= ny && z <= nz)
          f0->cbx += px * (f0->div_b_err - f[VOX(x - 1, y, z, nx, ny)].div_b_err);
        if (x <= nx && z <= nz)
          f0->cby += py * (f0->div_b_err - f[VOX(x, y - 1, z, nx, ny)].div_b_err);
        if (x <= nx && y <= ny)
          f0->cbz += pz * (f0->div_b_err - f[VOX(x, y, z - 1, nx, ny)].div_b_err);
      }
  local_adjust_norm_b(f, g);
}

void orc_compute_curl_b(vpic_field_t *f, const vpic_grid_t *g,
                        const vpic_field_params_t *m) {
  /* vacuum_compute_curl_b_pipeline.h:17-66, .cc:50-326 */
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const float px_muz = ((nx > 1) ? g->cvac * g->dt * g->rdx : 0) * m->rmuz;
  const float px_muy = ((nx > 1) ? g->cvac * g->dt * g->rdx : 0) * m->rmuy;
  const float py_mux = ((ny > 1) ? g->cvac * g->dt * g->rdy : 0) * m->rmux;
  const float py_muz = ((ny > 1) ? g->cvac * g->dt * g->rdy : 0) * m->rmuz;
  const float pz_muy = ((nz > 1) ? g->cvac * g->dt * g->rdz : 0) * m->rmuy;
  const float pz_mux = ((nz > 1) ? g->cvac * g->dt * g->rdz : 0) * m->rmux;
  remote_ghost_tang_b(f, g);
  local_ghost_tang_b(f, g);
  for (int z = 1; z <= nz + 1; z++)
    for (int y = 1; y <= ny + 1; y++)
      for (int x = 1; x <= nx + 1; x++) {
        vpic_field_t *f0 = f + VOX(x, y, z, nx, ny);
        const vpic_field_t *fx = f + VOX(x - 1, y, z, nx, ny);
        const vpic_field_t *fy = f + VOX(x, y - 1, z, nx, ny);
        const vpic_field_t *fz = f + VOX(x, y, z - 1, nx, ny);
        if (x <= nx)
          f0->tcax = (py_muz * (f0->cbz - fy->cbz) - pz_muy * (f0->cby - fz->cby));
        if (y <= ny)
          f0->tcay = (pz_mux * (f0->cbx - fz->cbx) - px_muz * (f0->cbz - fx->cbz));
        if (z <= nz)
          f0->tcaz = (px_muy * (f0->cby - fx->cby) - py_mux * (f0->cbx - fy->cbx));
      }
  local_adjust_tang_e(f, g);
}

/* synchronize_tang_e_norm_b for one domain (remote.cc:298-416): local
 * adjusts, then for the x, y and z periodic face pairs in turn both faces
 * take 0.5*(w1+w2) in FP64 of cb normal, e and tca tangential; err sums
 * (w1-w2)^2 of cb and e, once from each face. */
double orc_synchronize_tang_e_norm_b(vpic_field_t *f, const vpic_grid_t *g) {
  double err = 0;
  local_adjust_tang_e(f, g);
  local_adjust_norm_b(f, g);
  for (int X = 0; X < 3; X++) {
    if (g->fbc[X] != VPIC_BC_PERIODIC || g->fbc[X + 3] != VPIC_BC_PERIODIC)
      continue;
    const int Y = (X + 1) % 3, Z = (X + 2) % 3;
    const int lo = 1, hi = nd(g, X) + 1;
    int a[3], b[3];
    /* (slot block, component, Y extra, Z extra, counts in err) */
    const int parts[5][5] = {{1, X, 0, 0, 1}, {0, Y, 0, 1, 1}, {2, Y, 0, 1, 0},
                             {0, Z, 1, 0, 1}, {2, Z, 1, 0, 0}};
    for (int q = 0; q < 5; q++) {
      const int blk = parts[q][0], comp = parts[q][1];
      for (int zz = 1; zz <= nd(g, Z) + parts[q][3]; zz++)
        for (int yy = 1; yy <= nd(g, Y) + parts[q][2]; yy++) {
          a[X] = lo; a[Y] = yy; a[Z] = zz;
          b[X] = hi; b[Y] = yy; b[Z] = zz;
          float *pa = fcomp(f + fvox(g, a), blk, comp);
          float *pb = fcomp(f + fvox(g, b), blk, comp);
          const double w1 = *pa, w2 = *pb;
          *pa = 0.5 * (w1 + w2);
          *pb = 0.5 * (w2 + w1);
          if (parts[q][4]) err += 2 * ((w1 - w2) * (w1 - w2));
        }
    }
  }
  return err;
}

/* ------------------------------------------------------------------------ */
/* Hydro moments (hydro_p_pipeline.cc, hydro_array.cc)                      */
/* ------------------------------------------------------------------------ */

/* accumulate_hydro_p_pipeline_scalar for particles [n0, n1) into h
 * (hydro_p_pipeline.cc:17-210). */
static void hydro_range(vpic_hydro_t *h, const vpic_particle_t *p, int n0,
                        int n1, float qsp, float msp,
                        const vpic_interpolator_t *f, const vpic_grid_t *g) {
  const float qdt_2mc = (qsp * g->dt) / (2 * msp * g->cvac);
  const float qdt_4mc2 = qdt_2mc / (2 * g->cvac);
  const float mspc = g->cvac * msp;
  const float c = g->cvac;
  const float r8V = g->r8V;
  const float one = 1.0, one_third = 1.0 / 3.0;
  const int stride_10 = 1, stride_21 = g->sy - 1, stride_43 = g->sz - g->sy - 1;
  for (int n = n0; n < n1; n++) {
    float dx = p[n].dx, dy = p[n].dy, dz = p[n].dz;
    int i = p[n].i;
    float ux = p[n].ux, uy = p[n].uy, uz = p[n].uz, w = p[n].w;
    const vpic_interpolator_t *fi = f + i;
    float vx, vy, vz, ke_mc, t, w0, w1, w2, w3, w4, w5, w6, w7;
    ux += qdt_2mc * ((fi->ex + dy * fi->dexdy) + dz * (fi->dexdz + dy * fi->d2exdydz));
    uy += qdt_2mc * ((fi->ey + dz * fi->deydz) + dx * (fi->deydx + dz * fi->d2eydzdx));
    uz += qdt_2mc * ((fi->ez + dx * fi->dezdx) + dy * (fi->dezdy + dx * fi->d2ezdxdy));
    w5 = fi->cbx + dx * fi->dcbxdx;
    w6 = fi->cby + dy * fi->dcbydy;
    w7 = fi->cbz + dz * fi->dcbzdz;
    ke_mc = ux * ux + uy * uy + uz * uz;
    vz = sqrtf(one + ke_mc);
    ke_mc *= c / (vz + one);
    vz = c / vz;
    w0 = qdt_4mc2 * vz;
    w1 = w5 * w5 + w6 * w6 + w7 * w7;
    w2 = w0 * w0 * w1;
    w3 = w0 * (one + (one_third) * w2 * (one + 0.4f * w2));
    w4 = w3 / (one + w1 * w3 * w3);
    w4 += w4;
    w0 = ux + w3 * (uy * w7 - uz * w6);
    w1 = uy + w3 * (uz * w5 - ux * w7);
    w2 = uz + w3 * (ux * w6 - uy * w5);
    ux += w4 * (w1 * w7 - w2 * w6);
    uy += w4 * (w2 * w5 - w0 * w7);
    uz += w4 * (w0 * w6 - w1 * w5);
    vx = ux * vz;
    vy = uy * vz;
    vz = uz * vz;
    w0 = r8V * w;
    dx *= w0;
    w1 = w0 + dx;
    w0 -= dx;
    w3 = one + dy;
    w2 = w0 * w3;
    w3 *= w1;
    dy = one - dy;
    w0 *= dy;
    w1 *= dy;
    w7 = one + dz;
    w4 = w0 * w7;
    w5 = w1 * w7;
    w6 = w2 * w7;
    w7 *= w3;
    dz = one - dz;
    w0 *= dz;
    w1 *= dz;
    w2 *= dz;
    w3 *= dz;
    const float wn[8] = {w0, w1, w2, w3, w4, w5, w6, w7};
    const int step[8] = {0, stride_10, stride_21, stride_10,
                         stride_43, stride_10, stride_21, stride_10};
    for (int k = 0; k < 8; k++) {
      i += step[k];
      vpic_hydro_t *hh = h + i;
      t = qsp * wn[k];
      hh->jx += t * vx;
      hh->jy += t * vy;
      hh->jz += t * vz;
      hh->rho += t;
      t = mspc * wn[k];
      dx = t * ux;
      dy = t * uy;
      dz = t * uz;
      hh->px += dx;
      hh->py += dy;
      hh->pz += dz;
      hh->ke += t * ke_mc;
      hh->txx += dx * vx;
      hh->tyy += dy * vy;
      hh->tzz += dz * vz;
      hh->tyz += dy * vz;
      hh->tzx += dz * vx;
      hh->txy += dx * vy;
    }
  }
}

void orc_accumulate_hydro_p(vpic_hydro_t *h, const vpic_particle_t *p, int np,
                            float q, float m, const vpic_interpolator_t *f,
                            const vpic_grid_t *g) {
  /* one pipeline: rank 0 accumulates into h (copy 0), the straggler rank
   * into a cleared copy 1, then reduce_hydro_array sets h = h + copy 1
   * (hydro_p_pipeline.cc:214-242; reduce_array_pipeline.cc, O1) */
  int i0, n0, i1, n1;
  distribute(np, 16, 0, 1, &i0, &n0);
  distribute(np, 16, 1, 1, &i1, &n1);
  vpic_hydro_t *c1 = (vpic_hydro_t *)calloc((size_t)g->nv, sizeof(*h));
  hydro_range(h, p, i0, i0 + n0, q, m, f, g);
  hydro_range(c1, p, i1, i1 + n1, q, m, f, g);
  for (int v = 0; v < g->nv; v++) {
    float *a = (float *)(h + v);
    const float *b = (const float *)(c1 + v);
    for (int k = 0; k < 14; k++) a[k] = a[k] + b[k];
  }
  free(c1);
}

void orc_clear_hydro(vpic_hydro_t *h, const vpic_grid_t *g) {
  memset(h, 0, sizeof(*h) * (size_t)g->nv);
}

/* synchronize_hydro_array for one domain (hydro_array.cc:131-311): local
 * faces doubled (order -x,-y,-z,+x,+y,+z), then the x, y, z periodic face
 * pairs each take lw*h + rw*remote with lw = rw = 1. */
void orc_synchronize_hydro(vpic_hydro_t *h, const vpic_grid_t *g) {
  for (int k = 0; k < 6; k++) {
    if (g->fbc[k] == VPIC_BC_PERIODIC) continue;
    const int X = FACE_AXIS(k), Y = (X + 1) % 3, Z = (X + 2) % 3;
    const int face = FACE_HIGH(k) ? nd(g, X) + 1 : 1;
    int c[3];
    for (int zz = 1; zz <= nd(g, Z) + 1; zz++)
      for (int yy = 1; yy <= nd(g, Y) + 1; yy++) {
        c[X] = face; c[Y] = yy; c[Z] = zz;
        float *a = (float *)(h + fvox(g, c));
        for (int q = 0; q < 14; q++) a[q] *= 2;
      }
  }
  for (int X = 0; X < 3; X++) {
    if (g->fbc[X] != VPIC_BC_PERIODIC || g->fbc[X + 3] != VPIC_BC_PERIODIC)
      continue;
    const int Y = (X + 1) % 3, Z = (X + 2) % 3;
    const float d = X == 0 ? g->dx : X == 1 ? g->dy : g->dz;
    float rw = d, lw = rw + d;
    rw /= lw;
    lw = d / lw;
    lw += lw;
    rw += rw;
    int a[3], b[3];
    for (int zz = 1; zz <= nd(g, Z) + 1; zz++)
      for (int yy = 1; yy <= nd(g, Y) + 1; yy++) {
        a[X] = 1; a[Y] = yy; a[Z] = zz;
        b[X] = nd(g, X) + 1; b[Y] = yy; b[Z] = zz;
        float *pa = (float *)(h + fvox(g, a)), *pb = (float *)(h + fvox(g, b));
        for (int q = 0; q < 14; q++) {
          const float va = pa[q], vb = pb[q];
          pb[q] = lw * vb + rw * va;
          pa[q] = lw * va + rw * vb;
        }
      }
  }
}

/* Takizuka-Abe binary collisions -------------------------------------- */

void orc_philox4x32(const uint32_t ctr[4], const uint32_t key[2],
                    uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; r++) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
    const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
    c1 = (uint32_t)p1;
    c3 = (uint32_t)p0;
    c0 = n0;
    c2 = n2;
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

typedef struct {
  uint32_t key[2], ctr[4], buf[4];
  uint32_t n; /* draws taken in this voxel */
} ta_rng_t;

static void ta_rng_voxel(ta_rng_t *r, int v) {
  r->ctr[1] = (uint32_t)v;
  r->n = 0;
}

static uint32_t ta_u32(ta_rng_t *r) {
  if ((r->n & 3) == 0) {
    r->ctr[0] = r->n >> 2;
    orc_philox4x32(r->ctr, r->key, r->buf);
  }
  return r->buf[r->n++ & 3];
}

static float ta_frand_c0(ta_rng_t *r) {
  return (float)(ta_u32(r) >> 8) * (1.0f / 16777216.0f);
}

static float ta_frandn(ta_rng_t *r) {
  const double u1 = ((double)(ta_u32(r) >> 8) + 1.0) * (1.0 / 16777216.0);
  const double u2 = (double)(ta_u32(r) >> 8) * (1.0 / 16777216.0);
  return (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
}

/* takizuka_abe_collision (takizuka_abe_pipeline.cc:14-79), with the same
 * operations in the same order (float, except the double sin/cos of the
 * scattering phase as there). */
static void ta_collide(vpic_particle_t *pi, vpic_particle_t *pj, float mu_i,
                       float mu_j, float std, ta_rng_t *rng) {
  float dd, ur, urx, ury, urz, tx, ty, tz, t0, t1, t2, wi, wj, stack[3];
  int d0, d1, d2;
  urx = pi->ux - pj->ux;
  ury = pi->uy - pj->uy;
  urz = pi->uz - pj->uz;
  wi = pi->w;
  wj = pj->w;
  /* the axis of the smallest |ur| component (:30-33) */
  t0 = urx * urx; d0 = 0; d1 = 1; d2 = 2; t1 = t0; ur = t0;
  t0 = ury * ury;
  if (t0 < t1) { d0 = 1; d1 = 2; d2 = 0; t1 = t0; }
  ur += t0;
  t0 = urz * urz;
  if (t0 < t1) { d0 = 2; d1 = 0; d2 = 1; }
  ur += t0;
  ur = sqrtf(ur);
  /* a unit vector perpendicular to ur (:35-46) */
  stack[0] = urx; stack[1] = ury; stack[2] = urz;
  t1 = stack[d1];
  t2 = stack[d2];
  t0 = 1 / sqrtf(t1 * t1 + t2 * t2 + FLT_MIN);
  stack[d0] = 0;
  stack[d1] = t0 * t2;
  stack[d2] = -t0 * t1;
  tx = stack[0]; ty = stack[1]; tz = stack[2];
  /* tan(theta/2) = dd ~ N(0, std^2/ur^3), at most 1 standard unit (:48-52) */
  t0 = 1;
  t2 = 1 / ur;
  t1 = std * sqrtf(t2) * t2;
  if (t0 < t1) t1 = t0;
  dd = t1 * ta_frandn(rng);
  t0 = 2 * dd / (1 + dd * dd);
  t1 = (float)(2 * M_PI * ta_frand_c0(rng));
  t2 = (float)(t0 * sin((double)t1));
  t1 = (float)(t0 * ur * cos((double)t1));
  t0 *= -dd;
  /* (1 - cos theta) u + |u| sin theta T (:60-63) */
  stack[0] = (t0 * urx + t1 * tx) + t2 * (ury * tz - urz * ty);
  stack[1] = (t0 * ury + t1 * ty) + t2 * (urz * tx - urx * tz);
  stack[2] = (t0 * urz + t1 * tz) + t2 * (urx * ty - ury * tx);
  /* unequal weights (:65-70) */
  t0 = ta_frand_c0(rng);
  t1 = mu_i;
  t2 = mu_j;
  if (wj < wi && wi * t0 > wj) t1 = 0;
  if (wi < wj && wj * t0 > wi) t2 = 0;
  pi->ux += t1 * stack[0];
  pi->uy += t1 * stack[1];
  pi->uz += t1 * stack[2];
  pj->ux -= t2 * stack[0];
  pj->uy -= t2 * stack[1];
  pj->uz -= t2 * stack[2];
}

void orc_takizuka_abe(vpic_particle_t *spi_p, const int *spi_partition,
                      float qi, float mi, vpic_particle_t *spj_p,
                      const int *spj_partition, float qj, float mj,
                      double cvar0, int interval, uint64_t seed,
                      uint64_t call, const vpic_grid_t *g) {
  const int same = spi_p == spj_p;
  const float dtinterval_dV = (g->dt * (float)interval) / g->dV;
  const float mu = (mi * mj) / (mi + mj);
  const float mu_i = mj / (mi + mj);
  const float mu_j = mi / (mi + mj);
  const double cvar = cvar0 * (qi * qi * qj * qj) / (mu * mu);
  ta_rng_t rng;
  rng.key[0] = (uint32_t)seed;
  rng.key[1] = (uint32_t)(seed >> 32);
  rng.ctr[2] = (uint32_t)call;
  rng.ctr[3] = (uint32_t)(call >> 32);
  const int v1 = g->nx + g->sy * g->ny + g->sz * g->nz + 1; /* VOXEL(nx,ny,nz)+1 */
  for (int v = 0; v < v1; v++) {
    ta_rng_voxel(&rng, v);
    int k0 = spi_partition[v];
    int nk = spi_partition[v + 1] - k0;
    if (!nk) continue;
    /* Fisher-Yates shuffle of species i in this voxel, summing its
     * weights (:113-127) */
    float density_k = 0, density_l, std;
    const int k1 = k0 + nk;
    int i, j, l0, nl;
    for (i = k0; i < k1 - 1; ++i) {
      const uint32_t rn = UINT32_MAX / (uint32_t)(k1 - i);
      do { j = i + (int)(ta_u32(&rng) / rn); } while (j >= k1);
      const vpic_particle_t t = spi_p[j];
      spi_p[j] = spi_p[i];
      spi_p[i] = t;
      density_k += spi_p[i].w;
    }
    density_k += spi_p[i].w;
    if (same) {
      /* an odd count: the first three collide pairwise (:131-144) */
      if (nk % 2 && nk >= 3) {
        std = sqrtf((float)(0.5 * density_k * cvar * dtinterval_dV));
        ta_collide(spi_p + k0, spi_p + k0 + 1, mu_i, mu_j, std, &rng);
        ta_collide(spi_p + k0, spi_p + k0 + 2, mu_i, mu_j, std, &rng);
        ta_collide(spi_p + k0 + 1, spi_p + k0 + 2, mu_i, mu_j, std, &rng);
        nk -= 3;
        k0 += 3;
      }
      nl = nk = nk / 2;
      l0 = k0 + nk;
      if (!nk) continue;
      std = sqrtf((float)(cvar * density_k * dtinterval_dV));
    } else {
      l0 = spj_partition[v];
      nl = spj_partition[v + 1] - l0;
      if (!nl) continue;
      density_l = 0;
      for (i = 0; i < nl; ++i) density_l += spj_p[l0 + i].w;
      std = sqrtf((float)(cvar * (density_l > density_k ? density_k
                                                        : density_l) *
                          dtinterval_dV));
    }
    /* pair the larger set's particles round the smaller's (:167-212) */
    int ii, rn;
    if (nk > nl) {
      ii = nk / nl;
      rn = nk - ii * nl;
      for (i = 0; i < rn; ++i, ++l0)
        for (j = 0; j <= ii; ++j, ++k0)
          ta_collide(spi_p + k0, spj_p + l0, mu_i, mu_j, std, &rng);
      for (; i < nl; ++i, ++l0)
        for (j = 0; j < ii; ++j, ++k0)
          ta_collide(spi_p + k0, spj_p + l0, mu_i, mu_j, std, &rng);
    } else {
      ii = nl / nk;
      rn = nl - ii * nk;
      for (i = 0; i < rn; ++i, ++k0)
        for (j = 0; j <= ii; ++j, ++l0)
          ta_collide(spi_p + k0, spj_p + l0, mu_i, mu_j, std, &rng);
      for (; i < nk; ++i, ++k0)
        for (j = 0; j < ii; ++j, ++l0)
          ta_collide(spi_p + k0, spj_p + l0, mu_i, mu_j, std, &rng);
    }
  }
}
