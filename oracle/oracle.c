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

/* Face k in -x,-y,-z,+x,+y,+z order -> axis and side. */
#define FACE_AXIS(k) ((k) % 3)
#define FACE_HIGH(k) ((k) >= 3)

/* The step functions: FP32 (orc_*), then the same code in FP64 (orc64_*). */
#define ORC_FP64 0
#include "oracle_step.inc"
#undef ORC_FP64
#define ORC_FP64 1
#include "oracle_step.inc"
#undef ORC_FP64

void orc64_grid_from(orc64_grid_t *g64, const vpic_grid_t *g) {
  memset(g64, 0, sizeof(*g64));
  g64->dt = g->dt;
  g64->cvac = g->cvac;
  g64->eps0 = g->eps0;
  g64->dx = g->dx;
  g64->dy = g->dy;
  g64->dz = g->dz;
  g64->dV = g->dV;
  g64->rdx = g->rdx;
  g64->rdy = g->rdy;
  g64->rdz = g->rdz;
  g64->r8V = g->r8V;
  g64->nx = g->nx;
  g64->ny = g->ny;
  g64->nz = g->nz;
  g64->sx = g->sx;
  g64->sy = g->sy;
  g64->sz = g->sz;
  g64->nv = g->nv;
  g64->rangel = g->rangel;
  g64->rangeh = g->rangeh;
  g64->neighbor = g->neighbor;
  for (int k = 0; k < 6; k++) {
    g64->pbc[k] = g->pbc[k];
    g64->fbc[k] = g->fbc[k];
  }
}

/* Float-only parts below use the FP32 instantiation's names. */

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

void orc_field_params_vacuum(vpic_field_params_t *mp, float damp) {
  /* sfa.cc:111-126 with sigma = 0, eps = mu = 1. */
  mp->decayx = mp->decayy = mp->decayz = (float)exp(-0.0);
  mp->drivex = mp->drivey = mp->drivez = (float)(1. / 1.0);
  mp->rmux = mp->rmuy = mp->rmuz = (float)(1. / 1.0);
  mp->epsx = mp->epsy = mp->epsz = 1.0f;
  mp->damp = damp;
  mp->nonconductive = 1.0f; /* sfa.cc:144 with ax = ay = az = 0 */
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
 * operations in the same order, all in float.  The macro takes three draws
 * per collision, in this order: frandn (the scattering angle), frand_c0
 * (the phase) and frand_c0 (the unequal-weight test); they are passed in
 * as dn, u_phase and u_w. */
static void ta_collide_draws(vpic_particle_t *pi, vpic_particle_t *pj,
                             float mu_i, float mu_j, float std, float dn,
                             float u_phase, float u_w) {
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
  dd = t1 * dn;
  t0 = 2 * dd / (1 + dd * dd);
  t1 = (float)(2 * M_PI * u_phase);
  /* sin(t1), cos(t1) of a float are the float overloads in the reference's
   * C++; here the double functions rounded to float (the correctly rounded
   * value, as the device forms it), multiplied in float. */
  t2 = t0 * (float)sin((double)t1);
  t1 = t0 * ur * (float)cos((double)t1);
  t0 *= -dd;
  /* (1 - cos theta) u + |u| sin theta T (:60-63) */
  stack[0] = (t0 * urx + t1 * tx) + t2 * (ury * tz - urz * ty);
  stack[1] = (t0 * ury + t1 * ty) + t2 * (urz * tx - urx * tz);
  stack[2] = (t0 * urz + t1 * tz) + t2 * (urx * ty - ury * tx);
  /* unequal weights (:65-70) */
  t0 = u_w;
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

static void ta_collide(vpic_particle_t *pi, vpic_particle_t *pj, float mu_i,
                       float mu_j, float std, ta_rng_t *rng) {
  /* the macro's draws, in the order it takes them */
  const float dn = ta_frandn(rng);
  const float u_phase = ta_frand_c0(rng);
  const float u_w = ta_frand_c0(rng);
  ta_collide_draws(pi, pj, mu_i, mu_j, std, dn, u_phase, u_w);
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
