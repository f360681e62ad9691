// dropin_check.cc -- a reference-side program: compiled against lanl/vpic's
// own headers (its particle_t, species_t, grid_t, ... definitions) and
// linked against libvpichip.so for the kernels, as a VPIC build using the
// drop-in would be.  Each kernel's result is compared with the reference's
// own scalar kernels (libvpicref.so, compiled from the lanl/vpic sources)
// or, for the field kernels' boundary handling, with the oracle.
//
// TEST INFRASTRUCTURE ONLY.  Prints "ok <check>" or "FAIL <check>: ..." per
// check; the exit status is the number of failed checks.  With --layout it
// only compares the struct layouts (no device needed).
#define IN_sfa
#include "src/field_advance/standard/sfa_private.h"
#include "src/species_advance/species_advance.h"
#define IN_boundary
#include "src/boundary/boundary_private.h"

#define VPIC_DROPIN_NO_TYPES
#include "../../include/vpic_dropin.h"
#include "../../include/vpic_dropin_layout.h"
#include "../../include/vpic_hip.h"
#include "../oracle.h"

#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <chrono>
#include <vector>

extern "C" {
int ref_advance_p(vpic_particle_t *p, int np, float q, float m,
                  vpic_particle_mover_t *pm, int max_nm, int *n_ignored,
                  vpic_accumulator_t *a, const vpic_interpolator_t *f,
                  const vpic_grid_t *g, int n_pipeline);
int ref_move_p(vpic_particle_t *p0, vpic_particle_mover_t *pm,
               vpic_accumulator_t *a0, const vpic_grid_t *g, float qsp);
double ref_energy_p(const vpic_particle_t *p, int np, float q, float m,
                    const vpic_interpolator_t *f, const vpic_grid_t *g);
void ref_center_p(vpic_particle_t *p, int np, float q, float m,
                  const vpic_interpolator_t *f, const vpic_grid_t *g,
                  int uncenter);
void ref_sort_p(vpic_particle_t *p, int np, int *partition,
                const vpic_grid_t *g);
void ref_load_interpolator(vpic_interpolator_t *fi, const vpic_field_t *f,
                           const vpic_grid_t *g);
void ref_unload_accumulator(vpic_field_t *f, const vpic_accumulator_t *a,
                            const vpic_grid_t *g);
void ref_energy_f(double en[6], const vpic_field_t *f, const vpic_grid_t *g,
                  const vpic_field_params_t *mp);
void ref_accumulate_hydro_p(vpic_hydro_t *h, const vpic_particle_t *p, int np,
                            float q, float m, const vpic_interpolator_t *f,
                            const vpic_grid_t *g);
void ref_accumulate_rhob(vpic_field_t *f, const vpic_particle_t *p, int np,
                         float qsp, const vpic_grid_t *g);
}

static int g_fail = 0;

static void report(bool ok, const char *name, const char *fmt = "", ...) {
  if (ok) {
    printf("ok %s\n", name);
  } else {
    char msg[400];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(msg, sizeof msg, fmt, ap);
    va_end(ap);
    printf("FAIL %s: %s\n", name, msg);
    g_fail++;
  }
  fflush(stdout);
}

// ---- inputs ----------------------------------------------------------------
struct Rng {
  uint64_t s;
  explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ull + 1) {}
  uint64_t next() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
  }
  double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
  double normal() {
    double u = uniform(), v = uniform();
    if (u < 1e-300) u = 1e-300;
    return sqrt(-2 * log(u)) * cos(6.283185307179586 * v);
  }
};

struct Box {
  vpic_grid_t G;               // the C-ABI grid (host neighbor table)
  grid_t rg;                   // the reference's grid_t over the same data
  std::vector<int64_t> nb;
  int64_t range[2];
};

// A periodic box; x_faces < 0 makes the x faces that particle BC.
static void make_box(Box &b, int nx, int ny, int nz, int x_faces) {
  const float dt = 0.5f, cvac = 1.0f, eps0 = 1.0f;
  vpic_hip_grid_periodic(&b.G, 0, 0, 0, nx, ny, nz, nx, ny, nz, dt, cvac,
                         eps0);
  b.nb.resize((size_t)6 * b.G.nv);
  vpic_hip_grid_neighbor_table(&b.G, b.nb.data());
  if (x_faces < 0) {
    for (int z = 1; z <= nz; z++)
      for (int y = 1; y <= ny; y++) {
        int64_t lo = 1 + (nx + 2) * (y + (int64_t)(ny + 2) * z);
        int64_t hi = nx + (nx + 2) * (y + (int64_t)(ny + 2) * z);
        b.nb[6 * lo + 0] = x_faces;
        b.nb[6 * hi + 3] = x_faces;
      }
    b.G.pbc[0] = b.G.pbc[3] = x_faces;
  }
  b.G.neighbor = b.nb.data();
  memset(&b.rg, 0, sizeof b.rg);
  grid_t &r = b.rg;
  r.dt = b.G.dt;
  r.cvac = b.G.cvac;
  r.eps0 = b.G.eps0;
  r.x1 = nx;
  r.y1 = ny;
  r.z1 = nz;
  r.nx = nx;
  r.ny = ny;
  r.nz = nz;
  r.dx = b.G.dx;
  r.dy = b.G.dy;
  r.dz = b.G.dz;
  r.dV = b.G.dV;
  r.rdx = b.G.rdx;
  r.rdy = b.G.rdy;
  r.rdz = b.G.rdz;
  r.r8V = b.G.r8V;
  r.sx = b.G.sx;
  r.sy = b.G.sy;
  r.sz = b.G.sz;
  r.nv = b.G.nv;
  for (int k = 0; k < 27; k++) r.bc[k] = 0;  // every face: this rank
  if (x_faces < 0) {
    r.bc[BOUNDARY(-1, 0, 0)] = anti_symmetric_fields;
    r.bc[BOUNDARY(1, 0, 0)] = anti_symmetric_fields;
  }
  b.range[0] = 0;
  b.range[1] = b.G.nv;
  r.range = b.range;
  r.neighbor = b.nb.data();
  r.rangel = 0;
  r.rangeh = b.G.nv - 1;
  r.mp = NULL;
  if (x_faces < 0) {
    b.G.fbc[0] = b.G.fbc[3] = VPIC_BC_ANTI_SYMMETRIC_FIELDS;
  }
}

static void fill_particles(const Box &b, particle_t *p, int ppc, float uth,
                           uint64_t seed) {
  Rng r(seed);
  int n = 0;
  const int nx = b.G.nx, ny = b.G.ny, nz = b.G.nz;
  for (int z = 1; z <= nz; z++)
    for (int y = 1; y <= ny; y++)
      for (int x = 1; x <= nx; x++)
        for (int k = 0; k < ppc; k++, n++) {
          p[n].dx = (float)(2 * r.uniform() - 1);
          p[n].dy = (float)(2 * r.uniform() - 1);
          p[n].dz = (float)(2 * r.uniform() - 1);
          p[n].i = x + (nx + 2) * (y + (ny + 2) * z);
          p[n].ux = (float)(uth * r.normal());
          p[n].uy = (float)(uth * r.normal());
          p[n].uz = (float)(uth * r.normal());
          p[n].w = (float)(0.5 + r.uniform());
        }
  // shuffle so that sort_p has work to do
  for (int k = n - 1; k > 0; k--) {
    int j = (int)(r.next() % (uint64_t)(k + 1));
    particle_t t = p[k];
    p[k] = p[j];
    p[j] = t;
  }
}

static void fill_fields(field_t *f, int nv, float amp, uint64_t seed) {
  Rng r(seed);
  memset(f, 0, sizeof(field_t) * nv);
  for (int v = 0; v < nv; v++) {
    float *x = &f[v].ex;
    for (int c = 0; c < 16; c++) x[c] = (float)(amp * (2 * r.uniform() - 1));
  }
}

static bool same(const void *a, const void *b, size_t n) {
  return memcmp(a, b, n) == 0;
}

static double acc_err(const accumulator_t *a, const vpic_accumulator_t *b,
                      int nv) {
  double num = 0, den = 0;
  for (int v = 0; v < nv; v++)
    for (int c = 0; c < 12; c++) {
      double x = (&a[v].jx[0])[c], y = (&b[v].jx[0])[c];
      num = fmax(num, fabs(x - y));
      den = fmax(den, fabs(y));
    }
  return den > 0 ? num / den : num;
}

// ---- the checks --------------------------------------------------------------
static void check_layout() {
  const int64_t mine[] = {VPIC_DROPIN_LAYOUT(VPIC_DROPIN_SIZE,
                                             VPIC_DROPIN_OFFSET)};
  const int n = (int)(sizeof mine / sizeof mine[0]);
  std::vector<int64_t> lib(n + 8, -1);
  int count = vpic_hip_dropin_layout(lib.data(), (int)lib.size());
  int bad = -1;
  for (int k = 0; k < n && bad < 0; k++)
    if (lib[k] != mine[k]) bad = k;
  report(count == n && bad < 0, "layout",
         "count %d vs %d, first mismatch at entry %d", count, n, bad);
}

struct Species {
  species_t sp;
  std::vector<particle_t> p;
  std::vector<particle_mover_t> pm;
  std::vector<int> part;
};

static void make_species(Species &s, Box &b, float q, float m, int ppc,
                         float uth, uint64_t seed) {
  const int np = ppc * b.G.nx * b.G.ny * b.G.nz;
  s.p.assign(np + 16, particle_t());
  s.pm.assign(np + 16, particle_mover_t());
  s.part.assign(b.G.nv + 1, 0);
  fill_particles(b, s.p.data(), ppc, uth, seed);
  memset(&s.sp, 0, sizeof s.sp);
  s.sp.name = (char *)"check";
  s.sp.q = q;
  s.sp.m = m;
  s.sp.np = np;
  s.sp.max_np = np + 16;
  s.sp.p = s.p.data();
  s.sp.nm = 0;
  s.sp.max_nm = np + 16;
  s.sp.pm = s.pm.data();
  s.sp.last_sorted = -1;
  s.sp.sort_interval = 20;
  s.sp.partition = s.part.data();
  s.sp.g = &b.rg;
}

struct Arrays {
  std::vector<interpolator_t> fi;
  std::vector<accumulator_t> acc;  // 2 copies (n_pipeline = 1)
  std::vector<field_t> f;
  interpolator_array_t ia;
  accumulator_array_t aa;
  field_array_t fa;
  material_coefficient_t mc;
  sfa_params_t params;
};

static void make_arrays(Arrays &A, Box &b, uint64_t seed) {
  const int nv = b.G.nv;
  const int stride = POW2_CEIL(nv, 2);
  A.fi.assign(nv, interpolator_t());
  A.acc.assign((size_t)2 * stride, accumulator_t());
  A.f.assign(nv, field_t());
  fill_fields(A.f.data(), nv, 0.3f, seed);
  A.ia.i = A.fi.data();
  A.ia.g = &b.rg;
  A.aa.a = A.acc.data();
  A.aa.n_pipeline = 1;
  A.aa.stride = stride;
  A.aa.g = &b.rg;
  vpic_field_params_t mp;
  vpic_hip_field_params_vacuum(&mp, 0.01f);
  memset(&A.mc, 0, sizeof A.mc);
  A.mc.decayx = mp.decayx;
  A.mc.drivex = mp.drivex;
  A.mc.decayy = mp.decayy;
  A.mc.drivey = mp.drivey;
  A.mc.decayz = mp.decayz;
  A.mc.drivez = mp.drivez;
  A.mc.rmux = mp.rmux;
  A.mc.rmuy = mp.rmuy;
  A.mc.rmuz = mp.rmuz;
  A.mc.nonconductive = 1;
  A.mc.epsx = mp.epsx;
  A.mc.epsy = mp.epsy;
  A.mc.epsz = mp.epsz;
  A.params.mc = &A.mc;
  A.params.n_mc = 1;
  A.params.damp = mp.damp;
  memset(&A.fa, 0, sizeof A.fa);
  A.fa.f = A.f.data();
  A.fa.g = &b.rg;
  A.fa.params = &A.params;
}

static vpic_field_params_t params_of(const Arrays &A) {
  vpic_field_params_t mp;
  vpic_hip_field_params_vacuum(&mp, A.params.damp);
  return mp;
}

// One push of a species through the drop-in and through the reference.
static void check_push(const char *name, Species &s, Arrays &A, Box &b) {
  const int np = s.sp.np, nv = b.G.nv;
  std::vector<particle_t> p0(s.p.begin(), s.p.begin() + np);
  std::vector<vpic_particle_mover_t> rpm(s.sp.max_nm);
  std::vector<vpic_accumulator_t> racc(nv);
  int ign = 0;
  clear_accumulator_array(&A.aa);
  advance_p(&s.sp, &A.aa, &A.ia);
  reduce_accumulator_array(&A.aa);
  int rnm = ref_advance_p((vpic_particle_t *)p0.data(), np, s.sp.q, s.sp.m,
                          rpm.data(), s.sp.max_nm, &ign, racc.data(),
                          (const vpic_interpolator_t *)A.fi.data(), &b.G, 1);
  bool ok_p = same(s.p.data(), p0.data(), sizeof(particle_t) * np);
  bool ok_m = s.sp.nm == rnm &&
              same(s.pm.data(), rpm.data(), sizeof(particle_mover_t) * rnm);
  double e = acc_err(A.acc.data(), racc.data(), nv);
  report(ok_p && ok_m && e <= 1e-5, name,
         "particles %s, movers %s (nm %d vs %d), accumulator rel err %.3g",
         ok_p ? "same" : "DIFFER", ok_m ? "same" : "DIFFER", s.sp.nm, rnm, e);
}

static void check_species_kernels() {
  Box b;
  make_box(b, 8, 6, 5, 0);
  Arrays A;
  make_arrays(A, b, 7);
  // load_interpolator_array
  std::vector<vpic_interpolator_t> rfi(b.G.nv);
  memset(rfi.data(), 0, sizeof(vpic_interpolator_t) * b.G.nv);
  load_interpolator_array(&A.ia, &A.fa);
  ref_load_interpolator(rfi.data(), (const vpic_field_t *)A.f.data(), &b.G);
  report(same(A.fi.data(), rfi.data(), sizeof(interpolator_t) * b.G.nv),
         "load_interpolator_array", "interpolators differ");

  Species e;
  make_species(e, b, -1.0f, 1.0f, 9, 0.3f, 11);
  check_push("advance_p (unsorted)", e, A, b);

  // sort_p
  {
    std::vector<particle_t> p0(e.p.begin(), e.p.begin() + e.sp.np);
    std::vector<int> rpart(b.G.nv + 1);
    sort_p(&e.sp);
    ref_sort_p((vpic_particle_t *)p0.data(), e.sp.np, rpart.data(), &b.G);
    bool okp = same(e.p.data(), p0.data(), sizeof(particle_t) * e.sp.np);
    // entries [0, nv): the reference leaves partition[nv] unwritten
    bool okq = same(e.part.data(), rpart.data(), sizeof(int) * b.G.nv);
    report(okp && okq, "sort_p", "particles %s, partition %s",
           okp ? "same" : "DIFFER", okq ? "same" : "DIFFER");
  }
  check_push("advance_p (after sort_p: tiled)", e, A, b);

  // energy_p
  {
    double en = energy_p(&e.sp, &A.ia);
    double ren = ref_energy_p((const vpic_particle_t *)e.p.data(), e.sp.np,
                              e.sp.q, e.sp.m,
                              (const vpic_interpolator_t *)A.fi.data(), &b.G);
    report(fabs(en - ren) <= 1e-12 * fabs(ren), "energy_p", "%.17g vs %.17g",
           en, ren);
  }
  // center_p / uncenter_p
  for (int unc = 0; unc <= 1; unc++) {
    std::vector<particle_t> p0(e.p.begin(), e.p.begin() + e.sp.np);
    if (unc)
      uncenter_p(&e.sp, &A.ia);
    else
      center_p(&e.sp, &A.ia);
    ref_center_p((vpic_particle_t *)p0.data(), e.sp.np, e.sp.q, e.sp.m,
                 (const vpic_interpolator_t *)A.fi.data(), &b.G, unc);
    report(same(e.p.data(), p0.data(), sizeof(particle_t) * e.sp.np),
           unc ? "uncenter_p" : "center_p", "momenta differ");
  }
  // move_p on one particle (the way inject_particle calls it, misc.cc:95)
  {
    particle_t *p = e.p.data() + 5;
    particle_t keep = *p;
    particle_mover_t pm = {1.7f, -0.9f, 0.4f, 5};
    vpic_particle_mover_t rpm = {1.7f, -0.9f, 0.4f, 5};
    std::vector<vpic_accumulator_t> racc(b.G.nv);
    memset(racc.data(), 0, sizeof(vpic_accumulator_t) * b.G.nv);
    clear_accumulator_array(&A.aa);
    int ret = move_p(e.p.data(), &pm, A.aa.a, &b.rg, e.sp.q);
    particle_t mine = *p;
    *p = keep;
    std::vector<particle_t> p0(e.p.begin(), e.p.begin() + e.sp.np);
    int rret = ref_move_p((vpic_particle_t *)p0.data(), &rpm, racc.data(),
                          &b.G, e.sp.q);
    bool okp = same(&mine, &p0[5], sizeof mine);
    bool okm = same(&pm, &rpm, sizeof pm);
    bool oka = same(A.acc.data(), racc.data(),
                    sizeof(vpic_accumulator_t) * b.G.nv);
    *p = mine;
    report(ret == rret && okp && okm && oka, "move_p",
           "ret %d vs %d, particle %d, mover %d, accumulator %d", ret, rret,
           okp, okm, oka);
  }
  // unload_accumulator_array after a push
  {
    clear_accumulator_array(&A.aa);
    advance_p(&e.sp, &A.aa, &A.ia);
    std::vector<field_t> f0(A.f);
    unload_accumulator_array(&A.fa, &A.aa);
    ref_unload_accumulator((vpic_field_t *)f0.data(),
                           (const vpic_accumulator_t *)A.acc.data(), &b.G);
    report(same(A.f.data(), f0.data(), sizeof(field_t) * b.G.nv),
           "unload_accumulator_array", "fields differ");
  }
  // accumulate_rho_p + synchronize_rho against the oracle
  {
    int rc = vpic_hip_install_field_kernels(&A.fa);
    report(rc == 0, "install_field_kernels", "%s", vpic_hip_last_error());
    std::vector<field_t> f0(A.f);
    A.fa.kernel->clear_rhof(&A.fa);
    accumulate_rho_p(&A.fa, &e.sp);
    A.fa.kernel->synchronize_rho(&A.fa);
    orc_clear_rhof((vpic_field_t *)f0.data(), &b.G);
    orc_accumulate_rho_p((vpic_field_t *)f0.data(),
                         (const vpic_particle_t *)e.p.data(), e.sp.np, e.sp.q,
                         &b.G);
    orc_synchronize_rho((vpic_field_t *)f0.data(), &b.G);
    double num = 0, den = 0;
    for (int v = 0; v < b.G.nv; v++) {
      num = fmax(num, fabs((double)A.f[v].rhof - f0[v].rhof));
      den = fmax(den, fabs((double)f0[v].rhof));
    }
    report(num <= 1e-5 * den, "accumulate_rho_p + synchronize_rho",
           "max |d rhof| %.3g of %.3g", num, den);
  }
  // accumulate_rhob of single particles (as advance.cc:78-101 calls it)
  // against the reference's own function, one particle after another
  {
    std::vector<field_t> f0(A.f);
    const int n = e.sp.np < 64 ? e.sp.np : 64;
    for (int k = 0; k < n; k++)
      accumulate_rhob(A.f.data(), &e.p[k], e.sp.g, e.sp.q);
    ref_accumulate_rhob((vpic_field_t *)f0.data(),
                        (const vpic_particle_t *)e.p.data(), n, e.sp.q,
                        &b.G);
    report(same(A.f.data(), f0.data(), sizeof(field_t) * b.G.nv),
           "accumulate_rhob", "fields differ");
  }
  // hydro moments in the sequence of dump.cc:240-242 (synchronize reduces
  // first, hydro_array.cc:145-147); sorted (tiled) and unsorted (flat)
  {
    Species u;
    make_species(u, b, 1.0f, 1836.0f, 5, 0.02f, 13);
    Species *sps[2] = {&e, &u};
    const char *names[2] = {"hydro moments (sorted species: tiled)",
                            "hydro moments (unsorted species: flat)"};
    const int nv = b.G.nv;
    const int stride = POW2_CEIL(nv, 2);
    for (int k = 0; k < 2; k++) {
      std::vector<hydro_t> hh((size_t)2 * stride);
      for (size_t j = 0; j < hh.size(); j++) hh[j].rho = 99.0f;  // stale
      hydro_array_t ha;
      ha.h = hh.data();
      ha.n_pipeline = 1;
      ha.stride = stride;
      ha.g = &b.rg;
      clear_hydro_array(&ha);
      accumulate_hydro_p(&ha, &sps[k]->sp, &A.ia);
      reduce_hydro_array(&ha);
      synchronize_hydro_array(&ha);
      std::vector<vpic_hydro_t> r(nv);
      memset(r.data(), 0, sizeof(vpic_hydro_t) * nv);
      ref_accumulate_hydro_p(r.data(), (const vpic_particle_t *)sps[k]->p.data(),
                             sps[k]->sp.np, sps[k]->sp.q, sps[k]->sp.m,
                             (const vpic_interpolator_t *)A.fi.data(), &b.G);
      orc_synchronize_hydro(r.data(), &b.G);
      double num = 0, den = 0;
      for (int v = 0; v < nv; v++)
        for (int c = 0; c < 14; c++) {
          const double x = (&hh[v].jx)[c], y = (&r[v].jx)[c];
          num = fmax(num, fabs(x - y));
          den = fmax(den, fabs(y));
        }
      bool copy1 = true;  // the host's second copy was cleared
      for (int v = 0; v < nv; v++) copy1 = copy1 && hh[stride + v].rho == 0;
      report(num <= 1e-5 * den && copy1, names[k],
             "max |d| %.3g of %.3g, copy 1 %s", num, den,
             copy1 ? "clear" : "NOT CLEARED");
    }
    vpic_hip_dropin_release(u.sp.p);
  }
  vpic_hip_dropin_release_all();
}

// Field kernels through fa->kernel after vpic_hip_install_field_kernels,
// against the oracle (whole step, including ghosts and boundary faces).
static void check_field_kernels(int x_faces, const char *name) {
  Box b;
  make_box(b, 7, 5, 6, x_faces);
  Arrays A;
  make_arrays(A, b, 21);
  int rc = vpic_hip_install_field_kernels(&A.fa);
  if (rc) {
    report(false, name, "install failed: %s", vpic_hip_last_error());
    return;
  }
  std::vector<field_t> f0(A.f);
  vpic_field_params_t mp = params_of(A);
  field_advance_kernels_t *k = A.fa.kernel;
  k->clear_jf(&A.fa);
  k->synchronize_jf(&A.fa);
  k->advance_b(&A.fa, 0.5f);
  k->advance_e(&A.fa, 1.0f);
  k->advance_b(&A.fa, 0.5f);
  vpic_field_t *o = (vpic_field_t *)f0.data();
  orc_clear_jf(o, &b.G);
  orc_synchronize_jf(o, &b.G);
  orc_advance_b(o, &b.G, 0.5f);
  orc_vacuum_advance_e(o, &b.G, &mp);
  orc_advance_b(o, &b.G, 0.5f);
  bool okf = same(A.f.data(), f0.data(), sizeof(field_t) * b.G.nv);
  double en[6], ren[6];
  k->energy_f(en, &A.fa);
  ref_energy_f(ren, o, &b.G, &mp);
  double worst = 0;
  for (int c = 0; c < 6; c++)
    worst = fmax(worst, fabs(en[c] - ren[c]) / fmax(1e-300, fabs(ren[c])));
  report(okf && worst <= 1e-12, name, "fields %s, energy_f rel err %.3g",
         okf ? "same" : "DIFFER", worst);
  // the divergence-cleaning / initialization entries of the table
  k->compute_curl_b(&A.fa);
  orc_compute_curl_b(o, &b.G, &mp);
  k->compute_rhob(&A.fa);
  orc_compute_rhob(o, &b.G, &mp);
  k->compute_div_e_err(&A.fa);
  orc_compute_div_e_err(o, &b.G, &mp);
  const double re = k->compute_rms_div_e_err(&A.fa);
  const double oe = orc_compute_rms_div_e_err(o, &b.G);
  k->clean_div_e(&A.fa);
  orc_clean_div_e(o, &b.G, &mp);
  k->compute_div_b_err(&A.fa);
  orc_compute_div_b_err(o, &b.G);
  const double rb = k->compute_rms_div_b_err(&A.fa);
  const double ob = orc_compute_rms_div_b_err(o, &b.G);
  k->clean_div_b(&A.fa);
  orc_clean_div_b(o, &b.G);
  const double rs = k->synchronize_tang_e_norm_b(&A.fa);
  const double os = orc_synchronize_tang_e_norm_b(o, &b.G);
  const bool okc = same(A.f.data(), f0.data(), sizeof(field_t) * b.G.nv);
  const double wc = fmax(fmax(fabs(re - oe) / fmax(1e-300, fabs(oe)),
                              fabs(rb - ob) / fmax(1e-300, fabs(ob))),
                         fabs(rs - os) / fmax(1e-300, fabs(os)));
  report(okc && wc <= 1e-12, name,
         "cleaning kernels: fields %s, rms/err rel err %.3g",
         okc ? "same" : "DIFFER", wc);
  vpic_hip_dropin_release_all();
}

// Movers returned by an absorbing x wall, and the reference's rollback.
static void check_absorb() {
  Box b;
  make_box(b, 6, 5, 4, absorb_particles);
  Arrays A;
  make_arrays(A, b, 31);
  load_interpolator_array(&A.ia, &A.fa);
  Species e;
  make_species(e, b, -1.0f, 1.0f, 7, 0.6f, 41);
  check_push("advance_p (absorbing x walls, movers)", e, A, b);
  report(e.sp.nm > 0, "absorbing walls return movers", "nm = %d", e.sp.nm);
  vpic_hip_dropin_release_all();
}

// A custom particle boundary for the check: the particle comes back as an
// injector with its normal momentum and remaining displacement reversed.
static int g_custom_hits = 0;
static int reflux_handler(void *params, species_t *sp, particle_t *p,
                          particle_mover_t *pm, particle_injector_t *pi,
                          int max_pi, int face) {
  g_custom_hits++;
  if (max_pi < 1) return 0;
  const int axis = face % 3;
  pi->dx = p->dx;
  pi->dy = p->dy;
  pi->dz = p->dz;
  pi->i = p->i;
  pi->ux = p->ux;
  pi->uy = p->uy;
  pi->uz = p->uz;
  pi->w = p->w;
  pi->dispx = pm->dispx;
  pi->dispy = pm->dispy;
  pi->dispz = pm->dispz;
  (&pi->ux)[axis] = -(&pi->ux)[axis];
  (&pi->dispx)[axis] = -(&pi->dispx)[axis];
  pi->sp_id = sp->id;
  return 1;
}

// boundary_p through the drop-in after a push against absorbing x walls
// and custom z walls, against the oracle's boundary_p, the same handler
// on the oracle's hits, and the oracle's injection.
static void check_boundary() {
  Box b;
  make_box(b, 6, 5, 4, absorb_particles);
  const int nx = b.G.nx, ny = b.G.ny, nz = b.G.nz;
  for (int y = 1; y <= ny; y++)
    for (int x = 1; x <= nx; x++) {
      b.nb[6 * (x + (nx + 2) * (y + (ny + 2) * 1)) + 2] = -3;
      b.nb[6 * (x + (nx + 2) * (y + (ny + 2) * nz)) + 5] = -3;
    }
  b.G.pbc[2] = b.G.pbc[5] = -3;
  // a face with a particle boundary has a local field boundary too
  b.rg.bc[BOUNDARY(0, 0, -1)] = anti_symmetric_fields;
  b.rg.bc[BOUNDARY(0, 0, 1)] = anti_symmetric_fields;
  b.G.fbc[2] = b.G.fbc[5] = VPIC_BC_ANTI_SYMMETRIC_FIELDS;
  Arrays A;
  make_arrays(A, b, 61);
  load_interpolator_array(&A.ia, &A.fa);
  int rc = vpic_hip_install_field_kernels(&A.fa);
  Species e;
  make_species(e, b, -1.0f, 1.0f, 7, 0.6f, 71);
  e.sp.id = 0;
  clear_accumulator_array(&A.aa);
  advance_p(&e.sp, &A.aa, &A.ia);
  const int np0 = e.sp.np, nm0 = e.sp.nm;
  std::vector<particle_t> p0(e.p);
  std::vector<particle_mover_t> pm0(e.pm);
  std::vector<field_t> f0(A.f);
  std::vector<accumulator_t> a0(A.acc);
  particle_bc_t pbc;
  memset(&pbc, 0, sizeof pbc);
  pbc.interact = (particle_bc_func_t)reflux_handler;
  pbc.id = -3;
  g_custom_hits = 0;
  boundary_p(&pbc, &e.sp, &A.fa, &A.aa);
  const int dev_hits = g_custom_hits;
  // the oracle, with the same handler on its hits
  std::vector<vpic_boundary_hit_t> hits(nm0 > 0 ? nm0 : 1);
  int32_t counts[2];
  int np = orc_boundary_p((vpic_particle_t *)p0.data(), np0,
                          (const vpic_particle_mover_t *)pm0.data(), nm0,
                          (vpic_field_t *)f0.data(), e.sp.q, &b.G,
                          hits.data(), counts);
  std::vector<particle_injector_t> ci;
  g_custom_hits = 0;
  for (int m = nm0 - 1; m >= 0; m--) {
    if (hits[m].code != -3) continue;
    particle_injector_t c;
    if (reflux_handler(nullptr, &e.sp, (particle_t *)&hits[m].p,
                       (particle_mover_t *)&hits[m].m, &c, 1, hits[m].face))
      ci.push_back(c);
  }
  int nm = 0;
  int32_t ic[2];
  orc_inject_p((vpic_particle_t *)p0.data(), &np, e.sp.max_np,
               (vpic_particle_mover_t *)pm0.data(), &nm, e.sp.max_nm,
               (const vpic_particle_injector_t *)ci.data(), (int)ci.size(),
               (vpic_accumulator_t *)a0.data(), e.sp.q, &b.G, ic);
  bool okn = e.sp.np == np && e.sp.nm == nm && dev_hits == g_custom_hits;
  bool okp = okn && same(e.p.data(), p0.data(), sizeof(particle_t) * np);
  bool okm = okn && same(e.pm.data(), pm0.data(),
                         sizeof(particle_mover_t) * nm);
  double num = 0, den = 0;
  for (int v = 0; v < b.G.nv; v++) {
    num = fmax(num, fabs((double)A.f[v].rhob - f0[v].rhob));
    den = fmax(den, fabs((double)f0[v].rhob));
  }
  const double ea = acc_err(A.acc.data(), (const vpic_accumulator_t *)a0.data(),
                            b.G.nv);
  report(rc == 0 && nm0 > 0 && g_custom_hits > 0 && okp && okm &&
             num <= 1e-5 * den && ea <= 1e-5,
         "boundary_p (absorbing x walls, custom z walls, injection)",
         "movers %d, np %d vs %d, nm %d vs %d, handler calls %d vs %d, "
         "particles %s, movers %s, rhob err %.3g of %.3g, acc err %.3g",
         nm0, e.sp.np, np, e.sp.nm, nm, dev_hits, g_custom_hits,
         okp ? "same" : "DIFFER", okm ? "same" : "DIFFER", num, den, ea);
  vpic_hip_dropin_release_all();
}

// Resident mode: several steps with no transfers in between; the state is
// read back after each step (a download does not disturb the device copy)
// and compared with the reference stepping in lockstep.
static void check_resident() {
  Box b;
  make_box(b, 8, 8, 4, 0);
  Arrays A;
  make_arrays(A, b, 51);
  load_interpolator_array(&A.ia, &A.fa);
  Species e;
  make_species(e, b, -1.0f, 1.0f, 6, 0.3f, 61);
  std::vector<particle_t> p0(e.p.begin(), e.p.begin() + e.sp.np);
  const std::vector<particle_t> pinit(p0);
  std::vector<vpic_accumulator_t> racc(b.G.nv);
  std::vector<vpic_particle_mover_t> rpm(e.sp.max_nm);
  std::vector<int> rpart(b.G.nv + 1);
  vpic_hip_dropin_resident(1);
  sort_p(&e.sp);
  ref_sort_p((vpic_particle_t *)p0.data(), e.sp.np, rpart.data(), &b.G);
  // no transfer yet: the host still holds the unsorted particles
  const bool untouched =
      same(e.p.data(), pinit.data(), sizeof(particle_t) * e.sp.np);
  int bad_step = -1;
  double worst = 0;
  bool okp = true;
  for (int step = 0; step < 4; step++) {
    clear_accumulator_array(&A.aa);
    advance_p(&e.sp, &A.aa, &A.ia);
    int ign = 0;
    // ref_advance_p adds into its accumulator, as advance_p does
    memset(racc.data(), 0, sizeof(vpic_accumulator_t) * b.G.nv);
    ref_advance_p((vpic_particle_t *)p0.data(), e.sp.np, e.sp.q, e.sp.m,
                  rpm.data(), e.sp.max_nm, &ign, racc.data(),
                  (const vpic_interpolator_t *)A.fi.data(), &b.G, 1);
    vpic_hip_species_sync(&e.sp, 0);
    vpic_hip_accumulator_array_sync(&A.aa, 0);
    const bool sp_ok =
        same(e.p.data(), p0.data(), sizeof(particle_t) * e.sp.np);
    const double err = acc_err(A.acc.data(), racc.data(), b.G.nv);
    worst = fmax(worst, err);
    if ((!sp_ok || err > 1e-5) && bad_step < 0) bad_step = step;
    okp = okp && sp_ok;
  }
  vpic_hip_dropin_resident(0);
  report(untouched && okp && bad_step < 0, "resident mode (sort + 4 pushes)",
         "host untouched before sync %d, particles %s, worst acc err %.3g "
         "(first bad step %d)",
         untouched, okp ? "same" : "DIFFER", worst, bad_step);
  vpic_hip_dropin_release_all();
}

static int g_handled = 0;
static void handler(const char *msg) { g_handled++; }

static void check_errors() {
  vpic_hip_dropin_error_handler(handler);
  advance_p(NULL, NULL, NULL);
  sort_p(NULL);
  Box b;
  make_box(b, 4, 4, 4, 0);
  Arrays A;
  make_arrays(A, b, 3);
  A.fa.kernel->advance_e = NULL;
  vpic_hip_fa_advance_e(&A.fa, 0.5f);  // frac != 1: the reference's ERROR
  vpic_hip_dropin_error_handler(NULL);
  report(g_handled == 3, "errors reach the handler", "%d of 3", g_handled);
}

// The reference's step (advance.cc:15-207 for a periodic vacuum box with
// two species) through the reference-named entry points in resident mode,
// timed: the drop-in at scale, with no particle crossing PCIe between
// steps.  Electrons (uth 0.1) and ions (m 25, uth 0.02), ppc each, sorted
// every 20 steps, density 1; `warm` untimed steps, then `steps` timed.
static int bench_resident(int nx, int ny, int nz, int ppc, int warm,
                          int steps) {
  Box b;
  make_box(b, nx, ny, nz, 0);
  Arrays A;
  make_arrays(A, b, 7);
  memset(A.f.data(), 0, sizeof(field_t) * A.f.size());
  if (vpic_hip_install_field_kernels(&A.fa)) {
    printf("FAIL install: %s\n", vpic_hip_last_error());
    return 1;
  }
  Species sp[2];
  make_species(sp[0], b, -1.0f, 1.0f, ppc, 0.1f, 71);
  make_species(sp[1], b, 1.0f, 25.0f, ppc, 0.02f, 72);
  // weights 1/ppc: density 1, plasma frequency 1, so dt = 0.5 resolves it
  // (the checks' weights near 1 would give omega_pe dt = 5, unstable)
  for (int s = 0; s < 2; s++)
    for (int k = 0; k < sp[s].sp.np; k++) sp[s].p[k].w = 1.0f / ppc;
  field_advance_kernels_t *k = A.fa.kernel;
  vpic_hip_dropin_resident(1);
  load_interpolator_array(&A.ia, &A.fa);
  auto step = [&](int n) {
    for (int s = 0; s < 2; s++)
      if (n % sp[s].sp.sort_interval == 0) sort_p(&sp[s].sp);
    clear_accumulator_array(&A.aa);
    for (int s = 0; s < 2; s++) advance_p(&sp[s].sp, &A.aa, &A.ia);
    reduce_accumulator_array(&A.aa);
    k->clear_jf(&A.fa);
    unload_accumulator_array(&A.fa, &A.aa);
    k->synchronize_jf(&A.fa);
    k->advance_b(&A.fa, 0.5f);
    k->advance_e(&A.fa, 1.0f);
    k->advance_b(&A.fa, 0.5f);
    load_interpolator_array(&A.ia, &A.fa);
  };
  for (int n = 0; n < warm; n++) step(n);
  vpic_hip_device_synchronize();
  const auto t0 = std::chrono::steady_clock::now();
  for (int n = warm; n < warm + steps; n++) step(n);
  vpic_hip_device_synchronize();
  const double el = std::chrono::duration<double>(
                        std::chrono::steady_clock::now() - t0).count();
  const double np = (double)sp[0].sp.np + sp[1].sp.np;
  int lost = 0;
  for (int s = 0; s < 2; s++) lost += sp[s].sp.nm;
  vpic_hip_dropin_resident(0);  // downloads the device state
  double umax = 0, emax = 0;
  for (int k = 0; k < sp[0].sp.np; k++) {
    const particle_t &q = sp[0].p[k];
    umax = fmax(umax, sqrt((double)q.ux * q.ux + (double)q.uy * q.uy +
                           (double)q.uz * q.uz));
  }
  for (size_t v = 0; v < A.f.size(); v++)
    emax = fmax(emax, fabs((double)A.f[v].ex));
  vpic_hip_dropin_release_all();
  printf("dropin resident: %dx%dx%d cells, %.0f particles, %d steps, "
         "%.4e particle-steps/s, %.3f ms/step, movers %d, electron |u| max "
         "%.3g, |ex| max %.3g\n", nx, ny, nz, np, steps, np * steps / el,
         el * 1e3 / steps, lost, umax, emax);
  return lost == 0 ? 0 : 1;
}

int main(int argc, char **argv) {
  if (argc > 1 && strcmp(argv[1], "--bench") == 0) {
    int a[6] = {128, 64, 64, 100, 5, 40};  // nx ny nz ppc warm steps
    for (int k = 0; k < 6 && k + 2 < argc; k++) a[k] = atoi(argv[k + 2]);
    if (vpic_hip_init(0) != 0) {
      printf("FAIL init: %s\n", vpic_hip_last_error());
      return 1;
    }
    return bench_resident(a[0], a[1], a[2], a[3], a[4], a[5]);
  }
  check_layout();
  if (argc > 1 && strcmp(argv[1], "--layout") == 0) return g_fail;
  if (vpic_hip_init(0) != 0) {
    printf("FAIL init: %s\n", vpic_hip_last_error());
    return 1;
  }
  check_species_kernels();
  check_field_kernels(0, "field kernels via fa->kernel (periodic)");
  check_field_kernels(absorb_particles, "field kernels via fa->kernel (pec x)");
  check_absorb();
  check_boundary();
  check_resident();
  check_errors();
  vpic_hip_dropin_release_all();
  printf("%d failed\n", g_fail);
  return g_fail;
}
