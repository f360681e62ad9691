// dump.cc -- the reference's output formats, written from host arrays:
// the energies text trace (src/vpic/dump.cc:38-77) and the binary dumps
// with the version-0 header (src/vpic/dumpmacros.h:7-53; dump.cc:138-335:
// grid, fields, hydro, particles).  Host code only; the caller downloads
// what it dumps (and time-centres particles first, as dump_particles
// does with center_p).  Files are written in the machine's byte order, as
// the reference's FileIO does (little endian on this platform).
#include <cerrno>
#include <cstdint>
#include <cstdio>
#include <cstring>

#include "../../include/vpic_hip.h"

int vpic_set_error(int code, const char *fmt, ...);

namespace {

// dump.cc:129-136
enum { GRID_DUMP = 0, FIELD_DUMP = 1, HYDRO_DUMP = 2, PARTICLE_DUMP = 3 };

struct Out {
  FILE *f = nullptr;
  bool ok = true;
  template <class T>
  void put(T v) {
    if (ok && fwrite(&v, sizeof v, 1, f) != 1) ok = false;
  }
  void raw(const void *p, size_t bytes) {
    if (ok && bytes && fwrite(p, 1, bytes, f) != bytes) ok = false;
  }
};

// WRITE_HEADER_V0 (dumpmacros.h:7-43): binary compatibility block, dump
// type and version, grid extents and spacing of the written array, the
// domain's low corner, cvac, eps0, damp (written as 0), rank / nproc, then
// the species id and q/m (-1 and 0 when not a species dump).
void header_v0(Out &o, int dump_type, int sp_id, float q_m,
               const vpic_grid_t *g, const vpic_dump_info_t *in) {
  o.put<char>(8);  // CHAR_BIT
  o.put<char>((char)sizeof(short));
  o.put<char>((char)sizeof(int));
  o.put<char>((char)sizeof(float));
  o.put<char>((char)sizeof(double));
  o.put<short>((short)0xcafe);
  o.put<int>((int)0xdeadbeef);
  o.put<float>(1.0f);
  o.put<double>(1.0);
  o.put<int>(0);  // header version
  o.put<int>(dump_type);
  o.put<int>((int)in->step);
  o.put<int>(g->nx);
  o.put<int>(g->ny);
  o.put<int>(g->nz);
  o.put<float>(g->dt);
  o.put<float>(g->dx);
  o.put<float>(g->dy);
  o.put<float>(g->dz);
  o.put<float>(in->x0);
  o.put<float>(in->y0);
  o.put<float>(in->z0);
  o.put<float>(g->cvac);
  o.put<float>(g->eps0);
  o.put<float>(0.0f);  // damp
  o.put<int>(in->rank);
  o.put<int>(in->nproc);
  o.put<int>(sp_id);
  o.put<float>(q_m);
}

// WRITE_ARRAY_HEADER (dumpmacros.h:47-51): element size, rank, extents.
void array_header(Out &o, int elem, int ndim, const int *dim) {
  o.put<int>(elem);
  o.put<int>(ndim);
  for (int k = 0; k < ndim; k++) o.put<int>(dim[k]);
}

bool args_ok(const char *what, const char *fname, const vpic_grid_t *g,
             const vpic_dump_info_t *in) {
  if (!fname || !g || !in) {
    vpic_set_error(-1, "%s: NULL argument", what);
    return false;
  }
  if (g->nx <= 0 || g->ny <= 0 || g->nz <= 0) {
    vpic_set_error(-1, "%s: empty grid", what);
    return false;
  }
  return true;
}

int open_out(const char *what, const char *fname, const char *mode, Out &o) {
  o.f = fopen(fname, mode);
  if (!o.f)
    return vpic_set_error(-1, "%s: Could not open \"%s\" (%s)", what, fname,
                          strerror(errno));
  return 0;
}

int close_out(const char *what, const char *fname, Out &o) {
  const bool wrote = o.ok;
  const bool closed = fclose(o.f) == 0;
  o.f = nullptr;
  if (!wrote || !closed)
    return vpic_set_error(-1, "%s: writing \"%s\" failed", what, fname);
  return 0;
}

int64_t voxels(const vpic_grid_t *g) {
  return (int64_t)(g->nx + 2) * (g->ny + 2) * (g->nz + 2);
}

}  // namespace

extern "C" int vpic_hip_write_energies(const char *fname, int append,
                                       const vpic_dump_info_t *info, float dt,
                                       const double en_f[6],
                                       const double *en_p,
                                       const char *const *sp_names,
                                       int32_t nsp) {
  const char *what = "dump_energies";
  if (!fname || !info || !en_f || nsp < 0 || (nsp && !en_p))
    return vpic_set_error(-1, "%s: Invalid arguments", what);
  Out o;
  if (int rc = open_out(what, fname, append ? "a" : "w", o)) return rc;
  // dump.cc:51-57: the layout comment opens a new file
  if (!append) {
    if (fprintf(o.f, "%% Layout\n%% step ex ey ez bx by bz") < 0) o.ok = false;
    for (int k = 0; k < nsp; k++)
      if (fprintf(o.f, " \"%s\"",
                  sp_names && sp_names[k] ? sp_names[k] : "") < 0)
        o.ok = false;
    if (fprintf(o.f, "\n%% timestep = %e\n", dt) < 0) o.ok = false;
  }
  // dump.cc:58-76: step, the six field energies, one energy per species
  if (fprintf(o.f, "%li ", (long)info->step) < 0 ||
      fprintf(o.f, "%e %e %e %e %e %e", en_f[0], en_f[1], en_f[2], en_f[3],
              en_f[4], en_f[5]) < 0)
    o.ok = false;
  for (int k = 0; k < nsp; k++)
    if (fprintf(o.f, " %e", en_p[k]) < 0) o.ok = false;
  if (fprintf(o.f, "\n") < 0) o.ok = false;
  return close_out(what, fname, o);
}

extern "C" int vpic_hip_write_fields(const char *fname, const vpic_field_t *f,
                                     const vpic_grid_t *g,
                                     const vpic_dump_info_t *info) {
  const char *what = "dump_fields";
  if (!args_ok(what, fname, g, info)) return -1;
  if (!f) return vpic_set_error(-1, "%s: NULL field array", what);
  Out o;
  if (int rc = open_out(what, fname, "wb", o)) return rc;
  // dump.cc:181-219
  header_v0(o, FIELD_DUMP, -1, 0.0f, g, info);
  const int dim[3] = {g->nx + 2, g->ny + 2, g->nz + 2};
  array_header(o, (int)sizeof(vpic_field_t), 3, dim);
  o.raw(f, sizeof(vpic_field_t) * (size_t)voxels(g));
  return close_out(what, fname, o);
}

extern "C" int vpic_hip_write_hydro(const char *fname, const vpic_hydro_t *h,
                                    const vpic_grid_t *g,
                                    const vpic_dump_info_t *info,
                                    int32_t sp_id, float q_m) {
  const char *what = "dump_hydro";
  if (!args_ok(what, fname, g, info)) return -1;
  if (!h) return vpic_set_error(-1, "%s: NULL hydro array", what);
  Out o;
  if (int rc = open_out(what, fname, "wb", o)) return rc;
  // dump.cc:221-270
  header_v0(o, HYDRO_DUMP, sp_id, q_m, g, info);
  const int dim[3] = {g->nx + 2, g->ny + 2, g->nz + 2};
  array_header(o, (int)sizeof(vpic_hydro_t), 3, dim);
  o.raw(h, sizeof(vpic_hydro_t) * (size_t)voxels(g));
  return close_out(what, fname, o);
}

extern "C" int vpic_hip_write_particles(const char *fname,
                                        const vpic_particle_t *p, int32_t n,
                                        int32_t np_total, const vpic_grid_t *g,
                                        const vpic_dump_info_t *info,
                                        int32_t sp_id, float q_m) {
  const char *what = "dump_particles";
  if (!args_ok(what, fname, g, info)) return -1;
  if (n < 0 || np_total < n || (n && !p))
    return vpic_set_error(-1, "%s: Invalid particle counts", what);
  Out o;
  if (int rc = open_out(what, fname, "wb", o)) return rc;
  // dump.cc:273-335: header, then np particles (time-centred by the caller)
  header_v0(o, PARTICLE_DUMP, sp_id, q_m, g, info);
  const int dim[1] = {np_total};
  array_header(o, (int)sizeof(vpic_particle_t), 1, dim);
  o.raw(p, sizeof(vpic_particle_t) * (size_t)n);
  return close_out(what, fname, o);
}

extern "C" int vpic_hip_write_grid(const char *fname, const vpic_grid_t *g,
                                   const vpic_dump_info_t *info,
                                   const int32_t bc[27], const int64_t *range,
                                   const int64_t *neighbor) {
  const char *what = "dump_grid";
  if (!args_ok(what, fname, g, info)) return -1;
  if (!bc || !range || !neighbor || info->nproc < 1)
    return vpic_set_error(-1, "%s: NULL argument", what);
  Out o;
  if (int rc = open_out(what, fname, "wb", o)) return rc;
  // dump.cc:138-179: bc[3][3][3], range[nproc+1], neighbor[6][nx+2][..]
  header_v0(o, GRID_DUMP, -1, 0.0f, g, info);
  const int dbc[3] = {3, 3, 3};
  array_header(o, (int)sizeof(int32_t), 3, dbc);
  o.raw(bc, sizeof(int32_t) * 27);
  const int dr[1] = {info->nproc + 1};
  array_header(o, (int)sizeof(int64_t), 1, dr);
  o.raw(range, sizeof(int64_t) * (size_t)dr[0]);
  const int dn[4] = {6, g->nx + 2, g->ny + 2, g->nz + 2};
  array_header(o, (int)sizeof(int64_t), 4, dn);
  o.raw(neighbor, sizeof(int64_t) * 6 * (size_t)voxels(g));
  return close_out(what, fname, o);
}
