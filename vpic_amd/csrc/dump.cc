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

// The extents and spacing a dump writes in its header: the grid's for the
// whole-array dumps, divided / multiplied by the strides for field_dump
// and hydro_dump (dump.cc:590-596).
struct OutGrid {
  int nx, ny, nz;
  float dx, dy, dz;
};

OutGrid whole(const vpic_grid_t *g) {
  return OutGrid{g->nx, g->ny, g->nz, g->dx, g->dy, g->dz};
}

// WRITE_HEADER_V0 (dumpmacros.h:7-43): binary compatibility block, dump
// type and version, grid extents and spacing of the written array, the
// domain's low corner, cvac, eps0, damp (written as 0), rank / nproc, then
// the species id and q/m (-1 and 0 when not a species dump).
void header_v0(Out &o, int dump_type, int sp_id, float q_m,
               const vpic_grid_t *g, const vpic_dump_info_t *in,
               const OutGrid &og, int64_t step) {
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
  o.put<int>((int)step);
  o.put<int>(og.nx);
  o.put<int>(og.ny);
  o.put<int>(og.nz);
  o.put<float>(g->dt);
  o.put<float>(og.dx);
  o.put<float>(og.dy);
  o.put<float>(og.dz);
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

void header_v0(Out &o, int dump_type, int sp_id, float q_m,
               const vpic_grid_t *g, const vpic_dump_info_t *in) {
  header_v0(o, dump_type, sp_id, q_m, g, in, whole(g), in->step);
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

// ---------------------------------------------------------------------------
// The strided dumps and their global header (dump.cc:337-880)
// ---------------------------------------------------------------------------
#include <sys/stat.h>

#include <string>
#include <vector>

namespace {

struct VarInfo {
  const char *name, *degree, *elements, *type;
  int size;
};

// The variable groups and their first word (vpic.h:53-79, dump.cc:343-371):
// the names the global header lists, word counts as the reference counts
// them (a material id, 2 bytes, is counted as one word).
const VarInfo kFieldInfo[12] = {
    {"Electric Field", "VECTOR", "3", "FLOATING_POINT", 4},
    {"Electric Field Divergence Error", "SCALAR", "1", "FLOATING_POINT", 4},
    {"Magnetic Field", "VECTOR", "3", "FLOATING_POINT", 4},
    {"Magnetic Field Divergence Error", "SCALAR", "1", "FLOATING_POINT", 4},
    {"TCA Field", "VECTOR", "3", "FLOATING_POINT", 4},
    {"Bound Charge Density", "SCALAR", "1", "FLOATING_POINT", 4},
    {"Free Current Field", "VECTOR", "3", "FLOATING_POINT", 4},
    {"Charge Density", "SCALAR", "1", "FLOATING_POINT", 4},
    {"Edge Material", "VECTOR", "3", "INTEGER", 2},
    {"Node Material", "SCALAR", "1", "INTEGER", 2},
    {"Face Material", "VECTOR", "3", "INTEGER", 2},
    {"Cell Material", "SCALAR", "1", "INTEGER", 2}};
const int kFieldFirst[12] = {0, 3, 4, 7, 8, 11, 12, 15, 16, 19, 20, 23};
const VarInfo kHydroInfo[5] = {
    {"Current Density", "VECTOR", "3", "FLOATING_POINT", 4},
    {"Charge Density", "SCALAR", "1", "FLOATING_POINT", 4},
    {"Momentum Density", "VECTOR", "3", "FLOATING_POINT", 4},
    {"Kinetic Energy Density", "SCALAR", "1", "FLOATING_POINT", 4},
    {"Stress Tensor", "TENSOR", "6", "FLOATING_POINT", 4}};
const int kHydroFirst[5] = {0, 3, 4, 7, 8};
constexpr int kFieldWords = 24, kHydroWords = 14;

bool bit_set(uint32_t m, int i) { return ((m >> i) & 1u) != 0; }

void hashed(FILE *f, bool &ok, const char *comment) {
  const char *bar =
      "################################################################"
      "################\n";
  if (fprintf(f, "%s# %s\n%s", bar, comment, bar) < 0) ok = false;
}

// One dump's variables: the selected groups' lines of the global header.
bool list_vars(FILE *f, const char *key, uint32_t mask, const VarInfo *info,
               const int *first, int ngroups) {
  int n = 0;
  for (int i = 0; i < ngroups; i++) n += bit_set(mask, first[i]);
  bool ok = fprintf(f, "%s %d\n", key, n) >= 0;
  for (int i = 0; i < ngroups; i++)
    if (bit_set(mask, first[i]) &&
        fprintf(f, "\"%s\" %s %s %s %d\n", info[i].name, info[i].degree,
                info[i].elements, info[i].type, info[i].size) < 0)
      ok = false;
  return ok;
}

// The output voxel index along one axis (dump.cc:643-645): the low ghost,
// every stride-th cell's last voxel, the high ghost; unit strides take the
// voxels as they are.
inline int out_index(int k, int nout, int n, int stride, bool unit) {
  if (unit) return k;
  return k == 0 ? 0 : k == nout + 1 ? n + 1 : k * stride - 1;
}

struct Strided {
  const char *what;
  int dump_type;
  const unsigned char *base;  // the array's records
  size_t rec;                 // bytes per record
  int nwords;                 // variables the band format may select
  bool ghost_dims;            // band_interleave dims include the ghosts
};

int strided_dump(const Strided &d, const vpic_grid_t *g,
                 const vpic_dump_info_t *info, int sp_id, float q_m,
                 const vpic_dump_params_t *p, int64_t step) {
  if (!args_ok(d.what, "", g, info)) return -1;
  if (!d.base || !p || !p->base_dir || !p->base_file_name)
    return vpic_set_error(-1, "%s: NULL argument", d.what);
  const int n[3] = {g->nx, g->ny, g->nz};
  const char axis[3] = {'x', 'y', 'z'};
  for (int a = 0; a < 3; a++)
    if (p->stride[a] <= 0 || n[a] % p->stride[a] != 0)
      return vpic_set_error(-1, "%s: %c stride must be an integer factor of "
                            "n%c", d.what, axis[a], axis[a]);
  if (p->format != VPIC_DUMP_BAND && p->format != VPIC_DUMP_BAND_INTERLEAVE)
    return vpic_set_error(-1, "%s: unknown format %d", d.what, p->format);
  const long dstep = (long)(step < 0 ? info->step : step);
  // dump.cc:540-562: the step's directory, then the rank's file
  const std::string dir =
      std::string(p->base_dir) + "/T." + std::to_string(dstep);
  mkdir(dir.c_str(), S_IRWXU);  // an existing directory is fine
  const std::string fname = dir + "/" + p->base_file_name + "." +
                            std::to_string(dstep) + "." +
                            std::to_string(info->rank);
  const int sx = p->stride[0], sy = p->stride[1], sz = p->stride[2];
  const OutGrid og{g->nx / sx, g->ny / sy, g->nz / sz, g->dx * (float)sx,
                   g->dy * (float)sy, g->dz * (float)sz};
  const bool unit = sx == 1 && sy == 1 && sz == 1;
  std::vector<int> vars;
  if (p->format == VPIC_DUMP_BAND) {
    int total = 0;
    for (int i = 0; i < 32; i++) total += bit_set(p->output_vars, i);
    for (int i = 0; i < d.nwords; i++)
      if (bit_set(p->output_vars, i)) vars.push_back(i);
    const int numvars = total < d.nwords ? total : d.nwords;
    if (numvars != (int)vars.size())
      return vpic_set_error(-1, "%s: output_vars 0x%x selects bits past "
                            "word %d", d.what, p->output_vars, d.nwords - 1);
  }
  Out o;
  if (int rc = open_out(d.what, fname.c_str(), "wb", o)) return rc;
  header_v0(o, d.dump_type, sp_id, q_m, g, info, og, dstep);
  const size_t sy_v = (size_t)g->nx + 2, sz_v = sy_v * ((size_t)g->ny + 2);
  const size_t nbytes = d.rec * (size_t)voxels(g);
  std::vector<unsigned char> row;
  if (p->format == VPIC_DUMP_BAND) {
    // dump.cc:604-662: one block per selected word
    const int dim[3] = {og.nx + 2, og.ny + 2, og.nz + 2};
    array_header(o, (int)d.rec, 3, dim);
    row.resize(4 * (size_t)dim[0]);
    for (int v : vars)
      for (int k = 0; k < dim[2]; k++) {
        const int kk = out_index(k, og.nz, g->nz, sz, unit);
        for (int j = 0; j < dim[1]; j++) {
          const int jj = out_index(j, og.ny, g->ny, sy, unit);
          for (int i = 0; i < dim[0]; i++) {
            const int ii = out_index(i, og.nx, g->nx, sx, unit);
            const size_t at =
                d.rec * (ii + sy_v * jj + sz_v * kk) + 4 * (size_t)v;
            uint32_t w = 0;  // past the array's end: see the header
            if (at + 4 <= nbytes) memcpy(&w, d.base + at, 4);
            memcpy(&row[4 * (size_t)i], &w, 4);
          }
          o.raw(row.data(), row.size());
        }
      }
  } else if (d.ghost_dims) {
    // dump.cc:668-702: whole records, ghosts included
    const int dim[3] = {og.nx + 2, og.ny + 2, og.nz + 2};
    array_header(o, (int)d.rec, 3, dim);
    if (unit) {
      o.raw(d.base, nbytes);
    } else {
      for (int k = 0; k < dim[2]; k++) {
        const int kk = out_index(k, og.nz, g->nz, sz, false);
        for (int j = 0; j < dim[1]; j++) {
          const int jj = out_index(j, og.ny, g->ny, sy, false);
          for (int i = 0; i < dim[0]; i++) {
            const int ii = out_index(i, og.nx, g->nx, sx, false);
            o.raw(d.base + d.rec * (ii + sy_v * jj + sz_v * kk), d.rec);
          }
        }
      }
    }
  } else {
    // dump.cc:851-876: nx/sx * ny/sy * nz/sz records; unit strides write
    // the array's first records, other strides start at the low ghost
    const int dim[3] = {og.nx, og.ny, og.nz};
    array_header(o, (int)d.rec, 3, dim);
    if (unit) {
      o.raw(d.base, d.rec * (size_t)dim[0] * dim[1] * dim[2]);
    } else {
      for (int k = 0; k < dim[2]; k++) {
        const int kk = out_index(k, og.nz, g->nz, sz, false);
        for (int j = 0; j < dim[1]; j++) {
          const int jj = out_index(j, og.ny, g->ny, sy, false);
          for (int i = 0; i < dim[0]; i++) {
            const int ii = out_index(i, og.nx, g->nx, sx, false);
            o.raw(d.base + d.rec * (ii + sy_v * jj + sz_v * kk), d.rec);
          }
        }
      }
    }
  }
  return close_out(d.what, fname.c_str(), o);
}

}  // namespace

extern "C" int vpic_hip_global_header(const char *base, const vpic_grid_t *g,
                                      const vpic_global_info_t *gi,
                                      const vpic_dump_params_t *params,
                                      int32_t nparams) {
  const char *what = "global_header";
  if (!base || !g || !gi || !params)
    return vpic_set_error(-1, "%s: NULL argument", what);
  if (nparams < 2)
    return vpic_set_error(-1, "%s: needs the field parameters and at least "
                          "one species' (dumpParams.size() >= 2)", what);
  for (int i = 0; i < nparams; i++)
    if (!params[i].base_dir || !params[i].base_file_name)
      return vpic_set_error(-1, "%s: NULL directory or file name", what);
  if (gi->rank) return 0;  // rank 0 writes the header
  const std::string fname = std::string(base) + ".vpc";
  Out o;
  if (int rc = open_out(what, fname.c_str(), "w", o)) return rc;
  FILE *f = o.f;
  bool &ok = o.ok;
  auto pr = [&](int rc) {
    if (rc < 0) ok = false;
  };
  // dump.cc:419-460
  hashed(f, ok, "Header version information");
  pr(fprintf(f, "VPIC_HEADER_VERSION 1.0.0\n\n"));
  hashed(f, ok, "Header size for data file headers in bytes");
  pr(fprintf(f, "DATA_HEADER_SIZE 123\n\n"));
  hashed(f, ok, "Time step increment");
  pr(fprintf(f, "GRID_DELTA_T %f\n\n", g->dt));
  hashed(f, ok, "GRID_CVAC");
  pr(fprintf(f, "GRID_CVAC %f\n\n", g->cvac));
  hashed(f, ok, "GRID_EPS0");
  pr(fprintf(f, "GRID_EPS0 %f\n\n", g->eps0));
  hashed(f, ok, "Grid extents in the x-dimension");
  pr(fprintf(f, "GRID_EXTENTS_X %f %f\n\n", gi->x0, gi->x1));
  hashed(f, ok, "Grid extents in the y-dimension");
  pr(fprintf(f, "GRID_EXTENTS_Y %f %f\n\n", gi->y0, gi->y1));
  hashed(f, ok, "Grid extents in the z-dimension");
  pr(fprintf(f, "GRID_EXTENTS_Z %f %f\n\n", gi->z0, gi->z1));
  hashed(f, ok, "Spatial step increment in x-dimension");
  pr(fprintf(f, "GRID_DELTA_X %f\n\n", g->dx));
  hashed(f, ok, "Spatial step increment in y-dimension");
  pr(fprintf(f, "GRID_DELTA_Y %f\n\n", g->dy));
  hashed(f, ok, "Spatial step increment in z-dimension");
  pr(fprintf(f, "GRID_DELTA_Z %f\n\n", g->dz));
  hashed(f, ok, "Domain partitions in x-dimension");
  pr(fprintf(f, "GRID_TOPOLOGY_X %d\n\n", gi->px));
  hashed(f, ok, "Domain partitions in y-dimension");
  pr(fprintf(f, "GRID_TOPOLOGY_Y %d\n\n", gi->py));
  hashed(f, ok, "Domain partitions in z-dimension");
  pr(fprintf(f, "GRID_TOPOLOGY_Z %d\n\n", gi->pz));
  // dump.cc:462-489: the field dump's variables
  hashed(f, ok, "Field data information");
  pr(fprintf(f, "FIELD_DATA_DIRECTORY %s\n", params[0].base_dir));
  pr(fprintf(f, "FIELD_DATA_BASE_FILENAME %s\n", params[0].base_file_name));
  if (!list_vars(f, "FIELD_DATA_VARIABLES", params[0].output_vars,
                 kFieldInfo, kFieldFirst, 12))
    ok = false;
  pr(fprintf(f, "\n"));
  // dump.cc:491-528: each species' hydro variables
  hashed(f, ok, "Number of species with output data");
  pr(fprintf(f, "NUM_OUTPUT_SPECIES %d\n\n", nparams - 1));
  for (int i = 1; i < nparams; i++) {
    char comment[128];
    snprintf(comment, sizeof comment, "Species(%d) data information", i);
    hashed(f, ok, comment);
    pr(fprintf(f, "SPECIES_DATA_DIRECTORY %s\n", params[i].base_dir));
    pr(fprintf(f, "SPECIES_DATA_BASE_FILENAME %s\n",
               params[i].base_file_name));
    if (!list_vars(f, "HYDRO_DATA_VARIABLES", params[i].output_vars,
                   kHydroInfo, kHydroFirst, 5))
      ok = false;
    if (i < nparams - 1) pr(fprintf(f, "\n"));
  }
  return close_out(what, fname.c_str(), o);
}

extern "C" int vpic_hip_field_dump(const vpic_field_t *f,
                                   const vpic_grid_t *g,
                                   const vpic_dump_info_t *info,
                                   const vpic_dump_params_t *p,
                                   int64_t step) {
  const Strided d{"field_dump", FIELD_DUMP,
                  reinterpret_cast<const unsigned char *>(f),
                  sizeof(vpic_field_t), kFieldWords, true};
  return strided_dump(d, g, info, -1, 0.0f, p, step);
}

extern "C" int vpic_hip_hydro_dump(const vpic_hydro_t *h,
                                   const vpic_grid_t *g,
                                   const vpic_dump_info_t *info,
                                   int32_t sp_id, float q_m,
                                   const vpic_dump_params_t *p,
                                   int64_t step) {
  const Strided d{"hydro_dump", HYDRO_DUMP,
                  reinterpret_cast<const unsigned char *>(h),
                  sizeof(vpic_hydro_t), kHydroWords, false};
  return strided_dump(d, g, info, sp_id, q_m, p, step);
}
