// boundary.hip -- boundary_p's mover loop on gfx950 for a domain with no
// face shared with another rank.
//
// Reference: src/boundary/boundary_p.cc:254-375 (the movers of one
// species, last first: absorb -> accumulate_rhob, custom handler, or drop;
// then the last particle fills the mover's slot) and
// src/species_advance/standard/rho_p.cc:126-173 (accumulate_rhob).
//
// The reference removes the movers' particles one at a time, last mover
// first, each time moving the array's last particle into the hole.  With
// the nm movers r_0 < r_1 < ... < r_{nm-1} (the order advance_p leaves
// them in) and N particles, the removal of r_j takes the particle then at
// position N - nm + j.  Positions [N - nm, N) form the tail; a hole below
// the tail (r_j < N - nm) receives the particle at tail position
// N - nm + j, unless that position is itself a mover r_w (w > j, already
// removed), in which case it had received the particle at N - nm + w, and
// so on.  These chains only move up the tail and end at a tail particle
// that is not a mover; no two chains share a position.  Here every mover
// writes its chain link, the links are resolved by pointer jumping
// (log2 nm passes), and the holes are filled in one parallel pass: the
// same particles land in the same slots as the reference's loop.
#include "common.h"

namespace {

inline int blocks_for(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

__global__ void k_bp_fill_i32(int *a, int n, int v) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += gridDim.x * blockDim.x)
    a[k] = v;
}

// accumulate_rhob (rho_p.cc:126-173): FMA/FNMS of the file's macros are
// z + x*y and z - x*y; weights doubled on the faces of the local domain.
__device__ void rhob_deposit(vpic_field_t *f, const vpic_particle_t &p,
                             float qsp, float r8V, int voxel, int nx, int ny,
                             int nz, int sy, int sz) {
  float w0 = p.dx, w1 = p.dy, w2, w3, w4, w5, w6, w7, dz = p.dz;
  w7 = (qsp * r8V) * p.w;
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
  int x = voxel;
  const int z = x / sz;
  if (z == 1) w0 += w0, w1 += w1, w2 += w2, w3 += w3;
  if (z == nz) w4 += w4, w5 += w5, w6 += w6, w7 += w7;
  x -= sz * z;
  const int y = x / sy;
  if (y == 1) w0 += w0, w1 += w1, w4 += w4, w5 += w5;
  if (y == ny) w2 += w2, w3 += w3, w6 += w6, w7 += w7;
  x -= sy * y;
  if (x == 1) w0 += w0, w2 += w2, w4 += w4, w6 += w6;
  if (x == nx) w1 += w1, w3 += w3, w5 += w5, w7 += w7;
  atomic_add_f32(&f[voxel].rhob, w0);
  atomic_add_f32(&f[voxel + 1].rhob, w1);
  atomic_add_f32(&f[voxel + sy].rhob, w2);
  atomic_add_f32(&f[voxel + sy + 1].rhob, w3);
  atomic_add_f32(&f[voxel + sz].rhob, w4);
  atomic_add_f32(&f[voxel + sz + 1].rhob, w5);
  atomic_add_f32(&f[voxel + sz + sy].rhob, w6);
  atomic_add_f32(&f[voxel + sz + sy + 1].rhob, w7);
}

// cnt[2] counts movers out of order or out of range, or whose particle
// index is not 8*voxel + face with the voxel's 8 nodes inside the grid; the
// call then fails before anything is written.
__global__ void k_bp_check(const vpic_particle_t *__restrict__ p, int np,
                           const vpic_particle_mover_t *__restrict__ pm,
                           int nm, int nv, int sy, int sz,
                           int *__restrict__ cnt) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nm) return;
  const int i = pm[m].i;
  bool ok = (unsigned)i < (unsigned)np && (m == 0 || pm[m - 1].i < i);
  if (ok) {
    const int pi = p[i].i;
    ok = pi >= 0 && (pi & 7) <= 5 && (pi >> 3) + sz + sy + 1 < nv;
  }
  if (!ok) atomicAdd(cnt + 2, 1);
}

// One lane per (checked) mover m: what its boundary is (absorbed charge to
// rhob, the hit record), and its chain link when its particle is in the
// tail.
__global__ void k_bp_classify(const vpic_particle_t *__restrict__ p, int np,
                              const vpic_particle_mover_t *__restrict__ pm,
                              int nm, vpic_field_t *__restrict__ f, float q,
                              float r8V, GridDev G,
                              vpic_boundary_hit_t *__restrict__ hits,
                              int *__restrict__ link, int *__restrict__ cnt) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nm) return;
  const int i = pm[m].i;
  vpic_particle_t pp = p[i];
  const int face = pp.i & 7;
  const int voxel = pp.i >> 3;
  const int tail0 = np - nm;
  if (i >= tail0) link[i - tail0] = m;
  pp.i = voxel;
  const int64_t nn = grid_neighbor(G, voxel, face);
  int code = VPIC_BOUNDARY_UNKNOWN;
  if (nn == VPIC_BC_ABSORB_PARTICLES) {
    code = VPIC_BC_ABSORB_PARTICLES;
  } else if (nn <= -3) {
    code = nn < -2147483647 ? -2147483647 - 1 : (int)nn;
    atomicAdd(cnt + 0, 1);
  } else {
    atomicAdd(cnt + 1, 1);
  }
  if (hits) {
    vpic_boundary_hit_t h;
    h.p = pp;
    h.m = pm[m];
    h.face = face;
    h.code = code;
    h.pad_[0] = h.pad_[1] = 0;
    hits[m] = h;
  }
  if (code == VPIC_BC_ABSORB_PARTICLES)
    rhob_deposit(f, pp, q, r8V, voxel, G.nx, G.ny, G.nz, G.sy, G.sz);
}

// Chain start: tail rank u leads to rank link[u] when a mover sits there,
// else it is the end of its chain.
__global__ void k_bp_jump_init(const int *__restrict__ link,
                               int *__restrict__ d, int nm) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u < nm) d[u] = link[u] >= 0 ? link[u] : u;
}

__global__ void k_bp_jump(const int *__restrict__ d, int *__restrict__ e,
                          int nm) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u < nm) e[u] = d[d[u]];
}

// Fill each hole below the tail with the particle at the end of its chain.
__global__ void k_bp_fill(vpic_particle_t *__restrict__ p, int np,
                          const vpic_particle_mover_t *__restrict__ pm, int nm,
                          const int *__restrict__ d) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= nm) return;
  const int tail0 = np - nm;
  const int i = pm[m].i;
  if (i >= tail0) return;
  const float4 *src = reinterpret_cast<const float4 *>(p + tail0 + d[m]);
  float4 *dst = reinterpret_cast<float4 *>(p + i);
  dst[0] = src[0];
  dst[1] = src[1];
}

struct BoundaryScratch {
  DeviceScratch cnt, link, d0, d1;
};

BoundaryScratch &boundary_scratch() {
  static thread_local BoundaryScratch s[16];
  int dev = 0;
  (void)hipGetDevice(&dev);
  return s[dev & 15];
}

}  // namespace

namespace {
// accumulate_rhob (rho_p.cc:126-173) of particles p[0..n), one lane each;
// a particle whose 8 nodes are not all on the grid is counted in bad[0].
__global__ void k_accumulate_rhob(vpic_field_t *__restrict__ f,
                                  const vpic_particle_t *__restrict__ p, int n,
                                  float qsp, float r8V, int nx, int ny, int nz,
                                  int sy, int sz, int nv, int *bad) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += gridDim.x * blockDim.x) {
    const vpic_particle_t pp = p[k];
    if (pp.i < 0 || pp.i + sz + sy + 1 >= nv) {
      atomicAdd(bad, 1);
      continue;
    }
    rhob_deposit(f, pp, qsp, r8V, pp.i, nx, ny, nz, sy, sz);
  }
}
}  // namespace

int launch_accumulate_rhob(vpic_field_t *f, const vpic_particle_t *p, int n,
                           float qsp, const vpic_grid_t *g, int *dev_bad,
                           hipStream_t s) {
  if (n <= 0) return VPIC_OK;
  int blk = (n + 255) / 256;
  if (blk > 4096) blk = 4096;
  hipLaunchKernelGGL(k_accumulate_rhob, dim3(blk), dim3(256), 0, s, f, p, n,
                     qsp, g->r8V, g->nx, g->ny, g->nz, g->sy, g->sz, g->nv,
                     dev_bad);
  VPIC_LAUNCHED("k_accumulate_rhob");
  return VPIC_OK;
}

int launch_boundary_p(vpic_particle_t *p, int *np_host,
                      const vpic_particle_mover_t *pm, int nm,
                      vpic_field_t *f, float q, const vpic_grid_t *g,
                      vpic_boundary_hit_t *hits, int *counts_host,
                      hipStream_t s) {
  counts_host[0] = counts_host[1] = 0;
  const int np = *np_host;
  if (nm == 0) return VPIC_OK;
  BoundaryScratch &S = boundary_scratch();
  const size_t ib = sizeof(int) * (size_t)nm;
  int rc;
  if ((rc = scratch_reserve(S.cnt, 64)) || (rc = scratch_reserve(S.link, ib)) ||
      (rc = scratch_reserve(S.d0, ib)) || (rc = scratch_reserve(S.d1, ib)))
    return rc;
  int *cnt = (int *)S.cnt.ptr, *link = (int *)S.link.ptr;
  int *d = (int *)S.d0.ptr, *e = (int *)S.d1.ptr;
  const GridDev G = make_grid_dev(g);
  const int nb = blocks_for(nm, 256);
  hipLaunchKernelGGL(k_bp_fill_i32, dim3(1), dim3(64), 0, s, cnt, 4, 0);
  VPIC_LAUNCHED("k_bp_fill_i32");
  hipLaunchKernelGGL(k_bp_check, dim3(nb), dim3(256), 0, s, p, np, pm, nm,
                     g->nv, g->sy, g->sz, cnt);
  VPIC_LAUNCHED("k_bp_check");
  int h[4];
  VPIC_HIP(hipMemcpyAsync(h, cnt, sizeof h, hipMemcpyDeviceToHost, s));
  VPIC_HIP(hipStreamSynchronize(s));
  if (h[2])
    return vpic_set_error(VPIC_EINVAL,
                          "boundary_p: %d movers out of order, out of range, "
                          "or whose particle is not on a voxel face", h[2]);
  hipLaunchKernelGGL(k_bp_fill_i32, dim3(nb < 4096 ? nb : 4096), dim3(256), 0,
                     s, link, nm, -1);
  VPIC_LAUNCHED("k_bp_fill_i32");
  hipLaunchKernelGGL(k_bp_classify, dim3(nb), dim3(256), 0, s, p, np, pm, nm,
                     f, q, g->r8V, G, hits, link, cnt);
  VPIC_LAUNCHED("k_bp_classify");
  hipLaunchKernelGGL(k_bp_jump_init, dim3(nb), dim3(256), 0, s, link, d, nm);
  VPIC_LAUNCHED("k_bp_jump_init");
  for (int span = 1; span < nm; span *= 2) {
    hipLaunchKernelGGL(k_bp_jump, dim3(nb), dim3(256), 0, s, d, e, nm);
    VPIC_LAUNCHED("k_bp_jump");
    int *t = d;
    d = e;
    e = t;
  }
  hipLaunchKernelGGL(k_bp_fill, dim3(nb), dim3(256), 0, s, p, np, pm, nm, d);
  VPIC_LAUNCHED("k_bp_fill");
  VPIC_HIP(hipMemcpyAsync(h, cnt, sizeof h, hipMemcpyDeviceToHost, s));
  VPIC_HIP(hipStreamSynchronize(s));
  *np_host = np - nm;
  counts_host[0] = h[0];
  counts_host[1] = h[1];
  return VPIC_OK;
}
