// hydro.hip -- hydro moments on gfx950: accumulate_hydro_p,
// clear / synchronize of the hydro array.
//
// Reference: src/species_advance/standard/pipeline/hydro_p_pipeline.cc
// (:17-210 per particle, :214-242 driver) and
// src/sf_interface/hydro_array.cc:131-311 (synchronize_hydro_array).
//
// Each particle's half E kick, half Boris rotation, velocity and trilinear
// node weights are the reference's expressions in the same order; the 14
// node sums differ only in summation order.  With the sort-time partition
// the particles of a tile of cells are pushed by one workgroup: a lane
// sums its consecutive particles of one voxel in registers (8 nodes x 14
// moments), adds the sums to an LDS window of the tile's nodes when the
// voxel changes, and the window is added to the array at the end.
#include "common.h"

namespace {

struct HydroCoef {
  float qsp, qdt_2mc, qdt_4mc2, mspc, c, r8V;
};

struct HydroTile {
  int tx, ty, tz, ntx, nty, ntz;
};

constexpr int HY_BS = 512;
constexpr int HY_MAXROWS = 256;

// The per-particle part of accumulate_hydro_p_pipeline_scalar
// (hydro_p_pipeline.cc:65-172): node weights wn[8] and the 14 moments per
// node, m[k][c], added to acc.
__device__ __forceinline__ void hydro_particle(const float4 A, const float4 B,
                                               const vpic_interpolator_t *fi,
                                               const HydroCoef &C,
                                               float acc[8][14]) {
  const float one = 1.0, one_third = 1.0 / 3.0;
  float dx = A.x, dy = A.y, dz = A.z;
  float ux = B.x, uy = B.y, uz = B.z;
  const float w = B.w;
  const float4 e0 = reinterpret_cast<const float4 *>(fi)[0];
  const float4 e1 = reinterpret_cast<const float4 *>(fi)[1];
  const float4 e2 = reinterpret_cast<const float4 *>(fi)[2];
  const float4 b0 = reinterpret_cast<const float4 *>(fi)[3];
  const float2 b1 = reinterpret_cast<const float2 *>(fi)[8];
  float vx, vy, vz, ke_mc, t, w0, w1, w2, w3, w4, w5, w6, w7;
  ux += C.qdt_2mc * ((e0.x + dy * e0.y) + dz * (e0.z + dy * e0.w));
  uy += C.qdt_2mc * ((e1.x + dz * e1.y) + dx * (e1.z + dz * e1.w));
  uz += C.qdt_2mc * ((e2.x + dx * e2.y) + dy * (e2.z + dx * e2.w));
  w5 = b0.x + dx * b0.y;
  w6 = b0.z + dy * b0.w;
  w7 = b1.x + dz * b1.y;
  ke_mc = ux * ux + uy * uy + uz * uz;
  vz = sqrtf(one + ke_mc);
  ke_mc *= C.c / (vz + one);
  vz = C.c / vz;
  w0 = C.qdt_4mc2 * vz;
  w1 = w5 * w5 + w6 * w6 + w7 * w7;
  w2 = w0 * w0 * w1;
  w3 = w0 * (one + (one_third)*w2 * (one + 0.4f * w2));
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
  w0 = C.r8V * w;
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
#pragma unroll
  for (int k = 0; k < 8; k++) {
    t = C.qsp * wn[k];
    acc[k][0] += t * vx;
    acc[k][1] += t * vy;
    acc[k][2] += t * vz;
    acc[k][3] += t;
    t = C.mspc * wn[k];
    const float px = t * ux, py = t * uy, pz = t * uz;
    acc[k][4] += px;
    acc[k][5] += py;
    acc[k][6] += pz;
    acc[k][7] += t * ke_mc;
    acc[k][8] += px * vx;
    acc[k][9] += py * vy;
    acc[k][10] += pz * vz;
    acc[k][11] += py * vz;
    acc[k][12] += pz * vx;
    acc[k][13] += px * vy;
  }
}

// Node offsets of the 8 trilinear weights (hydro_p_pipeline.cc:194-201):
// i, i+1, i+sy, i+sy+1, i+sz, i+sz+1, i+sz+sy, i+sz+sy+1.
__device__ __forceinline__ int node_off(int k, int sy, int sz) {
  return (k & 1) + ((k >> 1) & 1) * sy + (k >> 2) * sz;
}

// Flat: one lane per particle, global float atomics (no partition).
__global__ __launch_bounds__(256) void k_hydro_flat(
    const vpic_particle_t *__restrict__ p, const int *__restrict__ start,
    int np, const vpic_interpolator_t *__restrict__ f0, float *__restrict__ hy,
    HydroCoef C, int nv, int sy, int sz, int *__restrict__ nbad) {
  const int n0 = start ? min(max(*start, 0), np) : 0;
  for (int n = n0 + blockIdx.x * blockDim.x + threadIdx.x; n < np;
       n += gridDim.x * blockDim.x) {
    const float4 *pp = reinterpret_cast<const float4 *>(p + n);
    const float4 A = pp[0], B = pp[1];
    const int ii = __float_as_int(A.w);
    if ((unsigned)ii >= (unsigned)nv || ii + sz + sy + 1 >= nv) {
      atomicAdd(nbad, 1);
      continue;
    }
    float acc[8][14];
#pragma unroll
    for (int k = 0; k < 8; k++)
#pragma unroll
      for (int c = 0; c < 14; c++) acc[k][c] = 0.0f;
    hydro_particle(A, B, f0 + ii, C, acc);
#pragma unroll
    for (int k = 0; k < 8; k++) {
      float *h = hy + 16 * (int64_t)(ii + node_off(k, sy, sz));
#pragma unroll
      for (int c = 0; c < 14; c++) atomic_add_f32(h + c, acc[k][c]);
    }
  }
}

// Tiled: one workgroup per tile of cells over the partition rows; the
// window holds the tile's nodes (x0..x1+1, ...), component-major.
__global__ __launch_bounds__(HY_BS) void k_hydro_tiled(
    const vpic_particle_t *__restrict__ p, int np, const int *__restrict__ part,
    const vpic_interpolator_t *__restrict__ f0, float *__restrict__ hy,
    HydroCoef C, GridDev G, HydroTile T, int *__restrict__ nbad) {
  extern __shared__ float hwin[];  // 14 * nnode
  __shared__ int row_start[HY_MAXROWS];
  __shared__ int row_pref[HY_MAXROWS + 1];
  const int t = threadIdx.x;
  int b = blockIdx.x;
  const int btx = b % T.ntx;
  b /= T.ntx;
  const int bty = b % T.nty;
  const int btz = b / T.nty;
  const int x0 = 1 + btx * T.tx, y0 = 1 + bty * T.ty, z0 = 1 + btz * T.tz;
  const int x1 = min(x0 + T.tx - 1, G.nx), y1 = min(y0 + T.ty - 1, G.ny),
            z1 = min(z0 + T.tz - 1, G.nz);
  const int rows_y = y1 - y0 + 1;
  const int nrow = rows_y * (z1 - z0 + 1);
  const int wnx = x1 - x0 + 2, wny = y1 - y0 + 2, wnz = z1 - z0 + 2;
  const int nnode = wnx * wny * wnz;
  const int sx1 = G.nx + 2, sy1 = G.ny + 2;
  const int sy = sx1, sz = sx1 * sy1;
  for (int r = t; r < nrow; r += HY_BS) {
    const int y = y0 + r % rows_y, z = z0 + r / rows_y;
    int s = part[vox(x0, y, z, G.nx, G.ny)];
    int e = part[vox(x1 + 1, y, z, G.nx, G.ny)];
    s = min(max(s, 0), np);
    e = min(max(e, s), np);
    row_start[r] = s;
    row_pref[r + 1] = e - s;
  }
  for (int k = t; k < 14 * nnode; k += HY_BS) hwin[k] = 0.0f;
  __syncthreads();
  if (t == 0) {
    row_pref[0] = 0;
    for (int r = 0; r < nrow; r++) row_pref[r + 1] += row_pref[r];
  }
  __syncthreads();
  const int total = row_pref[nrow];

  float acc[8][14];
#pragma unroll
  for (int k = 0; k < 8; k++)
#pragma unroll
    for (int c = 0; c < 14; c++) acc[k][c] = 0.0f;
  int key = -1;
  int bad = 0;

  auto flush = [&]() {
    const int v = key;
    const int zz = v / sz, rem = v - zz * sz, yy = rem / sy, xx = rem - yy * sy;
    const int lx = xx - x0, ly = yy - y0, lz = zz - z0;
#pragma unroll
    for (int k = 0; k < 8; k++) {
      const int ax = lx + (k & 1), ay = ly + ((k >> 1) & 1), az = lz + (k >> 2);
      if ((unsigned)ax < (unsigned)wnx && (unsigned)ay < (unsigned)wny &&
          (unsigned)az < (unsigned)wnz) {
        const int s = ax + wnx * (ay + wny * az);
#pragma unroll
        for (int c = 0; c < 14; c++) atomicAdd(&hwin[c * nnode + s], acc[k][c]);
      } else {
        float *h = hy + 16 * (int64_t)(v + node_off(k, sy, sz));
#pragma unroll
        for (int c = 0; c < 14; c++) atomic_add_f32(h + c, acc[k][c]);
      }
#pragma unroll
      for (int c = 0; c < 14; c++) acc[k][c] = 0.0f;
    }
  };

  // lane t takes one contiguous run of the tile's slots (the rows of the
  // partition back to back), so its particles mostly share a voxel and
  // the register sums reach the LDS window about once per voxel
  const int per = (total + HY_BS - 1) / HY_BS;
  const int k0 = min(t * per, total), k1 = min(k0 + per, total);
  int r = 0;
  while (r + 1 < nrow && row_pref[r + 1] <= k0) r++;
  for (int k = k0; k < k1; k++) {
    while (k >= row_pref[r + 1]) r++;
    const int n = row_start[r] + (k - row_pref[r]);
    const float4 *pp = reinterpret_cast<const float4 *>(p + n);
    const float4 A = pp[0], B = pp[1];
    const int ii = __float_as_int(A.w);
    if ((unsigned)ii >= (unsigned)G.nv || ii + sz + sy + 1 >= G.nv) {
      bad++;
      continue;
    }
    if (ii != key) {
      if (key >= 0) flush();
      key = ii;
    }
    hydro_particle(A, B, f0 + ii, C, acc);
  }
  if (key >= 0) flush();
  if (bad) atomicAdd(nbad, bad);
  __syncthreads();
  for (int k = t; k < 14 * nnode; k += HY_BS) {
    const float v = hwin[k];
    if (v == 0.0f) continue;
    const int c = k / nnode, s = k - c * nnode;
    const int ax = s % wnx, ay = (s / wnx) % wny, az = s / (wnx * wny);
    atomic_add_f32(hy + 16 * (int64_t)vox(x0 + ax, y0 + ay, z0 + az, G.nx, G.ny) + c,
                   v);
  }
}

// synchronize_hydro_array (hydro_array.cc:131-311) for one domain: local
// faces doubled, then the periodic node planes of axis X take
// lw*h + rw*remote (lw = rw = 1 for equal cells).
__global__ void k_hydro_adjust(float *hy, int4 n, int4 s, int k) {
  const int X = k % 3, Y = (X + 1) % 3, Z = (X + 2) % 3;
  const int nn[3] = {n.x, n.y, n.z}, ss[3] = {s.x, s.y, s.z};
  const int plane = k >= 3 ? nn[X] + 1 : 1;
  const int wy = nn[Y] + 1, wz = nn[Z] + 1;
  for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < wy * wz;
       l += gridDim.x * blockDim.x) {
    const int yy = 1 + l % wy, zz = 1 + l / wy;
    float *h = hy + 16 * (int64_t)(plane * ss[X] + yy * ss[Y] + zz * ss[Z]);
#pragma unroll
    for (int q = 0; q < 14; q++) h[q] *= 2;
  }
}

__global__ void k_hydro_fold(float *hy, int4 n, int4 s, int X, float lw,
                             float rw) {
  const int Y = (X + 1) % 3, Z = (X + 2) % 3;
  const int nn[3] = {n.x, n.y, n.z}, ss[3] = {s.x, s.y, s.z};
  const int wy = nn[Y] + 1, wz = nn[Z] + 1;
  for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < wy * wz;
       l += gridDim.x * blockDim.x) {
    const int yy = 1 + l % wy, zz = 1 + l / wy;
    const int64_t off = yy * ss[Y] + zz * ss[Z];
    float *a = hy + 16 * (int64_t)(1 * ss[X] + off);
    float *b = hy + 16 * (int64_t)((nn[X] + 1) * ss[X] + off);
#pragma unroll
    for (int q = 0; q < 14; q++) {
      const float va = a[q], vb = b[q];
      b[q] = lw * vb + rw * va;
      a[q] = lw * va + rw * vb;
    }
  }
}

__global__ void k_hydro_clear_count(int *c) {
  if (threadIdx.x == 0) c[0] = 0;
}

__global__ __launch_bounds__(256) void k_hydro_zero(float4 *h, int64_t n4) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n4;
       i += (int64_t)gridDim.x * blockDim.x)
    h[i] = make_float4(0.f, 0.f, 0.f, 0.f);
}

struct HydroScratch {
  DeviceScratch nbad;
};
HydroScratch &hydro_scratch() {
  static thread_local HydroScratch s[16];
  int dev = 0;
  (void)hipGetDevice(&dev);
  return s[dev & 15];
}

inline int blocks_for(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

}  // namespace

int launch_clear_hydro(vpic_hydro_t *h, const vpic_grid_t *g, hipStream_t s) {
  const int64_t n4 = (int64_t)g->nv * 4;
  int nb = blocks_for(n4, 256);
  if (nb > 65536) nb = 65536;
  hipLaunchKernelGGL(k_hydro_zero, dim3(nb), dim3(256), 0, s, (float4 *)h, n4);
  VPIC_LAUNCHED("k_hydro_zero");
  return VPIC_OK;
}

int launch_accumulate_hydro_p(vpic_hydro_t *h, const vpic_particle_t *p,
                              int np, float q, float m,
                              const vpic_interpolator_t *f,
                              const vpic_grid_t *g, const int *partition,
                              hipStream_t s) {
  HydroScratch &S = hydro_scratch();
  int rc;
  if ((rc = scratch_reserve(S.nbad, 64))) return rc;
  int *nbad = (int *)S.nbad.ptr;
  hipLaunchKernelGGL(k_hydro_clear_count, dim3(1), dim3(64), 0, s, nbad);
  VPIC_LAUNCHED("k_hydro_clear_count");
  // hydro_p_pipeline.cc:31-36, 232-234
  HydroCoef C;
  C.qsp = q;
  C.qdt_2mc = (q * g->dt) / (2 * m * g->cvac);
  C.qdt_4mc2 = C.qdt_2mc / (2 * g->cvac);
  C.mspc = g->cvac * m;
  C.c = g->cvac;
  C.r8V = g->r8V;
  if (np > 0) {
    if (partition) {
      HydroTile T = {8, 8, 4, 0, 0, 0};
      T.ntx = (g->nx + T.tx - 1) / T.tx;
      T.nty = (g->ny + T.ty - 1) / T.ty;
      T.ntz = (g->nz + T.tz - 1) / T.tz;
      const size_t lds = sizeof(float) * 14 * (T.tx + 1) * (T.ty + 1) * (T.tz + 1);
      const GridDev G = make_grid_dev(g);
      hipLaunchKernelGGL(k_hydro_tiled, dim3(T.ntx * T.nty * T.ntz),
                         dim3(HY_BS), lds, s, p, np, partition, f,
                         (float *)h, C, G, T, nbad);
      VPIC_LAUNCHED("k_hydro_tiled");
    }
    // without a partition every particle, with one the particles past it
    // (added since the sort), one lane each
    int nb = blocks_for(np, 256);
    if (nb > 16384) nb = 16384;
    hipLaunchKernelGGL(k_hydro_flat, dim3(nb), dim3(256), 0, s, p,
                       partition ? partition + g->nv : (const int *)nullptr,
                       np, f, (float *)h, C, g->nv, g->sy, g->sz, nbad);
    VPIC_LAUNCHED("k_hydro_flat");
  }
  int h_bad = 0;
  VPIC_HIP(hipMemcpyAsync(&h_bad, nbad, sizeof(int), hipMemcpyDeviceToHost, s));
  VPIC_HIP(hipStreamSynchronize(s));
  if (h_bad)
    return vpic_set_error(VPIC_EINVAL,
                          "accumulate_hydro_p: %d particles with a voxel "
                          "index outside the interior were skipped", h_bad);
  return VPIC_OK;
}

int launch_synchronize_hydro(vpic_hydro_t *h, const vpic_grid_t *g,
                             hipStream_t s) {
  const int4 n = make_int4(g->nx, g->ny, g->nz, 0);
  const int4 st = make_int4(1, g->sy, g->sz, 0);
  int mm = g->nx > g->ny ? g->nx : g->ny;
  mm = mm > g->nz ? mm : g->nz;
  int fb = blocks_for((int64_t)(mm + 1) * (mm + 1), 256);
  if (fb > 1024) fb = 1024;
  for (int k = 0; k < 6; k++) {
    if (g->fbc[k] == VPIC_BC_PERIODIC) continue;
    hipLaunchKernelGGL(k_hydro_adjust, dim3(fb), dim3(256), 0, s, (float *)h,
                       n, st, k);
    VPIC_LAUNCHED("k_hydro_adjust");
  }
  for (int X = 0; X < 3; X++) {
    if (g->fbc[X] != VPIC_BC_PERIODIC || g->fbc[X + 3] != VPIC_BC_PERIODIC)
      continue;
    // hydro_array.cc END_RECV: the weights in float as there
    const float d = X == 0 ? g->dx : X == 1 ? g->dy : g->dz;
    float rw = d, lw = rw + d;
    rw /= lw;
    lw = d / lw;
    lw += lw;
    rw += rw;
    hipLaunchKernelGGL(k_hydro_fold, dim3(fb), dim3(256), 0, s, (float *)h, n,
                       st, X, lw, rw);
    VPIC_LAUNCHED("k_hydro_fold");
  }
  return VPIC_OK;
}
