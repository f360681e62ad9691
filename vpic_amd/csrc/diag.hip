// diag.hip -- diagnostics and loading on gfx950: energy_p (FP32 terms, FP64
// sums), accumulate_rho_p (trilinear charge deposit) and the uniform
// thermal-plasma loader used by the benchmarks.
//
// References: src/species_advance/standard/pipeline/energy_p_pipeline.cc
// :18-68,76-115; src/species_advance/standard/rho_p.cc:23-113;
// src/vpic/misc.cc:12-97 (cell-local position mapping of inject_particle).
#include "common.h"

namespace {

__global__ __launch_bounds__(256) void k_energy_p(
    const vpic_particle_t *__restrict__ p, int np,
    const vpic_interpolator_t *__restrict__ f, int nv, float qdt_2mc,
    float msp, double *__restrict__ part, int *__restrict__ nbad) {
  __shared__ double red[256];
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  double en = 0.0;
  if (n < np) {
    const float one = 1.0f;
    const float4 *pp = reinterpret_cast<const float4 *>(p + n);
    const float4 A = pp[0], B = pp[1];
    const float dx = A.x, dy = A.y, dz = A.z;
    const int i = __float_as_int(A.w);
    if ((unsigned)i < (unsigned)nv) {
      const vpic_interpolator_t &F = f[i];
      float v0 = B.x + qdt_2mc * ((F.ex + dy * F.dexdy) +
                                  dz * (F.dexdz + dy * F.d2exdydz));
      float v1 = B.y + qdt_2mc * ((F.ey + dz * F.deydz) +
                                  dx * (F.deydx + dz * F.d2eydzdx));
      float v2 = B.z + qdt_2mc * ((F.ez + dx * F.dezdx) +
                                  dy * (F.dezdy + dx * F.d2ezdxdy));
      v0 = v0 * v0 + v1 * v1 + v2 * v2;
      v0 = (msp * B.w) * (v0 / (one + sqrtf(one + v0)));
      en = (double)v0;
    } else {
      atomicAdd(nbad, 1);  // never dereferenced; the call reports it
    }
  }
  red[threadIdx.x] = en;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// center_p / uncenter_p, center_p_pipeline.cc:47-94 and
// uncenter_p_pipeline.cc:47-94: half E kick and half Boris rotation of the
// momenta (kick first when centering, rotation first when uncentering).
// sqrtf here is correctly rounded, which equals the reference's
// (float)sqrt of a float argument.
template <bool UNCENTER>
__global__ __launch_bounds__(256) void k_center_p(
    vpic_particle_t *__restrict__ p, int np,
    const vpic_interpolator_t *__restrict__ f, int nv, float qdt_2mc,
    float qdt_4mc, int *__restrict__ nbad) {
  const float one = 1.0f, one_third = 1.0f / 3.0f,
              two_fifteenths = 2.0f / 15.0f;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < np;
       n += gridDim.x * blockDim.x) {
    float4 *pp = reinterpret_cast<float4 *>(p + n);
    const float4 A = pp[0];
    float4 B = pp[1];
    const float dx = A.x, dy = A.y, dz = A.z;
    const int i = __float_as_int(A.w);
    if ((unsigned)i >= (unsigned)nv) {
      atomicAdd(nbad, 1);  // never dereferenced; the call reports it
      continue;
    }
    const vpic_interpolator_t &F = f[i];
    const float hax = qdt_2mc * ((F.ex + dy * F.dexdy) +
                                 dz * (F.dexdz + dy * F.d2exdydz));
    const float hay = qdt_2mc * ((F.ey + dz * F.deydz) +
                                 dx * (F.deydx + dz * F.d2eydzdx));
    const float haz = qdt_2mc * ((F.ez + dx * F.dezdx) +
                                 dy * (F.dezdy + dx * F.d2ezdxdy));
    const float cbx = F.cbx + dx * F.dcbxdx;
    const float cby = F.cby + dy * F.dcbydy;
    const float cbz = F.cbz + dz * F.dcbzdz;
    float ux = B.x, uy = B.y, uz = B.z;
    if (!UNCENTER) {
      ux += hax;
      uy += hay;
      uz += haz;
    }
    float v0 = qdt_4mc / sqrtf(one + (ux * ux + (uy * uy + uz * uz)));
    float v1 = cbx * cbx + (cby * cby + cbz * cbz);
    float v2 = (v0 * v0) * v1;
    const float v3 = v0 * (one + v2 * (one_third + v2 * two_fifteenths));
    float v4 = v3 / (one + v1 * (v3 * v3));
    v4 += v4;
    v0 = ux + v3 * (uy * cbz - uz * cby);
    v1 = uy + v3 * (uz * cbx - ux * cbz);
    v2 = uz + v3 * (ux * cby - uy * cbx);
    ux += v4 * (v1 * cbz - v2 * cby);
    uy += v4 * (v2 * cbx - v0 * cbz);
    uz += v4 * (v0 * cby - v1 * cbx);
    if (UNCENTER) {
      ux += hax;
      uy += hay;
      uz += haz;
    }
    B.x = ux;
    B.y = uy;
    B.z = uz;
    pp[1] = B;
  }
}

__global__ void k_clear_int(int *a) {
  if (threadIdx.x == 0) a[0] = 0;
}

__global__ __launch_bounds__(256) void k_fold(const double *part, int n,
                                              double *out) {
  __shared__ double red[256];
  double s = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += part[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

__global__ __launch_bounds__(256) void k_accumulate_rho_p(
    vpic_field_t *__restrict__ f, const vpic_particle_t *__restrict__ p,
    int np, float q_8V, int sy, int sz, int vmax, int *__restrict__ nbad) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= np) return;
  const float4 *pp = reinterpret_cast<const float4 *>(p + n);
  const float4 A = pp[0], B = pp[1];
  float w0 = A.x, w1 = A.y, dz = A.z;
  const int v = __float_as_int(A.w);
  if (v < 0 || v >= vmax) {  // the 8 corners must lie inside the array
    atomicAdd(nbad, 1);
    return;
  }
  float w7 = B.w * q_8V, w2, w3, w4, w5, w6;
  // rho_p.cc:64-74 with FMA(x,y,z) = z + x*y, FNMS(x,y,z) = z - x*y.
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
  atomic_add_f32(&f[v].rhof, w0);
  atomic_add_f32(&f[v + 1].rhof, w1);
  atomic_add_f32(&f[v + sy].rhof, w2);
  atomic_add_f32(&f[v + sy + 1].rhof, w3);
  atomic_add_f32(&f[v + sz].rhof, w4);
  atomic_add_f32(&f[v + sz + 1].rhof, w5);
  atomic_add_f32(&f[v + sz + sy].rhof, w6);
  atomic_add_f32(&f[v + sz + sy + 1].rhof, w7);
}

// splitmix64: a counter-based stream, so particle k's values depend only on
// (seed, k) and the loader is reproducible at any launch shape.
__device__ __forceinline__ uint64_t splitmix64(uint64_t &s) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ __forceinline__ double u01(uint64_t &s) {
  return (double)(splitmix64(s) >> 11) * (1.0 / 9007199254740992.0);
}

__global__ __launch_bounds__(256) void k_load_uniform_plasma(
    vpic_particle_t *__restrict__ p, int64_t np, int ppc, int nx, int ny,
    double3 uth, double3 ud, float w, uint64_t seed) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < np;
       k += (int64_t)gridDim.x * blockDim.x) {
    uint64_t s = seed ^ ((uint64_t)k * 0xD1B54A32D192ED03ull);
    (void)splitmix64(s);
    const int64_t cell = k / ppc;
    const int x = 1 + (int)(cell % nx);
    const int64_t t = cell / nx;
    const int y = 1 + (int)(t % ny);
    const int z = 1 + (int)(t / ny);
    // misc.cc:50-55: fraction u of the cell -> (u+u)-1, rounded to float.
    double u;
    float d[3];
    for (int a = 0; a < 3; a++) {
      u = u01(s);
      d[a] = (float)((u + u) - 1);
    }
    // Box-Muller normals.
    double n[4];
    for (int a = 0; a < 2; a++) {
      double u1 = u01(s), u2 = u01(s);
      if (u1 < 1e-300) u1 = 1e-300;
      const double r = sqrt(-2.0 * log(u1));
      n[2 * a] = r * cos(6.283185307179586 * u2);
      n[2 * a + 1] = r * sin(6.283185307179586 * u2);
    }
    float4 *pp = reinterpret_cast<float4 *>(p + k);
    pp[0] = make_float4(d[0], d[1], d[2], __int_as_float(vox(x, y, z, nx, ny)));
    pp[1] = make_float4((float)(ud.x + uth.x * n[0]),
                        (float)(ud.y + uth.y * n[1]),
                        (float)(ud.z + uth.z * n[2]), w);
  }
}

struct DiagScratch {
  DeviceScratch part, out;
};
DiagScratch &diag_scratch() {
  static thread_local DiagScratch s[16];
  int dev = 0;
  (void)hipGetDevice(&dev);
  return s[dev & 15];
}

inline int blocks_for(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

}  // namespace

int launch_energy_p(const vpic_particle_t *p, int np, float q, float m,
                    const vpic_interpolator_t *f, const vpic_grid_t *g,
                    double *host_en, hipStream_t s) {
  DiagScratch &S = diag_scratch();
  const int nblk = np > 0 ? blocks_for(np, 256) : 1;
  int rc;
  if ((rc = scratch_reserve(S.part, sizeof(double) * (size_t)nblk)) ||
      (rc = scratch_reserve(S.out, 64)))
    return rc;
  // energy_p_pipeline.cc:100: qdt_2mc in float.
  const float qdt_2mc = (q * g->dt) / (2 * m * g->cvac);
  int *nbad = (int *)((char *)S.out.ptr + 32);
  hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(64), 0, s, nbad);
  VPIC_LAUNCHED("k_clear_int");
  hipLaunchKernelGGL(k_energy_p, dim3(nblk), dim3(256), 0, s, p, np, f, g->nv,
                     qdt_2mc, m, (double *)S.part.ptr, nbad);
  VPIC_LAUNCHED("k_energy_p");
  hipLaunchKernelGGL(k_fold, dim3(1), dim3(256), 0, s,
                     (const double *)S.part.ptr, nblk, (double *)S.out.ptr);
  VPIC_LAUNCHED("k_fold");
  double out[5];
  VPIC_HIP(hipMemcpyAsync(out, S.out.ptr, sizeof(out), hipMemcpyDeviceToHost,
                          s));
  VPIC_HIP(hipStreamSynchronize(s));
  const int bad = ((const int *)out)[8];
  if (bad)
    return vpic_set_error(VPIC_EINVAL,
                          "energy_p: %d particles have a voxel index outside "
                          "[0,nv)", bad);
  *host_en = out[0] * ((double)g->cvac * (double)g->cvac);
  return VPIC_OK;
}

int launch_center_p(vpic_particle_t *p, int np, float q, float m,
                    const vpic_interpolator_t *f, const vpic_grid_t *g,
                    int uncenter, hipStream_t s) {
  if (np <= 0) return VPIC_OK;
  DiagScratch &S = diag_scratch();
  int rc;
  if ((rc = scratch_reserve(S.out, 64))) return rc;
  int *nbad = (int *)((char *)S.out.ptr + 32);
  hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(64), 0, s, nbad);
  VPIC_LAUNCHED("k_clear_int");
  // center_p_pipeline.cc:123 / uncenter_p_pipeline.cc:123 and :26-27.
  const float qdt_2mc = (q * g->dt) / (2 * m * g->cvac);
  const float hk = uncenter ? -qdt_2mc : qdt_2mc;
  const float hr = uncenter ? (float)(-0.5 * qdt_2mc) : (float)(0.5 * qdt_2mc);
  const int nblk = blocks_for(np, 256) < 8192 ? blocks_for(np, 256) : 8192;
  if (uncenter)
    hipLaunchKernelGGL(k_center_p<true>, dim3(nblk), dim3(256), 0, s, p, np,
                       f, g->nv, hk, hr, nbad);
  else
    hipLaunchKernelGGL(k_center_p<false>, dim3(nblk), dim3(256), 0, s, p, np,
                       f, g->nv, hk, hr, nbad);
  VPIC_LAUNCHED("k_center_p");
  int bad = 0;
  VPIC_HIP(hipMemcpyAsync(&bad, nbad, sizeof(int), hipMemcpyDeviceToHost, s));
  VPIC_HIP(hipStreamSynchronize(s));
  if (bad)
    return vpic_set_error(VPIC_EINVAL,
                          "center_p: %d particles have a voxel index outside "
                          "[0,nv)", bad);
  return VPIC_OK;
}

int launch_accumulate_rho_p(vpic_field_t *f, const vpic_particle_t *p, int np,
                            float q, const vpic_grid_t *g, hipStream_t s) {
  if (np <= 0) return VPIC_OK;
  DiagScratch &S = diag_scratch();
  int rc;
  if ((rc = scratch_reserve(S.out, 64))) return rc;
  int *nbad = (int *)((char *)S.out.ptr + 32);
  hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(64), 0, s, nbad);
  VPIC_LAUNCHED("k_clear_int");
  const float q_8V = q * g->r8V;
  const int vmax = g->nv - g->sz - g->sy - 1;
  hipLaunchKernelGGL(k_accumulate_rho_p, dim3(blocks_for(np, 256)), dim3(256),
                     0, s, f, p, np, q_8V, g->sy, g->sz, vmax, nbad);
  VPIC_LAUNCHED("k_accumulate_rho_p");
  int bad = 0;
  VPIC_HIP(hipMemcpyAsync(&bad, nbad, sizeof(int), hipMemcpyDeviceToHost, s));
  VPIC_HIP(hipStreamSynchronize(s));
  if (bad)
    return vpic_set_error(VPIC_EINVAL,
                          "accumulate_rho_p: %d particles outside the grid",
                          bad);
  return VPIC_OK;
}

int launch_load_uniform_plasma(vpic_particle_t *p, const vpic_grid_t *g,
                               int ppc, float uthx, float uthy, float uthz,
                               float udx, float udy, float udz, float w,
                               uint64_t seed, hipStream_t s) {
  const int64_t np = (int64_t)ppc * g->nx * g->ny * g->nz;
  push_invalidate_bins(p);
  int blk = blocks_for(np, 256);
  if (blk > 262144) blk = 262144;
  hipLaunchKernelGGL(k_load_uniform_plasma, dim3(blk), dim3(256), 0, s, p, np,
                     ppc, g->nx, g->ny, make_double3(uthx, uthy, uthz),
                     make_double3(udx, udy, udz), w, seed);
  VPIC_LAUNCHED("k_load_uniform_plasma");
  return VPIC_OK;
}
