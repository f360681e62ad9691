// collision.hip -- the Takizuka-Abe binary collision operator
// (src/collision/pipeline/takizuka_abe_pipeline.cc:81-215), one lane per
// voxel over the sort partition.
//
// The reference runs the voxels in order on a pipeline's sequential SFMT
// generator.  Here every voxel draws from its own counter-based stream
// (Philox4x32-10, Salmon et al., SC'11: key = the seed, counter = {draw/4,
// voxel, call}), so the voxels are independent and the result does not
// depend on the launch shape.  Within a voxel the operations are the
// reference's, in its order: the Fisher-Yates shuffle of species i with
// rejection sampling, the three-particle start of an odd intra-species
// voxel, the pairing of the larger set round the smaller, and
// takizuka_abe_collision's float arithmetic (the scattering phase's sin and
// cos as correctly rounded floats).  Compiled without FMA contraction, like
// the rest.
#include "common.h"

#include <cfloat>

namespace {

struct TaRng {
  uint32_t key0, key1, ctr[4], buf[4];
  uint32_t n;

  __device__ void philox() {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key0, k1 = key1;
#pragma unroll
    for (int r = 0; r < 10; r++) {
      const uint32_t hi0 = __umulhi(0xD2511F53u, c0), lo0 = 0xD2511F53u * c0;
      const uint32_t hi1 = __umulhi(0xCD9E8D57u, c2), lo1 = 0xCD9E8D57u * c2;
      c0 = hi1 ^ c1 ^ k0;
      c1 = lo1;
      c2 = hi0 ^ c3 ^ k1;
      c3 = lo0;
      k0 += 0x9E3779B9u;
      k1 += 0xBB67AE85u;
    }
    buf[0] = c0; buf[1] = c1; buf[2] = c2; buf[3] = c3;
  }
  __device__ uint32_t u32() {
    if ((n & 3) == 0) {
      ctr[0] = n >> 2;
      philox();
    }
    return buf[n++ & 3];
  }
  __device__ float frand_c0() {
    return (float)(u32() >> 8) * (1.0f / 16777216.0f);
  }
  __device__ float frandn() {  // Box-Muller in double
    const double u1 = ((double)(u32() >> 8) + 1.0) * (1.0 / 16777216.0);
    const double u2 = (double)(u32() >> 8) * (1.0 / 16777216.0);
    return (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
  }
};

// takizuka_abe_collision (takizuka_abe_pipeline.cc:14-79)
__device__ void ta_collide(vpic_particle_t *pi, vpic_particle_t *pj,
                           float mu_i, float mu_j, float std, TaRng &rng) {
  float dd, ur, urx, ury, urz, tx, ty, tz, t0, t1, t2, wi, wj, st[3];
  int d0, d1, d2;
  urx = pi->ux - pj->ux;
  ury = pi->uy - pj->uy;
  urz = pi->uz - pj->uz;
  wi = pi->w;
  wj = pj->w;
  t0 = urx * urx; d0 = 0; d1 = 1; d2 = 2; t1 = t0; ur = t0;
  t0 = ury * ury;
  if (t0 < t1) { d0 = 1; d1 = 2; d2 = 0; t1 = t0; }
  ur += t0;
  t0 = urz * urz;
  if (t0 < t1) { d0 = 2; d1 = 0; d2 = 1; }
  ur += t0;
  ur = sqrtf(ur);
  st[0] = urx; st[1] = ury; st[2] = urz;
  t1 = st[d1];
  t2 = st[d2];
  t0 = 1 / sqrtf(t1 * t1 + t2 * t2 + FLT_MIN);
  st[d0] = 0;
  st[d1] = t0 * t2;
  st[d2] = -t0 * t1;
  tx = st[0]; ty = st[1]; tz = st[2];
  t0 = 1;
  t2 = 1 / ur;
  t1 = std * sqrtf(t2) * t2;
  if (t0 < t1) t1 = t0;
  dd = t1 * rng.frandn();
  t0 = 2 * dd / (1 + dd * dd);
  t1 = (float)(2 * M_PI * rng.frand_c0());
  // The reference's sin(t1)/cos(t1) of a float take the float overloads
  // (C++ <cmath>); they are formed here as the double functions rounded to
  // float (the correctly rounded float value, identical on host and
  // device), then multiplied in float as there.
  {
    double sn, cs;
    sincos((double)t1, &sn, &cs);
    t2 = t0 * (float)sn;
    t1 = t0 * ur * (float)cs;
  }
  t0 *= -dd;
  const float s0 = (t0 * urx + t1 * tx) + t2 * (ury * tz - urz * ty);
  const float s1 = (t0 * ury + t1 * ty) + t2 * (urz * tx - urx * tz);
  const float s2 = (t0 * urz + t1 * tz) + t2 * (urx * ty - ury * tx);
  t0 = rng.frand_c0();
  t1 = mu_i;
  t2 = mu_j;
  if (wj < wi && wi * t0 > wj) t1 = 0;
  if (wi < wj && wj * t0 > wi) t2 = 0;
  pi->ux += t1 * s0;
  pi->uy += t1 * s1;
  pi->uz += t1 * s2;
  pj->ux -= t2 * s0;
  pj->uy -= t2 * s1;
  pj->uz -= t2 * s2;
}

__global__ void k_ta_clear(int *a) {
  if (threadIdx.x == 0) a[0] = 0;
}

DeviceScratch &collision_scratch() {
  static thread_local DeviceScratch s[16];
  int dev = 0;
  (void)hipGetDevice(&dev);
  return s[dev & 15];
}

struct TaArgs {
  vpic_particle_t *pi, *pj;
  const int *part_i, *part_j;
  int np_i, np_j, same, v1;
  float mu_i, mu_j, dtinterval_dV;
  double cvar;
  uint32_t key0, key1, call0, call1;
};

__global__ __launch_bounds__(256) void k_takizuka_abe(TaArgs a, int *nbad) {
  const int v = blockIdx.x * 256 + threadIdx.x;
  if (v >= a.v1) return;
  vpic_particle_t *spi_p = a.pi, *spj_p = a.pj;
  int k0 = a.part_i[v];
  int nk = a.part_i[v + 1] - k0;
  if (!nk) return;
  // a partition that does not index this species' particles: skip the
  // voxel and count it (reported as an error by the host)
  if (k0 < 0 || nk < 0 || k0 + nk > a.np_i) {
    atomicAdd(nbad, 1);
    return;
  }
  TaRng rng;
  rng.key0 = a.key0;
  rng.key1 = a.key1;
  rng.ctr[1] = (uint32_t)v;
  rng.ctr[2] = a.call0;
  rng.ctr[3] = a.call1;
  rng.n = 0;
  float density_k = 0, density_l, std;
  const int k1 = k0 + nk;
  int i, j, l0, nl;
  for (i = k0; i < k1 - 1; ++i) {
    const uint32_t rn = UINT32_MAX / (uint32_t)(k1 - i);
    do { j = i + (int)(rng.u32() / rn); } while (j >= k1);
    const vpic_particle_t t = spi_p[j];
    spi_p[j] = spi_p[i];
    spi_p[i] = t;
    density_k += spi_p[i].w;
  }
  density_k += spi_p[i].w;
  if (a.same) {
    if (nk % 2 && nk >= 3) {
      std = sqrtf((float)(0.5 * density_k * a.cvar * a.dtinterval_dV));
      ta_collide(spi_p + k0, spi_p + k0 + 1, a.mu_i, a.mu_j, std, rng);
      ta_collide(spi_p + k0, spi_p + k0 + 2, a.mu_i, a.mu_j, std, rng);
      ta_collide(spi_p + k0 + 1, spi_p + k0 + 2, a.mu_i, a.mu_j, std, rng);
      nk -= 3;
      k0 += 3;
    }
    nl = nk = nk / 2;
    l0 = k0 + nk;
    if (!nk) return;
    std = sqrtf((float)(a.cvar * density_k * a.dtinterval_dV));
  } else {
    l0 = a.part_j[v];
    nl = a.part_j[v + 1] - l0;
    if (!nl) return;
    if (l0 < 0 || nl < 0 || l0 + nl > a.np_j) {
      atomicAdd(nbad, 1);
      return;
    }
    density_l = 0;
    for (i = 0; i < nl; ++i) density_l += spj_p[l0 + i].w;
    std = sqrtf((float)(a.cvar * (density_l > density_k ? density_k
                                                        : density_l) *
                        a.dtinterval_dV));
  }
  int ii, rn;
  if (nk > nl) {
    ii = nk / nl;
    rn = nk - ii * nl;
    for (i = 0; i < rn; ++i, ++l0)
      for (j = 0; j <= ii; ++j, ++k0)
        ta_collide(spi_p + k0, spj_p + l0, a.mu_i, a.mu_j, std, rng);
    for (; i < nl; ++i, ++l0)
      for (j = 0; j < ii; ++j, ++k0)
        ta_collide(spi_p + k0, spj_p + l0, a.mu_i, a.mu_j, std, rng);
  } else {
    ii = nl / nk;
    rn = nl - ii * nk;
    for (i = 0; i < rn; ++i, ++k0)
      for (j = 0; j <= ii; ++j, ++l0)
        ta_collide(spi_p + k0, spj_p + l0, a.mu_i, a.mu_j, std, rng);
    for (; i < nk; ++i, ++k0)
      for (j = 0; j < ii; ++j, ++l0)
        ta_collide(spi_p + k0, spj_p + l0, a.mu_i, a.mu_j, std, rng);
  }
}

}  // namespace

int launch_takizuka_abe(vpic_particle_t *pi, const int *part_i, int np_i,
                        float qi, float mi, vpic_particle_t *pj,
                        const int *part_j, int np_j, float qj, float mj,
                        double cvar0, int interval, uint64_t seed,
                        uint64_t call, const vpic_grid_t *g, hipStream_t s) {
  TaArgs a;
  a.pi = pi;
  a.pj = pj;
  a.part_i = part_i;
  a.part_j = part_j;
  a.np_i = np_i;
  a.np_j = np_j;
  a.same = pi == pj;
  a.v1 = g->nx + g->sy * g->ny + g->sz * g->nz + 1;  // VOXEL(nx,ny,nz)+1
  // the coupling constants as takizuka_abe_pipeline.cc:95-99 forms them
  a.dtinterval_dV = (g->dt * (float)interval) / g->dV;
  const float mu = (mi * mj) / (mi + mj);
  a.mu_i = mj / (mi + mj);
  a.mu_j = mi / (mi + mj);
  a.cvar = cvar0 * (qi * qi * qj * qj) / (mu * mu);
  a.key0 = (uint32_t)seed;
  a.key1 = (uint32_t)(seed >> 32);
  a.call0 = (uint32_t)call;
  a.call1 = (uint32_t)(call >> 32);
  DeviceScratch &S = collision_scratch();
  if (int rc = scratch_reserve(S, 64)) return rc;
  int *nbad = (int *)S.ptr;
  hipLaunchKernelGGL(k_ta_clear, dim3(1), dim3(64), 0, s, nbad);
  VPIC_LAUNCHED("k_ta_clear");
  hipLaunchKernelGGL(k_takizuka_abe, dim3((a.v1 + 255) / 256), dim3(256), 0,
                     s, a, nbad);
  VPIC_LAUNCHED("k_takizuka_abe");
  int h = 0;
  VPIC_HIP(hipMemcpyAsync(&h, nbad, sizeof(int), hipMemcpyDeviceToHost, s));
  VPIC_HIP(hipStreamSynchronize(s));
  if (h)
    return vpic_set_error(VPIC_EINVAL,
                          "takizuka_abe: the partition of %d voxels does not "
                          "index the species' particles", h);
  return VPIC_OK;
}
