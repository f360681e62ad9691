// ref_ta.cc -- harness for the reference's Takizuka-Abe operator
// (src/collision/pipeline/takizuka_abe_pipeline.cc:14-215, compiled here
// from the lanl/vpic source where it lies).  The operator's random draws
// (uirand, frand_c0, frandn) are redirected to the oracle's counter-based
// stream (Philox4x32-10, key = seed, counter = {draw/4, voxel, lo(call),
// hi(call)}; uirand = a word, frand_c0 = (word >> 8) 2^-24, frandn =
// Box-Muller in double from two words), so the reference's own shuffle,
// pairing and collision arithmetic can be compared with the oracle's on the
// same numbers.  The stream restarts at every voxel, so the reference's
// pipeline is run once per voxel on partitions that hold only that voxel's
// particles.  apply_takizuka_abe_pipeline (the dispatcher glue) is dropped
// by --gc-sections.  TEST INFRASTRUCTURE ONLY.
// the collision sources include their private header only under IN_collision
#define IN_collision
#include "src/collision/pipeline/collision_pipeline.h"

#include <math.h>
#include <stdint.h>
#include <string.h>

#include <vector>

#include "ref_common.h"

namespace {

struct philox_stream {
  uint32_t key[2], ctr[4], buf[4];
  uint32_t n;
};

void philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
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
  out[0] = c0;
  out[1] = c1;
  out[2] = c2;
  out[3] = c3;
}

uint32_t stream_word(rng_t *r) {
  philox_stream *s = (philox_stream *)r;
  if ((s->n & 3) == 0) {
    s->ctr[0] = s->n >> 2;
    philox4x32(s->ctr, s->key, s->buf);
  }
  return s->buf[s->n++ & 3];
}

float stream_frand_c0(rng_t *r) {
  return (float)(stream_word(r) >> 8) * (1.0f / 16777216.0f);
}

float stream_frandn(rng_t *r) {
  const double u1 = ((double)(stream_word(r) >> 8) + 1.0) * (1.0 / 16777216.0);
  const double u2 = (double)(stream_word(r) >> 8) * (1.0 / 16777216.0);
  return (float)(sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2));
}

}  // namespace

#define uirand(r) stream_word(r)
#define frand_c0(r) stream_frand_c0(r)
#define frandn(r) stream_frandn(r)

#include "src/collision/pipeline/takizuka_abe_pipeline.cc"

// takizuka_abe_pipeline_scalar(cm, 0, 1) voxel by voxel: for voxel v the
// species' partitions are replaced by ones that are empty except at v, and
// the stream is started at {0, v, call}.  spi_p == spj_p (with part_i ==
// part_j) is the intra-species operator.
extern "C" void ref_takizuka_abe(vpic_particle_t *spi_p, const int *part_i,
                                 float qi, float mi, vpic_particle_t *spj_p,
                                 const int *part_j, float qj, float mj,
                                 double cvar0, int interval, uint64_t seed,
                                 uint64_t call, const vpic_grid_t *g) {
  grid_t rg;
  ref_fill_grid(&rg, g);
  const int same = spi_p == spj_p;
  species_t si, sj;
  memset(&si, 0, sizeof si);
  memset(&sj, 0, sizeof sj);
  std::vector<int> pi_v(g->nv + 1), pj_v(g->nv + 1);
  si.p = (particle_t *)spi_p;
  si.q = qi;
  si.m = mi;
  si.g = &rg;
  si.partition = pi_v.data();
  sj.p = (particle_t *)spj_p;
  sj.q = qj;
  sj.m = mj;
  sj.g = &rg;
  sj.partition = pj_v.data();
  philox_stream st;
  rng_t *rngs[1] = {(rng_t *)&st};
  rng_pool_t pool;
  pool.rng = rngs;
  pool.n_rng = 1;
  takizuka_abe_t cm;
  memset(&cm, 0, sizeof cm);
  cm.spi = &si;
  cm.spj = same ? &si : &sj;
  cm.rp = &pool;
  cm.interval = interval;
  cm.cvar0 = cvar0;
  const int v1 = g->nx + g->sy * g->ny + g->sz * g->nz + 1;
  for (int v = 0; v < v1; v++) {
    for (int u = 0; u <= g->nv; u++) {
      pi_v[u] = u <= v ? part_i[v] : part_i[v + 1];
      pj_v[u] = u <= v ? part_j[v] : part_j[v + 1];
    }
    st.key[0] = (uint32_t)seed;
    st.key[1] = (uint32_t)(seed >> 32);
    st.ctr[1] = (uint32_t)v;
    st.ctr[2] = (uint32_t)call;
    st.ctr[3] = (uint32_t)(call >> 32);
    st.n = 0;
    takizuka_abe_pipeline_scalar(&cm, 0, 1);
  }
}
