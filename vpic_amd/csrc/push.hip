// push.hip -- advance_p on gfx950: Boris push with energy-conserving
// interpolation, in-cell Villasenor-Buneman deposit, and move_p for the
// particles that leave their cell.
//
// Reference: src/species_advance/standard/pipeline/advance_p_pipeline.cc
// (scalar kernel :21-245, driver :253-340) and
// src/species_advance/standard/move_p.cc:215-378.  The FP32 arithmetic of
// each particle is the reference's, operation for operation, so positions,
// momenta and cell indices are bit-identical; only the order in which the
// current contributions are summed differs (atomics).
//
// Two push kernels share the per-particle code:
//
//  k_advance_p_tiled (used when the sort-time voxel partition is given):
//    one workgroup per 3D tile of cells.  The tile's particles are the rows
//    of the partition written by the last sort_p.  Currents go to an LDS
//    window (the tile plus a halo of one or two cells, 96 B of 64-bit fixed
//    point per cell) and the window is added to the global accumulator once
//    at the end (k_merge_windows).  Cell crossers are queued in LDS and
//    moved by the same workgroup, so their segments also land in the
//    window.  Deposits outside the window (drifted particles, periodic
//    wrap) go to per-tile sub-buckets that the merge adds.
//
//  k_advance_p (any particle order, and the particles after the partition):
//    grid-stride over the particles, deposits straight to global atomics;
//    crossers are compacted into a list (ballot + one atomic per wavefront)
//    and moved by k_move_p_list.
//
// Deposits of a wavefront (wave_deposit) take one of three forms, chosen by
// the number of runs of equal voxels among its 64 lanes: one run -> a plain
// DPP reduction and one deposit; many short runs -> every lane deposits its
// own 12 values; otherwise a run-segmented DPP scan and one deposit per run.
// Both kernels prefetch the next particle while pushing the current one.
//
// Both end with k_order_movers: movers the grid cannot handle, sorted by
// particle index, truncated at max_nm (excess rolled back to the voxel
// index, advance_p_pipeline.cc:228-236).
#include <mutex>
#include <vector>

#include "common.h"

namespace {

struct PushCoef {
  float qdt_2mc, cdt_dx, cdt_dy, cdt_dz, qsp;
};

// The four quarter-face updates of one axis, advance_p_pipeline.cc:186-208
// and move_p.cc:281-303 (identical arithmetic).
__device__ __forceinline__ void streak4(float out[4], float qd, float mY,
                                        float mZ, float v5) {
  float v0, v1, v2, v3, v4;
  v4 = qd;
  v1 = v4 * mY;
  v0 = v4 - v1;
  v1 += v4;
  v4 = 1.0f + mZ;
  v2 = v0 * v4;
  v3 = v1 * v4;
  v4 = 1.0f - mZ;
  v0 *= v4;
  v1 *= v4;
  v0 += v5;
  v1 -= v5;
  v2 -= v5;
  v3 += v5;
  out[0] = v0;
  out[1] = v1;
  out[2] = v2;
  out[3] = v3;
}

// --------------------------------------------------------------------------
// Deposit sinks
// --------------------------------------------------------------------------
struct GlobalSink {
  float *acc;
  __device__ __forceinline__ void deposit(int voxel, const float j[12]) const {
    float *a = acc + 12 * (int64_t)voxel;
#pragma unroll
    for (int k = 0; k < 12; k++) atomic_add_f32(a + k, j[k]);
  }
};

// q = v / d for v * d < 2^32 with m = ceil(2^32 / d) (exact: the rounding
// excess v*(m*d - 2^32) / 2^32 stays below 1/d).
__device__ __forceinline__ unsigned fast_div(unsigned v, unsigned m) {
  return __umulhi(v, m);
}

typedef __attribute__((address_space(3))) unsigned long long lds_u64;

// round(x * 2^k) as a two's-complement int64, exact for |x * 2^k| < 2^51:
// adding 1.5 * 2^52 puts the rounded integer in the low mantissa bits.
__device__ __forceinline__ long long to_fixed(double y) {
  const double magic = 6755399441055744.0;
  return __double_as_longlong(y + magic) - __double_as_longlong(magic);
}

// A box of voxels [w0, w0+wn) per axis held in LDS as 64-bit fixed point
// (value * scale, scale a power of two), component-major (component k of
// cell s at lds[k * ncell + s]).  gfx950's LDS adds 64-bit integers ~25
// times faster than it adds floats, and the integer sum does not depend
// on the order of the additions.  A value too large for the fixed-point
// range, and any cell outside the window, goes to the global accumulator
// with float atomics.
// Out-of-window deposits of the tiled push, kept by the pushing (source)
// tile in 27 sub-buckets, one per neighbouring destination tile (offset
// code (dx+1) + 3(dy+1) + 9(dz+1), tile indices taken periodically).
// Record k of source tile s, code d, is rec[4 * ((27 s + d) cap + k)] .. +3
// = {voxel (bits), j0..j2}, {j3..j6}, {j7..j10}, {j11, 0, 0, 0}; cnt[2 (27 s
// + d)] (the count) and cnt[2 (27 s + d) + 1] (the bits of the largest sum of
// the 12 |j| of its records, a bound on every component, so the reader can
// fix its fixed-point scale without a pass over the records) are written at
// the end of the push.  k_merge_windows of tile t reads,
// for every code d, the sub-bucket d of the tile t - offset(d): each
// sub-bucket has exactly one reader.  Slots are taken with LDS atomics; a
// deposit further than one tile away or into a full sub-bucket goes to the
// accumulator with float atomics.
struct OowBuckets {
  int *cnt;     // 27 (count, magnitude bits) per source tile (count may
                // exceed cap)
  float4 *rec;  // 27 * cap records per source tile
  int cap;
};

// Offset code of destination tile (x, y, z) seen from tile (bx, by, bz),
// periodic in the tile indices; -1 when further than one tile on an axis.
__device__ __forceinline__ int tile_code_1(int d, int n) {
  d = d < 0 ? d + n : d;  // 0 .. n-1
  if (d == 0) return 1;
  if (d == 1) return 2;
  if (d == n - 1) return 0;
  return -1;
}

struct WindowSink {
  lds_u64 *lds;  // ncell cells of 12 components (cell-major)
  int ncell;
  double scale;  // 2^k, set once per workgroup before the first deposit
  float fscale;  // the same power of two
  float ifscale;  // 2^-k
  bool pre;       // register sums arrive scaled by 2^k (flush_scaled)
  float *acc;    // global accumulator
  int wx0, wy0, wz0, wnx, wny, wnz;
  int sx1, sy1;     // nx+2, ny+2
  unsigned mx, my;  // fast_div multipliers for sx1, sy1 (0: divide)
  int *oow;         // LDS count of deposits outside the window
  OowBuckets ob;    // where those deposits go
  int *dcnt;        // LDS: this tile's 27 sub-bucket counts
  unsigned *dmax;   // LDS: and their largest record magnitudes (bits)
  int tx, ty, tz, ntx, nty, ntz;  // tile extent and tiles per axis
  int btx, bty, btz;              // this tile
  int self;                       // its index
  __device__ __forceinline__ int slot(int v) const {
    unsigned t, z;
    if (mx) {
      t = fast_div((unsigned)v, mx);
      z = fast_div(t, my);
    } else {
      t = (unsigned)v / (unsigned)sx1;
      z = t / (unsigned)sy1;
    }
    const int x = v - (int)t * sx1 - wx0;
    const int y = (int)t - (int)z * sy1 - wy0;
    const int zz = (int)z - wz0;
    if ((unsigned)x < (unsigned)wnx && (unsigned)y < (unsigned)wny &&
        (unsigned)zz < (unsigned)wnz)
      return x + wnx * (y + wny * zz);
    return -1;
  }
  __device__ __forceinline__ void deposit(int voxel, const float j[12]) const {
    const int s = slot(voxel);
    float *g = acc + 12 * (int64_t)voxel;
    if (s >= 0) {
      lds_u64 *a = lds + 12 * s;
      // Scaling by a power of two is exact in float, so when every scaled
      // component is below 2^31 the rounding to an integer is done in FP32
      // (the same integer the FP64 path gives).
      // (the sum of the magnitudes bounds every component and is NaN when
      // one is, which then takes the path below)
      const float fs = fscale;
      float big = fabsf(j[0]);
#pragma unroll
      for (int k = 1; k < 12; k++) big += fabsf(j[k]);
      if (big * fs < 2147483520.0f) {
#pragma unroll
        for (int k = 0; k < 12; k++)
          __hip_atomic_fetch_add(a + k,
                                 (unsigned long long)(long long)__float2int_rn(j[k] * fs),
                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        return;
      }
#pragma unroll
      for (int k = 0; k < 12; k++) {
        const double y = (double)j[k] * scale;
        if (fabs(y) < 2251799813685248.0)  // 2^51
          __hip_atomic_fetch_add(a + k,
                                 (unsigned long long)to_fixed(y),
                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        else
          atomic_add_f32(g + k, j[k]);
      }
    } else {
      atomicAdd(oow, 1);
      // to this tile's sub-bucket for the destination tile (an LDS atomic
      // and a 64-B store instead of 12 float atomics, which this part
      // performs beyond the L2)
      unsigned t1, zz;
      if (mx) {
        t1 = fast_div((unsigned)voxel, mx);
        zz = fast_div(t1, my);
      } else {
        t1 = (unsigned)voxel / (unsigned)sx1;
        zz = t1 / (unsigned)sy1;
      }
      const int x = voxel - (int)t1 * sx1, y = (int)t1 - (int)zz * sy1;
      const int cx = tile_code_1((x - 1) / tx - btx, ntx);
      const int cy = tile_code_1((y - 1) / ty - bty, nty);
      const int cz = tile_code_1(((int)zz - 1) / tz - btz, ntz);
      // a non-finite deposit keeps the float path (its NaN or infinity
      // reaches the accumulator as it would have; the merge's fixed point
      // needs finite records)
      float mag = fabsf(j[0]);
#pragma unroll
      for (int c = 1; c < 12; c++) mag += fabsf(j[c]);
      const int d = (cx < 0 || cy < 0 || cz < 0 || !(mag < 3.0e38f))
                        ? -1
                        : cx + 3 * cy + 9 * cz;
      // one LDS atomic per code and wavefront: the lanes depositing to the
      // same sub-bucket take consecutive slots
      int k = ob.cap;
      if (ob.cap > 0) {
        const int lane = lane_id();
        uint64_t todo = __ballot(d >= 0);
        while (todo) {
          const int leader = __ffsll((unsigned long long)todo) - 1;
          const int dl = __shfl(d, leader);
          const uint64_t same = __ballot(d == dl) & todo;
          int base = 0;
          if (lane == leader) base = atomicAdd(dcnt + dl, __popcll(same));
          base = __shfl(base, leader);
          if (d == dl) k = base + lanes_below(same);
          todo &= ~same;
        }
      }
      if (d >= 0 && k < ob.cap) {
        atomicMax(dmax + d, __float_as_uint(mag));  // mag >= 0, finite
        float4 *r = ob.rec + 4 * ((int64_t)(27 * self + d) * ob.cap + k);
        r[0] = make_float4(__int_as_float(voxel), j[0], j[1], j[2]);
        r[1] = make_float4(j[3], j[4], j[5], j[6]);
        r[2] = make_float4(j[7], j[8], j[9], j[10]);
        r[3] = make_float4(j[11], 0.0f, 0.0f, 0.0f);
      } else {
#pragma unroll
        for (int c = 0; c < 12; c++) atomic_add_f32(g + c, j[c]);
      }
    }
  }
  // A lane's register sum js (12 components) for voxel `voxel`, whose
  // window cell `cell` (-1 outside the window) the caller computed when
  // the voxel was first seen.  With pre set, js is already scaled by 2^k
  // (the deposits were formed with the charge times 2^k, an exact power-of-
  // two scaling), so a sum whose components are all below 2^31 in magnitude
  // converts with one truncating conversion per component (the truncation
  // error, below 2^-k, is far below FP32 rounding of the sum).  Anything
  // else is unscaled and takes deposit().
  __device__ __forceinline__ void flush_scaled(int voxel, int cell,
                                               const float js[12]) const {
    if (pre && cell >= 0) {
      // the sum of the magnitudes bounds every component and is NaN when
      // one is, which then takes deposit()
      float big = fabsf(js[0]);
#pragma unroll
      for (int k = 1; k < 12; k++) big += fabsf(js[k]);
      if (big < 2147483520.0f) {
        lds_u64 *a = lds + 12 * cell;
#pragma unroll
        for (int k = 0; k < 12; k++)
          __hip_atomic_fetch_add(a + k, (unsigned long long)(long long)(int)js[k],
                                 __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        return;
      }
    }
    float j[12];
#pragma unroll
    for (int k = 0; k < 12; k++) j[k] = pre ? js[k] * ifscale : js[k];
    deposit(voxel, j);
  }
};

// --------------------------------------------------------------------------
// move_p (move_p.cc:215-378).  The particle lives in registers for the walk;
// `mv` is the mover (remaining displacement, particle index).  Returns 1
// when the particle hit a face the grid cannot handle; its voxel index is
// then 8*voxel + face.
// --------------------------------------------------------------------------
template <class Sink>
__device__ int dev_move_p(vpic_particle_t *__restrict__ p0,
                          vpic_particle_mover_t &mv, const Sink &sink,
                          const GridDev &G, float qsp,
                          int *__restrict__ vkey = nullptr) {
  float4 *pp = reinterpret_cast<float4 *>(p0 + mv.i);
  float4 A = pp[0], B = pp[1];
  float pd[3] = {A.x, A.y, A.z};
  int pi = __float_as_int(A.w);
  float pu[3] = {B.x, B.y, B.z};
  float md[3] = {mv.dispx, mv.dispy, mv.dispz};
  const float q = qsp * B.w;
  int ret = 0;
  for (;;) {
    float s_mid[3] = {pd[0], pd[1], pd[2]};
    float s_disp[3] = {md[0], md[1], md[2]};
    float s_dir[3], v[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      s_dir[k] = (s_disp[k] > 0.0f) ? 1.0f : -1.0f;
      v[k] = (s_disp[k] == 0.0f) ? 3.4e38f : (s_dir[k] - s_mid[k]) / s_disp[k];
    }
    float v3 = 2.0f;
    int axis = 3;
    if (v[0] < v3) v3 = v[0], axis = 0;
    if (v[1] < v3) v3 = v[1], axis = 1;
    if (v[2] < v3) v3 = v[2], axis = 2;
    v3 *= 0.5f;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      s_disp[k] *= v3;
      s_mid[k] += s_disp[k];
    }
    // move_p.cc:277: promoted to double by the 1.0/3.0 constant.
    const float v5 =
        (float)((double)(q * s_disp[0] * s_disp[1] * s_disp[2]) * (1.0 / 3.0));
    float j[12];
    streak4(j + 0, q * s_disp[0], s_mid[1], s_mid[2], v5);
    streak4(j + 4, q * s_disp[1], s_mid[2], s_mid[0], v5);
    streak4(j + 8, q * s_disp[2], s_mid[0], s_mid[1], v5);
    sink.deposit(pi, j);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      md[k] -= s_disp[k];
      pd[k] += s_disp[k] + s_disp[k];
    }
    if (axis == 3) break;
    const float d0 = s_dir[axis];
    pd[axis] = d0;
    const int face = axis + ((d0 > 0.0f) ? 3 : 0);
    const int64_t nb = grid_neighbor(G, pi, face);
    if (nb == VPIC_BC_REFLECT_PARTICLES) {
      pu[axis] = -pu[axis];
      md[axis] = -md[axis];
      continue;
    }
    if (nb < G.rangel || nb > G.rangeh) {
      pi = 8 * pi + face;
      ret = 1;
      break;
    }
    pi = (int)(nb - G.rangel);
    pd[axis] = -d0;
  }
  pp[0] = make_float4(pd[0], pd[1], pd[2], __int_as_float(pi));
  pp[1] = make_float4(pu[0], pu[1], pu[2], B.w);
  // the voxel a sort would key this particle by (-1 for a returned mover,
  // whose index the caller rewrites)
  if (vkey) *vkey = ret ? -1 : pi;
  mv.dispx = md[0];
  mv.dispy = md[1];
  mv.dispz = md[2];
  return ret;
}

// Voxel coordinates of v (0..n+1 per axis), with the fast_div multipliers
// of nx+2 and ny+2 when they are exact for this grid (m == 0: divide).
struct VoxDiv {
  unsigned mx, my;
  __device__ __forceinline__ void coords(const GridDev &G, int v, int &x,
                                         int &y, int &z) const {
    const unsigned sx1 = G.nx + 2, sy1 = G.ny + 2;
    unsigned t, zz;
    if (mx) {
      t = fast_div((unsigned)v, mx);
      zz = fast_div(t, my);
    } else {
      t = (unsigned)v / sx1;
      zz = t / sy1;
    }
    x = v - (int)(t * sx1);
    y = (int)(t - zz * sy1);
    z = (int)zz;
  }
};

// move_p (move_p.cc:215-378) for a particle held in registers: A =
// {dx,dy,dz,i}, B = {ux,uy,uz,w} (momentum already updated), mv = its
// mover.  Same arithmetic as dev_move_p.  The voxel coordinates are tracked
// so that a face inside the domain goes to the adjacent voxel without a
// table lookup: interior entries of the reference's neighbor table are
// always v -/+ stride (size_grid, ops.cc:45-76); the table (or the particle
// boundary codes) is read only on the domain's faces.  Writes the particle
// back to p0[mv.i]; returns 1 for a mover the grid cannot handle.
template <class Sink>
__device__ int dev_move_p_reg(vpic_particle_t *__restrict__ p0, float4 A,
                              const float4 B, vpic_particle_mover_t &mv,
                              const Sink &sink, const GridDev &G,
                              const VoxDiv &VD, float qsp, int &vout) {
  float pd[3] = {A.x, A.y, A.z};
  int pi = __float_as_int(A.w);
  float pu[3] = {B.x, B.y, B.z};
  float md[3] = {mv.dispx, mv.dispy, mv.dispz};
  const float q = qsp * B.w;
  int c[3];
  VD.coords(G, pi, c[0], c[1], c[2]);
  const int nmax[3] = {G.nx, G.ny, G.nz};
  const int stride[3] = {1, G.sy, G.sz};
  int ret = 0;
  for (;;) {
    float s_mid[3] = {pd[0], pd[1], pd[2]};
    float s_disp[3] = {md[0], md[1], md[2]};
    float s_dir[3], v[3];
#pragma unroll
    for (int k = 0; k < 3; k++) {
      s_dir[k] = (s_disp[k] > 0.0f) ? 1.0f : -1.0f;
      v[k] = (s_disp[k] == 0.0f) ? 3.4e38f : (s_dir[k] - s_mid[k]) / s_disp[k];
    }
    float v3 = 2.0f;
    int axis = 3;
    if (v[0] < v3) v3 = v[0], axis = 0;
    if (v[1] < v3) v3 = v[1], axis = 1;
    if (v[2] < v3) v3 = v[2], axis = 2;
    v3 *= 0.5f;
#pragma unroll
    for (int k = 0; k < 3; k++) {
      s_disp[k] *= v3;
      s_mid[k] += s_disp[k];
    }
    // move_p.cc:277: promoted to double by the 1.0/3.0 constant.
    const float v5 =
        (float)((double)(q * s_disp[0] * s_disp[1] * s_disp[2]) * (1.0 / 3.0));
    float j[12];
    streak4(j + 0, q * s_disp[0], s_mid[1], s_mid[2], v5);
    streak4(j + 4, q * s_disp[1], s_mid[2], s_mid[0], v5);
    streak4(j + 8, q * s_disp[2], s_mid[0], s_mid[1], v5);
    sink.deposit(pi, j);
#pragma unroll
    for (int k = 0; k < 3; k++) {
      md[k] -= s_disp[k];
      pd[k] += s_disp[k] + s_disp[k];
    }
    if (axis == 3) break;
    const float d0 = s_dir[axis];
    pd[axis] = d0;
    const bool up = d0 > 0.0f;
    const int face = axis + (up ? 3 : 0);
    int ca = axis == 0 ? c[0] : axis == 1 ? c[1] : c[2];
    const int na = axis == 0 ? nmax[0] : axis == 1 ? nmax[1] : nmax[2];
    const int st = axis == 0 ? stride[0] : axis == 1 ? stride[1] : stride[2];
    if (up ? ca < na : ca > 1) {
      // interior face: the adjacent voxel
      pi += up ? st : -st;
      ca += up ? 1 : -1;
      if (axis == 0) c[0] = ca;
      else if (axis == 1) c[1] = ca;
      else c[2] = ca;
      pd[axis] = -d0;
      continue;
    }
    const int64_t nb = grid_neighbor(G, pi, face);
    if (nb == VPIC_BC_REFLECT_PARTICLES) {
      pu[axis] = -pu[axis];
      md[axis] = -md[axis];
      continue;
    }
    if (nb < G.rangel || nb > G.rangeh) {
      pi = 8 * pi + face;
      ret = 1;
      break;
    }
    pi = (int)(nb - G.rangel);
    VD.coords(G, pi, c[0], c[1], c[2]);
    pd[axis] = -d0;
  }
  float4 *pp = reinterpret_cast<float4 *>(p0 + mv.i);
  pp[0] = make_float4(pd[0], pd[1], pd[2], __int_as_float(pi));
  pp[1] = make_float4(pu[0], pu[1], pu[2], B.w);
  mv.dispx = md[0];
  mv.dispy = md[1];
  mv.dispz = md[2];
  vout = pi;
  return ret;
}

// Counters: [0] crossers listed, [1] movers returned, [2] movers lost for
// lack of list space, [3] particles with an out-of-range voxel index.
__global__ void k_clear_counters(int *c) {
  if (threadIdx.x < 4) c[threadIdx.x] = 0;
}

__device__ __forceinline__ void push_returned_mover(
    vpic_particle_t *p, const vpic_particle_mover_t &mv,
    vpic_particle_mover_t *rlist, int rcap, int *cnt) {
  int s = atomicAdd(cnt + 1, 1);
  if (s < rcap) {
    rlist[s] = mv;
  } else {
    atomicAdd(cnt + 2, 1);
    int *pi = reinterpret_cast<int *>(&p[mv.i].i);
    *pi = *pi >> 3;
  }
}

// --------------------------------------------------------------------------
// The per-particle push (advance_p_pipeline.cc:91-211), split into the load
// (so the next particle's load can be issued early) and the compute.  The
// compute stores the new momentum (and the new position when the particle
// stays in its cell) and returns the 12 current contributions (zero for
// crossers) and, for a crosser, its half displacement (the mover of
// :215-219).
// --------------------------------------------------------------------------
struct PushResult {
  bool active, cross;
  int voxel;
  float j[12];
  float dispx, dispy, dispz;
  float4 Bnew;  // updated momentum and weight (used for crossers)
};

// Slots without a particle read particle 0 (every launch has np >= 1) and
// are masked by `valid` in push_compute, so the load is branch-free.
__device__ __forceinline__ void load_particle(const vpic_particle_t *p, int n,
                                              bool valid, float4 &A,
                                              float4 &B) {
  const float4 *pp = reinterpret_cast<const float4 *>(p + (valid ? n : 0));
  A = pp[0];
  B = pp[1];
}

// The 18 interpolation coefficients of one voxel (advance_p_pipeline.cc:96-
// 109 reads them field by field; here as four float4 and a float2).
struct Interp {
  float4 e0, e1, e2, b0;
  float2 b1;
};

__device__ __forceinline__ void load_interp(
    const vpic_interpolator_t *__restrict__ f0, int ii, Interp &F) {
  const float4 *fp = reinterpret_cast<const float4 *>(f0 + ii);
  F.e0 = fp[0];
  F.e1 = fp[1];
  F.e2 = fp[2];
  F.b0 = fp[3];
  F.b1 = reinterpret_cast<const float2 *>(fp + 4)[0];
}

// A voxel index outside the grid (e.g. a mover left pending as
// 8*voxel+face) is never dereferenced: the particle is left untouched and
// counted, and the call reports an error.
__device__ __forceinline__ int push_voxel(const float4 A, bool valid_slot,
                                          int nv, int *__restrict__ nbad,
                                          bool &active) {
  const int iraw = __float_as_int(A.w);
  active = valid_slot && (unsigned)iraw < (unsigned)nv;
  if (valid_slot && !active) atomicAdd(nbad, 1);
  return active ? iraw : 0;
}

// STORE_CROSS: also store a crosser's new momentum (the position is left
// for move_p); the tiled push keeps crossers in registers instead.
template <bool STORE_CROSS = true>
__device__ __forceinline__ void push_compute_f(
    const float4 A, const float4 B, bool active, int ii,
    vpic_particle_t *__restrict__ p, int n, const Interp &F,
    const PushCoef &C, PushResult &R) {
  const float one = 1.0f, one_third = 1.0 / 3.0, two_fifteenths = 2.0 / 15.0;
  float dx = A.x, dy = A.y, dz = A.z;

  // Interpolate E and B (advance_p_pipeline.cc:96-109).
  const float4 fe0 = F.e0, fe1 = F.e1, fe2 = F.e2, fb0 = F.b0;
  const float2 fb1 = F.b1;
  const float hax =
      C.qdt_2mc * ((fe0.x + dy * fe0.y) + dz * (fe0.z + dy * fe0.w));
  const float hay =
      C.qdt_2mc * ((fe1.x + dz * fe1.y) + dx * (fe1.z + dz * fe1.w));
  const float haz =
      C.qdt_2mc * ((fe2.x + dx * fe2.y) + dy * (fe2.z + dx * fe2.w));
  const float cbx = fb0.x + dx * fb0.y;
  const float cby = fb0.z + dy * fb0.w;
  const float cbz = fb1.x + dz * fb1.y;

  float ux = B.x, uy = B.y, uz = B.z, q = B.w;
  float v0, v1, v2, v3, v4, v5;

  // Boris rotation (advance_p_pipeline.cc:116-143).
  ux += hax;
  uy += hay;
  uz += haz;
  v0 = C.qdt_2mc / sqrtf(one + (ux * ux + (uy * uy + uz * uz)));
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
  const float4 Bnew = make_float4(ux, uy, uz, q);

  // Normalised displacement (advance_p_pipeline.cc:145-162).
  v0 = one / sqrtf(one + (ux * ux + (uy * uy + uz * uz)));
  ux *= C.cdt_dx;
  uy *= C.cdt_dy;
  uz *= C.cdt_dz;
  ux *= v0;
  uy *= v0;
  uz *= v0;
  v0 = dx + ux;
  v1 = dy + uy;
  v2 = dz + uz;
  v3 = v0 + ux;
  v4 = v1 + uy;
  v5 = v2 + uz;
  const bool inbnds = v3 <= one && v4 <= one && v5 <= one && -v3 <= one &&
                      -v4 <= one && -v5 <= one;

  float4 *pp = reinterpret_cast<float4 *>(p + (active ? n : 0));
  R.active = active;
  R.voxel = ii;
  R.cross = active && !inbnds;
  R.dispx = ux;
  R.dispy = uy;
  R.dispz = uz;
  R.Bnew = Bnew;
  if (active && inbnds) {
    q *= C.qsp;
    pp[0] = make_float4(v3, v4, v5, A.w);
    pp[1] = Bnew;
    dx = v0;
    dy = v1;
    dz = v2;
    v5 = q * ux * uy * uz * one_third;
    streak4(R.j + 0, q * ux, dy, dz, v5);
    streak4(R.j + 4, q * uy, dz, dx, v5);
    streak4(R.j + 8, q * uz, dx, dy, v5);
  } else {
#pragma unroll
    for (int k = 0; k < 12; k++) R.j[k] = 0.0f;
    if (STORE_CROSS && active) pp[1] = Bnew;  // position left for move_p
  }
}

// sqrtf(x) for x >= 1, +inf or NaN, correctly rounded: the hardware square
// root (within 1 ulp) followed by the one-ulp correction that sqrtf lowers
// to on this target, without the rescaling of arguments below 2^-96 that
// such x never need (the result is bit-identical to sqrtf).  An infinite
// argument is returned as it is, as sqrtf does.
__device__ __forceinline__ float sqrt_ge1(float x) {
  const float s = __builtin_amdgcn_sqrtf(x);
  const float sm = __int_as_float(__float_as_int(s) - 1);
  const float sp = __int_as_float(__float_as_int(s) + 1);
  float r = (fmaf(-sm, s, x) <= 0.0f) ? sm : s;
  r = (fmaf(-sp, s, x) > 0.0f) ? sp : r;
  return x == __builtin_inff() ? x : r;
}

// The push of push_compute_f with the particle stored unconditionally to
// dst (p + n, or a scratch slot for a lane without a particle): a lane that
// moves nothing writes back what it read.  Keeping the same memory
// operations on every path lets the compiler wait only for the loads it
// needs (stores and prefetches stay in flight across iterations).
__device__ __forceinline__ void push_compute_u(
    const float4 A, const float4 B, bool active, const Interp &F,
    const PushCoef &C, float4 *dst, PushResult &R) {
  const float one = 1.0f, one_third = 1.0 / 3.0, two_fifteenths = 2.0 / 15.0;
  float dx = A.x, dy = A.y, dz = A.z;
  const float4 fe0 = F.e0, fe1 = F.e1, fe2 = F.e2, fb0 = F.b0;
  const float2 fb1 = F.b1;
  const float hax =
      C.qdt_2mc * ((fe0.x + dy * fe0.y) + dz * (fe0.z + dy * fe0.w));
  const float hay =
      C.qdt_2mc * ((fe1.x + dz * fe1.y) + dx * (fe1.z + dz * fe1.w));
  const float haz =
      C.qdt_2mc * ((fe2.x + dx * fe2.y) + dy * (fe2.z + dx * fe2.w));
  const float cbx = fb0.x + dx * fb0.y;
  const float cby = fb0.z + dy * fb0.w;
  const float cbz = fb1.x + dz * fb1.y;
  float ux = B.x, uy = B.y, uz = B.z, q = B.w;
  float v0, v1, v2, v3, v4, v5;
  ux += hax;
  uy += hay;
  uz += haz;
  v0 = C.qdt_2mc / sqrt_ge1(one + (ux * ux + (uy * uy + uz * uz)));
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
  const float4 Bnew = make_float4(ux, uy, uz, q);
  v0 = one / sqrt_ge1(one + (ux * ux + (uy * uy + uz * uz)));
  ux *= C.cdt_dx;
  uy *= C.cdt_dy;
  uz *= C.cdt_dz;
  ux *= v0;
  uy *= v0;
  uz *= v0;
  v0 = dx + ux;
  v1 = dy + uy;
  v2 = dz + uz;
  v3 = v0 + ux;
  v4 = v1 + uy;
  v5 = v2 + uz;
  const bool inbnds = v3 <= one && v4 <= one && v5 <= one && -v3 <= one &&
                      -v4 <= one && -v5 <= one;
  const bool stay = active && inbnds;
  R.active = active;
  R.cross = active && !inbnds;
  R.dispx = ux;
  R.dispy = uy;
  R.dispz = uz;
  R.Bnew = Bnew;
  // advance_p_pipeline.cc:172-208 for a particle that stays in its cell
  const float4 Anew = make_float4(v3, v4, v5, A.w);
  q *= C.qsp;
  const float v6 = q * ux * uy * uz * one_third;
  // R.j is meaningful only for a particle that stays (the caller adds it
  // under that condition)
  streak4(R.j + 0, q * ux, v1, v2, v6);
  streak4(R.j + 4, q * uy, v2, v0, v6);
  streak4(R.j + 8, q * uz, v0, v1, v6);
  // per-component selects (a select of whole float4 values is lowered
  // through private memory)
  dst[0] = make_float4(stay ? Anew.x : A.x, stay ? Anew.y : A.y,
                       stay ? Anew.z : A.z, A.w);
  dst[1] = make_float4(active ? Bnew.x : B.x, active ? Bnew.y : B.y,
                       active ? Bnew.z : B.z, B.w);
}

__device__ __forceinline__ void push_compute(
    const float4 A, const float4 B, bool valid_slot,
    vpic_particle_t *__restrict__ p, int n,
    const vpic_interpolator_t *__restrict__ f0, const PushCoef &C, int nv,
    int *__restrict__ nbad, PushResult &R) {
  bool active;
  const int ii = push_voxel(A, valid_slot, nv, nbad, active);
  Interp F;
  load_interp(f0, ii, F);
  push_compute_f(A, B, active, ii, p, n, F, C, R);
}

// --------------------------------------------------------------------------
// Wavefront deposit
// --------------------------------------------------------------------------

// DPP move of a float across lanes (v_mov_b32_dpp); lanes without a valid
// source (outside the row, or rows masked off) read 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp_f32(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(
      0, __float_as_int(v), CTRL, ROW_MASK, 0xf, true));
}

// Wavefronts with at least this many runs of equal voxels deposit lane by
// lane instead of scanning.
__constant__ int c_direct_runs = 65;

// Sum the 12 contributions of the lanes sharing a voxel and deposit them.
// key < 0 marks lanes with nothing to deposit.
template <class Sink>
__device__ __forceinline__ void wave_deposit(int key, float j[12],
                                             const Sink &sink) {
  const int lane = lane_id();
  const int kprev = __shfl_up(key, 1);
  const bool head = (lane == 0) || (kprev != key);
  const uint64_t heads = __ballot(head);
  const int runs = __popcll(heads);
  bool dep;
  if (runs == 1) {
    // one run over the whole wavefront: plain scan, lane 63 holds the sum
#pragma unroll
    for (int k = 0; k < 12; k++) {
      float v = j[k];
      v += dpp_f32<0x111, 0xf>(v);
      v += dpp_f32<0x112, 0xf>(v);
      v += dpp_f32<0x114, 0xf>(v);
      v += dpp_f32<0x118, 0xf>(v);
      v += dpp_f32<0x142, 0xa>(v);
      v += dpp_f32<0x143, 0xc>(v);
      j[k] = v;
    }
    dep = lane == 63;
  } else if (runs >= c_direct_runs) {
    dep = true;
  } else {
    // Run-segmented inclusive scan: Hillis-Steele steps of 1,2,4,8 lanes
    // inside each 16-lane row (DPP row_shr), then the row carries (DPP
    // row_bcast:15 to rows 1 and 3, row_bcast:31 to rows 2 and 3); a step
    // adds only when the source lane lies in the same run.  The last lane of
    // each run ends with the run's sum.
    const uint64_t upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
    const int seg_start = 63 - __clzll(heads & upto);
    const int r = lane & 15;
#define VPIC_SCAN_STEP(CTRL, RMASK, TAKE)            \
  {                                                  \
    const bool take_ = (TAKE);                       \
    _Pragma("unroll") for (int k = 0; k < 12; k++) { \
      const float t_ = dpp_f32<CTRL, RMASK>(j[k]);   \
      j[k] += take_ ? t_ : 0.0f;                     \
    }                                                \
  }
    VPIC_SCAN_STEP(0x111, 0xf, r >= 1 && lane - 1 >= seg_start)
    VPIC_SCAN_STEP(0x112, 0xf, r >= 2 && lane - 2 >= seg_start)
    VPIC_SCAN_STEP(0x114, 0xf, r >= 4 && lane - 4 >= seg_start)
    VPIC_SCAN_STEP(0x118, 0xf, r >= 8 && lane - 8 >= seg_start)
    VPIC_SCAN_STEP(0x142, 0xa,
                   ((lane >> 4) & 1) && seg_start <= (lane & ~15) - 1)
    VPIC_SCAN_STEP(0x143, 0xc, lane >= 32 && seg_start <= 31)
#undef VPIC_SCAN_STEP
    dep = (lane == 63) || (((heads >> lane) >> 1) & 1ull);
  }
  if (dep && key >= 0) sink.deposit(key, j);
}

// --------------------------------------------------------------------------
// Batched boxes: a persistent launch over a work queue of (box, particle
// range) items.  Each box's particles are pushed by the flat kernel's
// arithmetic with the box's coupling constants, deposits run-summed into
// the box's accumulators; a crosser is moved right away by its lane (the
// boxes are periodic: no movers come back).
// --------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_ens_push(
    vpic_particle_t *__restrict__ P, const vpic_interpolator_t *__restrict__ FI,
    float *__restrict__ ACC, const EnsPushBox *__restrict__ boxes,
    const EnsItem *__restrict__ items, int nitems, int *queue,
    int *__restrict__ nbad) {
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(queue, 1);
    __syncthreads();
    const int it = s_item;
    __syncthreads();
    if (it >= nitems) break;
    const EnsItem I = items[it];
    const EnsPushBox b = boxes[I.box];
    vpic_particle_t *p = P + b.poff;
    const vpic_interpolator_t *f0 = FI + b.voff;
    const GlobalSink sink{ACC + 12 * b.voff};
    PushCoef C;
    C.qdt_2mc = b.qdt_2mc;
    C.cdt_dx = b.cdt_dx;
    C.cdt_dy = b.cdt_dy;
    C.cdt_dz = b.cdt_dz;
    C.qsp = b.qsp;
    const int end = I.first + I.count;
    for (int base = I.first; base < end; base += blockDim.x) {
      const int n = base + threadIdx.x;
      const bool valid = n < end;
      float4 A, B;
      load_particle(p, n, valid, A, B);
      PushResult R;
      push_compute(A, B, valid, p, n, f0, C, b.G.nv, nbad, R);
      wave_deposit(R.active ? R.voxel : -1, R.j, sink);
      if (R.cross) {
        vpic_particle_mover_t mv;
        mv.dispx = R.dispx;
        mv.dispy = R.dispy;
        mv.dispz = R.dispz;
        mv.i = n;
        if (dev_move_p(p, mv, sink, b.G, C.qsp)) atomicAdd(nbad + 1, 1);
      }
    }
  }
}

// --------------------------------------------------------------------------
// Flat kernel: particles [n0, np), any order, grid-stride over whole
// wavefronts.  n0 = *start clamped to [0,np] (the end of the sort-time
// partition when the tiled kernel ran first), or 0 when start is NULL.
// --------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_advance_p(
    vpic_particle_t *__restrict__ p, const int *__restrict__ start, int np,
    const vpic_interpolator_t *__restrict__ f0, float *__restrict__ acc,
    vpic_particle_mover_t *__restrict__ xlist, int xcap,
    vpic_particle_mover_t *__restrict__ rlist, int *__restrict__ cnt,
    PushCoef C, GridDev G, int *__restrict__ vkeys) {
  const int n0 = start ? min(max(*start, 0), np) : 0;
  const int lane = lane_id();
  const GlobalSink sink{acc};
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  int64_t nn = n0 + (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  float4 A, B;
  load_particle(p, (int)nn, nn < np, A, B);
  for (; nn - lane < np; nn += stride) {
    const int n = (int)nn;
    const bool valid = nn < np;
    float4 An, Bn;  // prefetch of the next particle
    load_particle(p, (int)(nn + stride), nn + stride < np, An, Bn);
    PushResult R;
    push_compute(A, B, valid, p, n, f0, C, G.nv, cnt + 3, R);
    wave_deposit(R.active ? R.voxel : -1, R.j, sink);
    // a particle that stays in its cell keeps its voxel index
    if (vkeys && valid && !R.cross) vkeys[n] = __float_as_int(A.w);

    // Cell crossers: compact into the crosser list (ballot + one atomic per
    // wavefront); the rare overflow of the list is moved right here.
    const uint64_t cm = __ballot(R.cross);
    if (cm) {
      const int leader = __ffsll((unsigned long long)cm) - 1;
      int base = 0;
      if (lane == leader) base = atomicAdd(cnt, __popcll(cm));
      base = __shfl(base, leader);
      if (R.cross) {
        const int slot = base + __popcll(cm & ((1ull << lane) - 1ull));
        vpic_particle_mover_t mv;
        mv.dispx = R.dispx;
        mv.dispy = R.dispy;
        mv.dispz = R.dispz;
        mv.i = n;
        if (slot < xcap) {
          xlist[slot] = mv;
        } else if (dev_move_p(p, mv, sink, G, C.qsp,
                              vkeys ? vkeys + n : nullptr)) {
          push_returned_mover(p, mv, rlist, xcap, cnt);
        }
      }
    }
    A = An;
    B = Bn;
  }
}

__global__ __launch_bounds__(256) void k_move_p_list(
    vpic_particle_t *__restrict__ p, const vpic_particle_mover_t *xlist,
    int xcap, vpic_particle_mover_t *__restrict__ rlist,
    float *__restrict__ acc, int *__restrict__ cnt, float qsp, GridDev G,
    int *__restrict__ vkeys) {
  const int n = min(cnt[0], xcap);
  const GlobalSink sink{acc};
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += gridDim.x * blockDim.x) {
    vpic_particle_mover_t mv = xlist[k];
    if (dev_move_p(p, mv, sink, G, qsp, vkeys ? vkeys + mv.i : nullptr))
      push_returned_mover(p, mv, rlist, xcap, cnt);
  }
}

// --------------------------------------------------------------------------
// Tiled kernel
// --------------------------------------------------------------------------
struct TileGeom {
  int tx, ty, tz;     // tile extent in cells
  int ntx, nty, ntz;  // tiles per axis
  unsigned mx, my;    // fast_div multipliers for nx+2, ny+2 (0: divide)
  int stage_win;      // coefficients staged for the tile and a one-cell
                      // ring (1) or the tile only (0)
  int halo;           // window cells beyond the tile on every side (1 or 2)
  int hx, hy, hz;     // the halo per axis: `halo`, or 0 along an axis of
                      // one cell (no particle leaves it, so no deposit
                      // lands beside it)
};

#ifndef VPIC_TILE_BS
#define VPIC_TILE_BS 1024
#endif
#ifndef VPIC_PUSH_DEPTH
#define VPIC_PUSH_DEPTH 1
#endif
// A box of a batched ensemble as the tiled push sees it: a single grid at
// offsets of the packed arrays, with its partition (box-relative, nv + 1
// entries), window slots and out-of-window sub-buckets.
struct EnsTileBox {
  vpic_particle_t *p;
  const int *part;
  const vpic_interpolator_t *f0;
  float *acc;
  float *wbuf;
  OowBuckets OB;
  PushCoef C;
  GridDev G;
  TileGeom T;
  int np, wstride;
};

struct EnsTileItem {
  int box, tile;
};

// One species of a tiled push: its particles, sort-time partition,
// couplings, returned-mover list and counters, and (when not null) the
// voxel keys to write for a following sort_p.
struct TileSpecies {
  vpic_particle_t *p;
  const int *part;
  vpic_particle_mover_t *rlist;
  int *cnt;
  int *vkeys;
  PushCoef C;
  int np, rcap;
};

// The species one launch of the tiled push advances into the same windows
// (NSP of them, fixed per kernel instance).
constexpr int TILE_MAXSP = 2;
struct TileSpeciesSet {
  TileSpecies s[TILE_MAXSP];
};

constexpr int TILE_BS = VPIC_TILE_BS;     // threads per workgroup
constexpr int TILE_WAVES = TILE_BS / 64;
// waves per SIMD the launch bounds ask for (one workgroup per CU)
constexpr int TILE_MIN_WAVES = TILE_BS / 256;
// slots a lane loads ahead of the one it pushes (1 to 3)
constexpr int PUSH_DEPTH = VPIC_PUSH_DEPTH;
static_assert(PUSH_DEPTH >= 1 && PUSH_DEPTH <= 3, "prefetch depth 1..3");
static_assert(TILE_BS % 256 == 0, "whole waves on every SIMD");
constexpr int TILE_MAXROWS = 256;         // ty * tz
constexpr int TILE_MAXCELLS = 512;        // tx * ty * tz
// The dynamic LDS of the tiled push: the window (96 B per window cell)
// followed by the staged interpolation coefficients (72 B per cell) of the
// whole window when both fit here, else of the tile's cells only.
constexpr int TILE_DYN_BYTES = 140 * 1024;
constexpr int WAVE_Q = 64;                // crosser queue of one wavefront

// Walks tile slots in increasing order: particle index of slot k (rows of
// the sort-time partition), with the current row's slot range and offset
// kept in registers so that LDS is read only when a row ends.
struct RowWalk {
  int r, end, base;  // row, first slot past it, particle index - slot
  __device__ __forceinline__ void start(const int *row_start,
                                        const int *row_pref) {
    r = 0;
    end = row_pref[1];
    base = row_start[0];
  }
  __device__ __forceinline__ int particle(const int *row_start,
                                          const int *row_pref, int nrow,
                                          int k) {
    while (k >= end && r + 1 < nrow) {
      r++;
      end = row_pref[r + 1];
      base = row_start[r] - row_pref[r];
    }
    return base + k;
  }
};

// Order this wavefront's LDS traffic: a wavefront's LDS operations complete
// in issue order, so waiting for them (and keeping the compiler from moving
// memory operations across) makes one lane's queue writes visible to the
// other lanes' reads.
__device__ __forceinline__ void wave_lds_sync() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_s_waitcnt(0xc07f);  // lgkmcnt(0)
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// Move the first `cnt` crossers of a wavefront's queue, one per lane.  The
// push stored each crosser's new momentum and left its position; the
// particle is read back (from cache) and walked in registers.
// Tile binning.  The tile that pushes a particle records, per particle,
// which of the 27 tiles around it holds the particle's voxel after the push
// (offset code (dx+1) + 3(dy+1) + 9(dz+1); 13 for its own tile and for a
// voxel further away), and counts the codes.  The next push can then give
// every tile the particles now in its cells (lists built by k_bin_scatter),
// so deposits stay inside the tile's window however far the particles have
// drifted since the sort.  Which tile pushes a particle never changes the
// result: a deposit outside the window takes the global path.
struct TileBox {
  int x0, x1, y0, y1, z0, z1;  // the tile's cells
  int lx, hx, ly, hy, lz, hz;  // cells of the tile and of its neighbours
  __device__ __forceinline__ int code(int x, int y, int z) const {
    if (x < lx || x > hx || y < ly || y > hy || z < lz || z > hz) return 13;
    const int dx = x < x0 ? 0 : x > x1 ? 2 : 1;
    const int dy = y < y0 ? 0 : y > y1 ? 2 : 1;
    const int dz = z < z0 ? 0 : z > z1 ? 2 : 1;
    return dx + 3 * dy + 9 * dz;
  }
};

// Binning control words (device ints):
//   [0] codes valid: the last tiled push wrote codes and counts
//   [2] [3] the covered range [part[0], part[nv]) of that push
//   [4] its deposits outside the window
//   [5] this push takes its particles from the lists
//   [6] the last tiled push took its particles from the lists
//   [7] lists stay on until the state is reset
//   [8] the lists failed a consistency check (diagnostic)
//   [9] this push records codes and counts
constexpr int BIN_CTL = 16;

struct BinArgs {
  const int *ctl;
  const int *list;    // particle indices grouped by tile
  const int *lstart;  // ntiles + 1 offsets into list
  unsigned char *key; // per particle offset code (written); null: no binning
  int *counts;        // per tile 27 code counts (written)
  int *oow;           // out-of-window deposits (accumulated)
};

struct BinSink {
  unsigned char *key;  // null: no binning
  int *cnt;            // LDS code counts
  TileBox box;
  int *vkeys;          // per particle voxel index after the push, for a
                       // following sort_p (written); null: not asked for
};

__device__ __forceinline__ void drain_wave_queue(
    const vpic_particle_mover_t *q, int cnt, int lane,
    vpic_particle_t *__restrict__ p, const WindowSink &W, const GridDev &G,
    const VoxDiv &VD, float qsp, vpic_particle_mover_t *__restrict__ rlist,
    int rcap, int *__restrict__ cnt_g, const BinSink &BS) {
  wave_lds_sync();
  vpic_particle_mover_t mv;
  if (lane < cnt) mv = q[lane];
  wave_lds_sync();
  if (lane < cnt) {
    const int n = mv.i;
    const float4 *pp = reinterpret_cast<const float4 *>(p + n);
    const float4 A = pp[0], B = pp[1];
    int vout;
    const int ret = dev_move_p_reg(p, A, B, mv, W, G, VD, qsp, vout);
    if (ret) push_returned_mover(p, mv, rlist, rcap, cnt_g);
    if (BS.vkeys) BS.vkeys[n] = ret ? -1 : vout;
    if (BS.key) {
      int o = 13;
      if (!ret) {
        int x, y, z;
        VD.coords(G, vout, x, y, z);
        o = BS.box.code(x, y, z);
      }
      BS.key[n] = (unsigned char)o;
      atomicAdd(BS.cnt + o, 1);
    }
  }
}

// Tiled push: one workgroup per tile of cells.  Its particles are either the
// rows of the sort-time partition (sorted by voxel) or, when binning is on,
// the tile's list (the particles the previous push left in its cells).
// They are split into one contiguous run per group of GS lanes; the GS
// lanes of a group take the run's particles in turn, so a group reads whole
// cache lines.  A lane sums the current of its particles in registers while
// they stay in one voxel and adds the sum to the LDS window when the voxel
// changes; the interpolation coefficients of the last voxel are kept the
// same way.  The window's fixed-point scale is set from the largest |q w| of
// the first particles each lane pushes.  Cell crossers go to a queue of
// their wavefront, drained 64 at a time by the wavefront (no workgroup
// barrier in the particle loop).  The finished window is written, as
// floats, to this tile's slot of `wbuf`; k_merge_windows then adds the
// windows into the accumulator, so no tile needs atomics for the cells it
// shares.
// (launch bounds: at least 4 waves per SIMD, i.e. one 1024-lane workgroup
// per CU, which caps the kernel at 128 VGPRs)
//
// MODE is the binning state the instance is compiled for: 0 no codes, 1
// codes recorded (slots from the partition rows), 2 codes recorded and slots
// from the tile lists.  The state lives on the device (control words set by
// k_bin_decide), so a binning push launches all three instances and the two
// that do not match return at once; compiling each state separately keeps
// the unused paths' values out of the scalar registers of the particle loop.
//
// NSP species (each with its own particles, partition, couplings and
// mover lists) are pushed in turn by every tile into the same window, so
// the species of one step share one launch, one staging of the
// coefficients and one merge.
//
// The body is a device function of the tile index, so that a batched
// launch over many grids (k_ens_push_tiled) runs the same code per (grid,
// tile) work item.
template <int GS, int MODE, int NSP>
__device__ __forceinline__ void tiled_push(
    int tile, const TileSpecies *SP, const vpic_interpolator_t *__restrict__ f0,
    float *__restrict__ acc, float *__restrict__ wbuf, int wstride,
    vpic_particle_t *__restrict__ trash, const GridDev &G, const TileGeom &T,
    const BinArgs &Bn, const OowBuckets &OB) {
  static_assert(NSP >= 1 && NSP <= TILE_MAXSP, "species per launch");
  static_assert(NSP == 1 || MODE == 0, "tile lists push one species");
  extern __shared__ unsigned long long lds_win[];  // 12 * ncell
  __shared__ vpic_particle_mover_t wq[TILE_WAVES][WAVE_Q];
  __shared__ int row_start[TILE_MAXROWS];
  __shared__ int row_pref[TILE_MAXROWS + 1];
  __shared__ unsigned qw_bits;  // largest |q w| of the first round (bits)
  __shared__ int bin_cnt[27];
  __shared__ int oow_cnt;
  __shared__ int oow_dcnt[27];
  __shared__ unsigned oow_dmax[27];

  const int t = threadIdx.x;
  const int lane = lane_id();
  int b = tile;
  const int btx = b % T.ntx;
  b /= T.ntx;
  const int bty = b % T.nty;
  const int btz = b / T.nty;
  const int x0 = 1 + btx * T.tx, y0 = 1 + bty * T.ty, z0 = 1 + btz * T.tz;
  const int x1 = min(x0 + T.tx - 1, G.nx), y1 = min(y0 + T.ty - 1, G.ny),
            z1 = min(z0 + T.tz - 1, G.nz);
  const int rows_y = y1 - y0 + 1;
  const int nrow = rows_y * (z1 - z0 + 1);
  {
    const bool dkeyed =
        Bn.key != nullptr && __builtin_amdgcn_readfirstlane(Bn.ctl[9]) != 0;
    const bool dlisted =
        dkeyed && __builtin_amdgcn_readfirstlane(Bn.ctl[5]) != 0;
    if ((dlisted ? 2 : dkeyed ? 1 : 0) != MODE) return;
  }
  constexpr bool keyed = MODE >= 1, listed = MODE == 2;

  WindowSink W;
  W.lds = (lds_u64 *)lds_win;
  W.acc = acc;
  W.wx0 = x0 - T.hx;
  W.wy0 = y0 - T.hy;
  W.wz0 = z0 - T.hz;
  W.wnx = x1 - x0 + 1 + 2 * T.hx;
  W.wny = y1 - y0 + 1 + 2 * T.hy;
  W.wnz = z1 - z0 + 1 + 2 * T.hz;
  W.sx1 = G.nx + 2;
  W.sy1 = G.ny + 2;
  W.mx = T.mx;
  W.my = T.my;
  W.oow = &oow_cnt;
  W.ob = OB;
  W.dcnt = oow_dcnt;
  W.dmax = oow_dmax;
  W.tx = T.tx;
  W.ty = T.ty;
  W.tz = T.tz;
  W.ntx = T.ntx;
  W.nty = T.nty;
  W.ntz = T.ntz;
  W.btx = btx;
  W.bty = bty;
  W.btz = btz;
  W.self = tile;
  const int ncell = W.wnx * W.wny * W.wnz;
  W.ncell = ncell;
  VoxDiv VD;
  VD.mx = T.mx;
  VD.my = T.my;
  BinSink BS;
  BS.key = keyed ? Bn.key : nullptr;
  BS.cnt = bin_cnt;
  BS.vkeys = nullptr;
  TileBox &TB = BS.box;
  TB.x0 = x0;
  TB.x1 = x1;
  TB.y0 = y0;
  TB.y1 = y1;
  TB.z0 = z0;
  TB.z1 = z1;
  TB.lx = btx > 0 ? x0 - T.tx : x0;
  TB.hx = btx + 1 < T.ntx ? min(x1 + T.tx, G.nx) : x1;
  TB.ly = bty > 0 ? y0 - T.ty : y0;
  TB.hy = bty + 1 < T.nty ? min(y1 + T.ty, G.ny) : y1;
  TB.lz = btz > 0 ? z0 - T.tz : z0;
  TB.hz = btz + 1 < T.ntz ? min(z1 + T.tz, G.nz) : z1;

  for (int k = t; k < 12 * ncell; k += TILE_BS) lds_win[k] = 0ull;
  // The window's fixed-point scale (below) samples the first particle of
  // every row of every species; those reads (the row bounds, then the
  // particle's voxel and weight) are issued here, so that the particle
  // reads overlap the staging reads that follow, and are used after the
  // barrier.  The same row bounds are the first species' rows of the
  // sort-time partition (clamped to [0, np)).  One row per lane.
  static_assert(TILE_MAXROWS <= TILE_BS, "one row per lane");
  int s_vox[NSP], s_a[NSP], s_e[NSP];
  float s_w[NSP];
  bool s_has[NSP];
#pragma unroll
  for (int sp = 0; sp < NSP; sp++) {
    const TileSpecies &S = SP[sp];
    s_has[sp] = false;
    s_vox[sp] = 0;
    s_a[sp] = s_e[sp] = 0;
    s_w[sp] = 0.0f;
    if (t < nrow) {
      const int y = y0 + t % rows_y, z = z0 + t / rows_y;
      int a = S.part[vox(x0, y, z, G.nx, G.ny)];
      int e = S.part[vox(x1 + 1, y, z, G.nx, G.ny)];
      a = min(max(a, 0), S.np);
      e = min(max(e, a), S.np);
      s_a[sp] = a;
      s_e[sp] = e;
      if (sp == 0) {
        row_start[t] = a;
        row_pref[t + 1] = e - a;
      }
      if (a < e) {
        s_vox[sp] = S.p[a].i;
        s_w[sp] = S.p[a].w;
        s_has[sp] = true;
      }
    }
  }
  // Interpolation coefficients staged in LDS (after the window) for the
  // cells of the staging box: the tile and a one-cell ring around it (which
  // stays within the ghost layers) when the launch made room for it, else
  // the tile.  A particle in the box reads its 18 coefficients from LDS,
  // any other from global memory (a dependent global load, which the
  // wavefront waits for, so the box should cover the drifted particles).
  const int swx = T.stage_win ? min(T.hx, 1) : 0,
            swy = T.stage_win ? min(T.hy, 1) : 0,
            swz = T.stage_win ? min(T.hz, 1) : 0;
  const int sx0 = x0 - swx, sy0 = y0 - swy, sz0 = z0 - swz;
  const int snx = x1 - x0 + 1 + 2 * swx, sny = y1 - y0 + 1 + 2 * swy,
            snz = z1 - z0 + 1 + 2 * swz;
  const int sncell = snx * sny * snz;
  // the staging box is the window itself (a halo of at most one cell
  // staged whole): a voxel's staged cell is then its window cell
  const bool stage_is_window = swx == T.hx && swy == T.hy && swz == T.hz;
  float4 *ic_e0 = reinterpret_cast<float4 *>(lds_win + 12 * ncell);
  float4 *ic_e1 = ic_e0 + sncell;
  float4 *ic_e2 = ic_e1 + sncell;
  float4 *ic_b0 = ic_e2 + sncell;
  float2 *ic_b1 = reinterpret_cast<float2 *>(ic_b0 + sncell);
  for (int c = t; c < sncell; c += TILE_BS) {
    const int v = vox(sx0 + c % snx, sy0 + (c / snx) % sny,
                      sz0 + c / (snx * sny), G.nx, G.ny);
    const float4 *fp = reinterpret_cast<const float4 *>(f0 + v);
    ic_e0[c] = fp[0];
    ic_e1[c] = fp[1];
    ic_e2[c] = fp[2];
    ic_b0[c] = fp[3];
    ic_b1[c] = reinterpret_cast<const float2 *>(fp + 4)[0];
  }
  if (t == 0) qw_bits = 0u;
  if (t < 27) bin_cnt[t] = 0;
  if (t == 0) oow_cnt = 0;
  if (t < 27) oow_dcnt[t] = 0;
  if (t < 27) oow_dmax[t] = 0u;
  __syncthreads();
  // Fixed-point scale of the window: 2^k with 2^k * 8|q w|max <= 2^31 for
  // the first particle of every row of every species (read above).  One
  // particle's current is below 4.4|q w|, so its components scale below
  // 2^31 and convert in FP32; a register sum of many of them may exceed
  // 2^31 and converts in FP64 (exact below 2^51: sums of up to 2^20
  // particles; a larger value takes the float path), as does a particle of
  // larger |q w|.  The resolution, 8|q w|max / 2^31 per deposit, is far
  // below FP32 rounding of the cell sums.
#pragma unroll
  for (int sp = 0; sp < NSP; sp++)
    if (s_has[sp] &&
        (unsigned)s_vox[sp] < (unsigned)G.nv)
      atomicMax(&qw_bits, __float_as_uint(fabsf(SP[sp].C.qsp * s_w[sp])));
  __syncthreads();
  {
    const float m = __uint_as_float(qw_bits);
    int e = 0;
    if (m > 0.0f) frexpf(8.0f * m, &e);  // 8|q w|max < 2^e
    W.scale = ldexp(1.0, 31 - e);
    W.fscale = (float)W.scale;
    W.ifscale = (float)(1.0 / W.scale);
  }

  // The species in turn, into the same window.
#pragma unroll
  for (int sp = 0; sp < NSP; sp++) {
    vpic_particle_t *__restrict__ p = SP[sp].p;
    const int np = SP[sp].np;
    vpic_particle_mover_t *__restrict__ rlist = SP[sp].rlist;
    const int rcap = SP[sp].rcap;
    int *__restrict__ cnt = SP[sp].cnt;
    const PushCoef C = SP[sp].C;
    BS.vkeys = SP[sp].vkeys;
    if (sp > 0) {
      // all wavefronts are done with the last species' rows
      __syncthreads();
      // this species' rows, read in the prologue
      if (t < nrow) {
        row_start[t] = s_a[sp];
        row_pref[t + 1] = s_e[sp] - s_a[sp];
      }
      __syncthreads();
    }
    if (t < 64) {
      // inclusive prefix of the row lengths by the first wavefront: four rows
      // per lane (nrow <= TILE_MAXROWS = 256), then the lanes' totals
      static_assert(TILE_MAXROWS <= 256, "four rows per lane");
      int v[4], run = 0;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int r = 4 * t + i;
        run += r < nrow ? row_pref[r + 1] : 0;
        v[i] = run;
      }
      int incl = run;
#pragma unroll
      for (int d = 1; d < 64; d <<= 1) {
        const int o = __shfl_up(incl, d);
        if (t >= d) incl += o;
      }
      const int off = incl - run;
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int r = 4 * t + i;
        if (r < nrow) row_pref[r + 1] = v[i] + off;
      }
      if (t == 0) row_pref[0] = 0;
    }
    __syncthreads();
    // the slots: the partition rows, or this tile's list
    int total = row_pref[nrow];
    const int *L = Bn.list;
    if (listed) {
      int ls = Bn.lstart[tile], le = Bn.lstart[tile + 1];
      ls = min(max(ls, 0), np);
      le = min(max(le, ls), np);
      L += ls;
      total = le - ls;
    }

    // This group's run of slots [g0, g1); lane j of the group takes slots
    // g0 + j, g0 + j + GS, ...
    constexpr int NGROUP = TILE_BS / GS;
    const int per = (total + NGROUP - 1) / NGROUP;  // slots per group
    const int steps = (per + GS - 1) / GS;          // same for every lane
    const int grp = t / GS, j = t % GS;
    const int g0 = min(grp * per, total), g1 = min(g0 + per, total);
    // this wavefront's crosser queue, indexed by a wave-uniform value so
    // that its address is a scalar (kept in a vector register, it was
    // spilled to scratch, and every reload waited for all memory traffic)
    const int wv = __builtin_amdgcn_readfirstlane(t >> 6);
    const vpic_particle_mover_t *qr = wq[wv];
    vpic_particle_mover_t *q = wq[wv];
    int qn = 0;  // entries in this wavefront's queue (wavefront-uniform)

    float jacc[12];  // this lane's current, not yet in the window
#pragma unroll
    for (int k = 0; k < 12; k++) jacc[k] = 0.0f;
    int jkey = -1;   // its voxel (-1: nothing held)
    int jcell = -1;  // jkey's window cell (-1: outside the window)
    Interp F;        // coefficients of voxel fkey
    F.e0 = F.e1 = F.e2 = F.b0 = make_float4(0.f, 0.f, 0.f, 0.f);
    F.b1 = make_float2(0.f, 0.f);
    int fkey = -1;
    int fcell = -1;  // fkey's window cell (-1: outside the window)
    int nbad = 0;  // slots with a voxel index outside the grid

    RowWalk rw;
    rw.start(row_start, row_pref);
    int k = g0 + j;  // the slot being pushed
    // The particle of slot ks (false past the group's run; a list entry
    // outside [0, np) is skipped).  Asked for increasing ks only.
    auto particle_of = [&](int ks, int &nn) -> bool {
      nn = 0;
      if (ks >= g1) return false;
      nn = listed ? L[ks] : rw.particle(row_start, row_pref, nrow, ks);
      const bool ok = (unsigned)nn < (unsigned)np;
      if (!ok) nn = 0;
      return ok;
    };
    // A lane's particles in flight: the one it pushes and the next
    // PUSH_DEPTH - 1 loaded ahead; each slot loads the particle PUSH_DEPTH
    // slots on into the set it frees, so a load is waited for PUSH_DEPTH
    // slots after it is issued.
    struct PSlot {
      float4 A, B;
      int n;
      bool ok;
    };
    PSlot P0, P1, P2, P3;
    P0.ok = particle_of(k, P0.n);
    load_particle(p, P0.n, P0.ok, P0.A, P0.B);
    if (PUSH_DEPTH >= 2) {
      P1.ok = particle_of(k + GS, P1.n);
      load_particle(p, P1.n, P1.ok, P1.A, P1.B);
    }
    if (PUSH_DEPTH >= 3) {
      P2.ok = particle_of(k + 2 * GS, P2.n);
      load_particle(p, P2.n, P2.ok, P2.A, P2.B);
    }
    PushCoef Cs = C;  // C with the deposit charge scaled by the window's 2^k
    {
      // The deposits of particles that stay in their cell are formed with the
      // charge times 2^k (exact while the product is a normal float), so the
      // register sums need no scaling when they are flushed.
      Cs.qsp = C.qsp * W.fscale;
      const float aq = fabsf(Cs.qsp);
      W.pre = W.fscale > 0.0f && aq <= 3.0e38f && W.ifscale > 0.0f &&
              (C.qsp == 0.0f || aq >= 1.0e-30f);
      if (!W.pre) Cs.qsp = C.qsp;
    }
    int it = 0;
    auto slot = [&](const PSlot &cur, PSlot &pref) {
      const float4 A = cur.A, B = cur.B;
      const int n = cur.n;
      const bool valid = cur.ok;
      const int iraw = __float_as_int(A.w);
      const bool active = valid && (unsigned)iraw < (unsigned)G.nv;
      nbad += valid && !active;
      const int ii = active ? iraw : 0;
      if (active && ii != fkey) {
        int x, y, z;
        VD.coords(G, ii, x, y, z);
        const int wx = x - W.wx0, wy = y - W.wy0, wz = z - W.wz0;
        const bool inw = (unsigned)wx < (unsigned)W.wnx &&
                         (unsigned)wy < (unsigned)W.wny &&
                         (unsigned)wz < (unsigned)W.wnz;
        // (window and staging extents are far below 2^24: 24-bit products)
        fcell = inw ? wx + (int)__umul24(W.wnx, wy + (int)__umul24(W.wny, wz))
                    : -1;
        bool staged;
        int c;
        if (stage_is_window) {
          staged = inw;
          c = fcell;
        } else {
          const int sx = x - sx0, sy = y - sy0, sz = z - sz0;
          staged = (unsigned)sx < (unsigned)snx &&
                   (unsigned)sy < (unsigned)sny &&
                   (unsigned)sz < (unsigned)snz;
          c = sx + (int)__umul24(snx, sy + (int)__umul24(sny, sz));
        }
        if (staged) {
          F.e0 = ic_e0[c];
          F.e1 = ic_e1[c];
          F.e2 = ic_e2[c];
          F.b0 = ic_b0[c];
          F.b1 = ic_b1[c];
        } else {
          load_interp(f0, ii, F);
        }
        fkey = ii;
      }
      // prefetch this lane's next particle (the compiler places the loads;
      // pinning them right after the coefficient loads measured 0.3%
      // slower)
      const int kn = k + PUSH_DEPTH * GS;
      pref.ok = particle_of(kn, pref.n);
      load_particle(p, pref.n, pref.ok, pref.A, pref.B);

      // every lane stores (an idle lane to its own trash slot), so the
      // number of memory operations per iteration does not depend on the
      // particle and the loop never waits for its own stores
      PushResult R;
      push_compute_u(A, B, active, F, Cs,
                     valid ? reinterpret_cast<float4 *>(p + n)
                           : reinterpret_cast<float4 *>(trash + t),
                     R);
      // a particle that stays in its cell keeps its voxel index (a crosser's
      // is written when it is moved)
      if (BS.vkeys && valid && !R.cross) BS.vkeys[n] = iraw;
      if (keyed) {
        // the tile holding the particle now (a crosser's is set when it is
        // moved)
        int o = 13;
        if (active && !R.cross) {
          int x, y, z;
          VD.coords(G, ii, x, y, z);
          o = TB.code(x, y, z);
        }
        unsigned char *kp = valid ? Bn.key + n
                                  : reinterpret_cast<unsigned char *>(trash) + t;
        *kp = (unsigned char)o;
        const bool counted = valid && !R.cross;
        const uint64_t m13 = __ballot(counted && o == 13);
        if (m13 && lane == __ffsll((unsigned long long)m13) - 1)
          atomicAdd(&bin_cnt[13], __popcll(m13));
        if (counted && o != 13) atomicAdd(&bin_cnt[o], 1);
      }
      if (R.active && !R.cross) {
        if (ii != jkey) {
          if (jkey >= 0) W.flush_scaled(jkey, jcell, jacc);
#pragma unroll
          for (int c = 0; c < 12; c++) jacc[c] = R.j[c];
          jkey = ii;
          jcell = fcell;
        } else {
#pragma unroll
          for (int c = 0; c < 12; c++) jacc[c] += R.j[c];
        }
      }
      // queue the crossers; drain first if they would not fit
      const uint64_t cm = __ballot(R.cross);
      if (cm) {
        const int c = __popcll(cm);
        if (qn + c > WAVE_Q) {
          drain_wave_queue(qr, qn, lane, p, W, G, VD, C.qsp, rlist, rcap, cnt,
                           BS);
          qn = 0;
        }
        if (R.cross) {
          vpic_particle_mover_t mv;
          mv.dispx = R.dispx;
          mv.dispy = R.dispy;
          mv.dispz = R.dispz;
          mv.i = n;
          q[qn + lanes_below(cm)] = mv;
        }
        qn += c;
      }
      k += GS;
      it++;
    };
    // the sets rotate: slot `it` pushes set it % (PUSH_DEPTH + 1) and loads
    // into set (it + PUSH_DEPTH) % (PUSH_DEPTH + 1)
    if (PUSH_DEPTH == 1) {
      while (it + 1 < steps) {
        slot(P0, P1);
        slot(P1, P0);
      }
      if (it < steps) slot(P0, P1);
    } else if (PUSH_DEPTH == 2) {
      while (it + 2 < steps) {
        slot(P0, P2);
        slot(P1, P0);
        slot(P2, P1);
      }
      if (it < steps) slot(P0, P2);
      if (it < steps) slot(P1, P0);
    } else {
      while (it + 3 < steps) {
        slot(P0, P3);
        slot(P1, P0);
        slot(P2, P1);
        slot(P3, P2);
      }
      if (it < steps) slot(P0, P3);
      if (it < steps) slot(P1, P0);
      if (it < steps) slot(P2, P1);
    }
    if (nbad) atomicAdd(cnt + 3, nbad);
    if (jkey >= 0) W.flush_scaled(jkey, jcell, jacc);
    if (qn)
      drain_wave_queue(qr, qn, lane, p, W, G, VD, C.qsp, rlist, rcap, cnt, BS);
  }  // species
  __syncthreads();
  if (keyed && t < 27) Bn.counts[tile * 27 + t] = bin_cnt[t];
  if (Bn.oow && t == 0 && oow_cnt) atomicAdd(Bn.oow, oow_cnt);
  if (OB.cap > 0 && t < 27) {
    OB.cnt[2 * (27 * tile + t)] = oow_dcnt[t];
    OB.cnt[2 * (27 * tile + t) + 1] = (int)oow_dmax[t];
  }
  // The window, as floats, to this tile's slot (component-major, ncell).
  float *wout = wbuf + (int64_t)tile * wstride;
  const double inv = 1.0 / W.scale;
  for (int c = t; c < ncell; c += TILE_BS) {
#pragma unroll
    for (int k = 0; k < 12; k++)
      wout[k * ncell + c] =
          (float)((double)(long long)lds_win[12 * c + k] * inv);
  }
}

template <int GS, int MODE, int NSP>
__global__ __launch_bounds__(TILE_BS, TILE_MIN_WAVES) void k_advance_p_tiled(
    TileSpeciesSet SS, const vpic_interpolator_t *__restrict__ f0,
    float *__restrict__ acc, float *__restrict__ wbuf, int wstride,
    vpic_particle_t *__restrict__ trash, GridDev G, TileGeom T, BinArgs Bn,
    OowBuckets OB) {
  tiled_push<GS, MODE, NSP>(blockIdx.x, SS.s, f0, acc, wbuf, wstride, trash,
                            G, T, Bn, OB);
}

// Add the tiles' windows into the accumulator: voxel v receives the
// window of its own tile and, when it lies within the halo (H cells) of a
// face of that tile, the halo of the neighbouring tiles across that face
// (windows do not wrap).  One workgroup per tile, one thread per voxel of
// the tile, so a workgroup reads its own window and the face layers of its
// neighbours' windows; every
// cell of acc is written by exactly one thread.
//
// The tile's bucket of out-of-window deposits (OowBuckets) is summed first,
// in LDS as 64-bit fixed point with a scale set from its largest component
// and record count (sums stay below 2^50, so the rounding of each record
// is exact to 2^-50 of that bound), and added with the windows.
__device__ __forceinline__ void merge_tile(int tile, float *__restrict__ acc,
                                           const float *__restrict__ wbuf,
                                           int wstride, const GridDev &G,
                                           const TileGeom &T,
                                           const OowBuckets &OB) {
  __shared__ unsigned long long bsum[12 * TILE_MAXCELLS];
  int b = tile;
  const int bx = b % T.ntx;
  b /= T.ntx;
  const int by = b % T.nty;
  const int bz = b / T.nty;
  const int x0 = 1 + bx * T.tx, y0 = 1 + by * T.ty, z0 = 1 + bz * T.tz;
  const int cx = min(T.tx, G.nx - bx * T.tx), cy = min(T.ty, G.ny - by * T.ty),
            cz = min(T.tz, G.nz - bz * T.tz);
  const int ncell = cx * cy * cz;
  const int t = threadIdx.x;
  // the sub-buckets addressed to this tile: code d of source tile
  // (this tile - offset(d)), periodic in the tile indices
  __shared__ int sb_n[27];
  __shared__ int64_t sb_at[27];
  __shared__ float sb_max[27];
  __shared__ int sb_total;
  if (t < 27) {
    int n = 0;
    int64_t at = 0;
    float mx = 0.0f;
    if (OB.cap > 0) {
      const int ox = t % 3 - 1, oy = (t / 3) % 3 - 1, oz = t / 9 - 1;
      const int sx = (bx - ox + T.ntx) % T.ntx, sy = (by - oy + T.nty) % T.nty,
                sz = (bz - oz + T.ntz) % T.ntz;
      const int src = sx + T.ntx * (sy + T.nty * sz);
      at = (int64_t)(27 * src + t) * OB.cap;
      n = min(OB.cnt[2 * (27 * src + t)], OB.cap);
      mx = __uint_as_float((unsigned)OB.cnt[2 * (27 * src + t) + 1]);
    }
    sb_n[t] = n;
    sb_at[t] = at;
    sb_max[t] = n > 0 ? mx : 0.0f;
  }
  __syncthreads();
  if (t == 0) {
    int tot = 0;
    for (int d = 0; d < 27; d++) tot += sb_n[d];
    sb_total = tot;
  }
  __syncthreads();
  const int nrec = sb_total;
  (void)tile;
  double bscale = 0.0;
  if (nrec > 0) {
    // every component of every record is at most the largest magnitude
    // the pushing tiles recorded
    float m = 0.0f;
    for (int d = 0; d < 27; d++) m = fmaxf(m, sb_max[d]);
    for (int k = t; k < 12 * ncell; k += blockDim.x) bsum[k] = 0ull;
    __syncthreads();
    int ex = 0;
    const double bound = (double)m * (double)nrec;
    if (bound > 0.0) frexp(bound, &ex);  // every sum is below 2^ex
    bscale = ldexp(1.0, 50 - ex);
    for (int d = 0; d < 27; d++) {
      const float4 *rec = OB.rec + 4 * sb_at[d];
      for (int r = t; r < sb_n[d]; r += blockDim.x) {
        const float4 a = rec[4 * r], c = rec[4 * r + 1], e = rec[4 * r + 2],
                     f = rec[4 * r + 3];
        const int v = __float_as_int(a.x);
        const int z = v / G.sz, rem = v - z * G.sz, y = rem / G.sy,
                  x = rem - y * G.sy;
        const int lx = x - x0, ly = y - y0, lz = z - z0;
        if ((unsigned)lx >= (unsigned)cx || (unsigned)ly >= (unsigned)cy ||
            (unsigned)lz >= (unsigned)cz)
          continue;  // not this tile's voxel (cannot happen)
        const int cell = lx + cx * (ly + cy * lz);
        const float jv[12] = {a.y, a.z, a.w, c.x, c.y, c.z,
                              c.w, e.x, e.y, e.z, e.w, f.x};
#pragma unroll
        for (int k = 0; k < 12; k++)
          atomicAdd(&bsum[k * ncell + cell],
                    (unsigned long long)to_fixed((double)jv[k] * bscale));
      }
    }
    __syncthreads();
  }
  for (int c = t; c < ncell; c += blockDim.x) {
    const int x = x0 + c % cx, y = y0 + (c / cx) % cy, z = z0 + c / (cx * cy);
    // tiles whose window holds v, per axis: those holding a cell within
    // the halo of v, [tile(v - h), tile(v + h)] clamped to the grid
    const int hx = T.hx, hy = T.hy, hz = T.hz;
    const int tx0 = (max(x - hx, 1) - 1) / T.tx, tx1 = (min(x + hx, G.nx) - 1) / T.tx;
    const int ty0 = (max(y - hy, 1) - 1) / T.ty, ty1 = (min(y + hy, G.ny) - 1) / T.ty;
    const int tz0 = (max(z - hz, 1) - 1) / T.tz, tz1 = (min(z + hz, G.nz) - 1) / T.tz;
    float s[12];
#pragma unroll
    for (int k = 0; k < 12; k++) s[k] = 0.0f;
    for (int tz = tz0; tz <= tz1; tz++)
      for (int ty = ty0; ty <= ty1; ty++)
        for (int tx = tx0; tx <= tx1; tx++) {
          const int wx0 = tx * T.tx, wy0 = ty * T.ty, wz0 = tz * T.tz;
          const int wnx = min(T.tx, G.nx - wx0) + 2 * hx;
          const int wny = min(T.ty, G.ny - wy0) + 2 * hy;
          const int wnz = min(T.tz, G.nz - wz0) + 2 * hz;
          const int nc = wnx * wny * wnz;
          // the window of tile (tx, ty, tz) starts at cell 1 + wx0 - hx
          const int pos = (x - 1 - wx0 + hx) +
                          wnx * ((y - 1 - wy0 + hy) + wny * (z - 1 - wz0 + hz));
          const float *w =
              wbuf + (int64_t)(tx + T.ntx * (ty + T.nty * tz)) * wstride + pos;
#pragma unroll
          for (int k = 0; k < 12; k++) s[k] += w[k * nc];
        }
    if (nrec > 0) {
      const double inv = 1.0 / bscale;
#pragma unroll
      for (int k = 0; k < 12; k++)
        s[k] += (float)((double)(long long)bsum[k * ncell + c] * inv);
    }
    float4 *a4 = reinterpret_cast<float4 *>(
        acc + 12 * (int64_t)vox(x, y, z, G.nx, G.ny));
    float4 v0 = a4[0], v1 = a4[1], v2 = a4[2];
    v0.x += s[0]; v0.y += s[1]; v0.z += s[2]; v0.w += s[3];
    v1.x += s[4]; v1.y += s[5]; v1.z += s[6]; v1.w += s[7];
    v2.x += s[8]; v2.y += s[9]; v2.z += s[10]; v2.w += s[11];
    a4[0] = v0;
    a4[1] = v1;
    a4[2] = v2;
  }
}

__global__ __launch_bounds__(256) void k_merge_windows(
    float *__restrict__ acc, const float *__restrict__ wbuf, int wstride,
    GridDev G, TileGeom T, OowBuckets OB) {
  merge_tile(blockIdx.x, acc, wbuf, wstride, G, T, OB);
}

// --------------------------------------------------------------------------
// Batched boxes, tiled: one work item per (box, tile).  Each box is a
// single-grid tiled push at an offset (its particles sorted by the batched
// sort, its partition box-relative), run by the same device functions as
// k_advance_p_tiled / k_merge_windows.  The boxes are periodic, so no mover
// comes back; cnt[3] counts particles with a voxel off their box.
// --------------------------------------------------------------------------
__global__ __launch_bounds__(TILE_BS, TILE_MIN_WAVES) void k_ens_push_tiled(
    const EnsTileBox *__restrict__ boxes, const EnsTileItem *__restrict__ items,
    int nitems, int *__restrict__ queue, vpic_particle_t *__restrict__ trash,
    vpic_particle_mover_t *__restrict__ rlist, int *__restrict__ cnt) {
  __shared__ int s_item;
  const BinArgs Bn = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(queue, 1);
    __syncthreads();
    const int it = __builtin_amdgcn_readfirstlane(s_item);
    __syncthreads();
    if (it >= nitems) break;
    const EnsTileItem I = items[it];
    const EnsTileBox &B = boxes[__builtin_amdgcn_readfirstlane(I.box)];
    const TileSpecies S = {B.p, B.part, rlist, cnt, nullptr, B.C, B.np, 0};
    tiled_push<16, 0, 1>(__builtin_amdgcn_readfirstlane(I.tile), &S, B.f0,
                         B.acc, B.wbuf, B.wstride, trash, B.G, B.T, Bn, B.OB);
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_ens_merge(
    const EnsTileBox *__restrict__ boxes, const EnsTileItem *__restrict__ items,
    int nitems) {
  const EnsTileItem I = items[blockIdx.x];
  const EnsTileBox &B = boxes[__builtin_amdgcn_readfirstlane(I.box)];
  merge_tile(__builtin_amdgcn_readfirstlane(I.tile), B.acc, B.wbuf, B.wstride,
             B.G, B.T, B.OB);
}

// --------------------------------------------------------------------------
// Tile lists for the next push (from the codes and counts of the last one)
// --------------------------------------------------------------------------

__global__ void k_fill_int32(int *a, int n, int v) {
  for (int i = threadIdx.x; i < n; i += blockDim.x) a[i] = v;
}

__global__ void k_fill_i32_grid(int *a, int n, int v) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += gridDim.x * blockDim.x)
    a[i] = v;
}

__device__ __forceinline__ void covered_range(const int *part, int nv, int np,
                                              int &lo, int &hi) {
  lo = min(max(part[0], 0), np);
  hi = min(max(part[nv], lo), np);
}

// Whether this push takes its particles from the lists: the codes are
// those of the last tiled push over the same covered range, and binning is
// forced (mode 2) or the last push made out-of-window deposits for at least
// 1/20 of its particles (then it stays on until the state is reset; below
// that the lists cost more than the global atomics they save).  Codes are
// recorded from 1/40 on, so that they are ready when the lists are wanted.
__global__ void k_bin_decide(int *ctl, const int *part, int nv, int np,
                             int mode) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int lo, hi;
  covered_range(part, nv, np, lo, hi);
  const bool ok = ctl[0] != 0 && ctl[2] == lo && ctl[3] == hi;
  const int64_t oow = ctl[4], n = hi - lo;
  const bool want =
      ok && mode > 0 && (mode == 2 || ctl[7] != 0 || oow * 20 >= n);
  const bool keys = mode == 2 || want || (mode > 0 && oow * 40 >= n);
  ctl[5] = want ? 1 : 0;
  ctl[7] = want ? 1 : 0;
  ctl[9] = keys ? 1 : 0;
  ctl[4] = 0;
  ctl[8] = 0;
}

// After the push: its codes (if it recorded them) are valid for its
// covered range.
__global__ void k_bin_post(int *ctl, const int *part, int nv, int np) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  int lo, hi;
  covered_range(part, nv, np, lo, hi);
  ctl[0] = ctl[9];
  ctl[2] = lo;
  ctl[3] = hi;
  ctl[6] = ctl[5];
}

__device__ __forceinline__ bool tile_src(const TileGeom &T, int d, int o,
                                         int &s) {
  const int dtx = d % T.ntx, dty = (d / T.ntx) % T.nty,
            dtz = d / (T.ntx * T.nty);
  const int sx = dtx - (o % 3 - 1), sy = dty - ((o / 3) % 3 - 1),
            sz = dtz - (o / 9 - 1);
  if (sx < 0 || sx >= T.ntx || sy < 0 || sy >= T.nty || sz < 0 ||
      sz >= T.ntz)
    return false;
  s = sx + T.ntx * (sy + T.nty * sz);
  return true;
}

// total[d] = particles the last push left in tile d (total[ntiles] = 0).
__global__ __launch_bounds__(256) void k_bin_total(const int *ctl,
                                                   const int *counts,
                                                   int *total, TileGeom T) {
  const int ntiles = T.ntx * T.nty * T.ntz;
  const int d = blockIdx.x * 256 + threadIdx.x;
  if (d > ntiles) return;
  int sum = 0;
  if (d < ntiles && ctl[5]) {
    for (int o = 0; o < 27; o++) {
      int s;
      if (tile_src(T, d, o, s)) sum += counts[s * 27 + o];
    }
  }
  total[d] = sum;
}

// base[s*27+o] = where the particles tile s sends to tile d = s + off(o)
// start in d's list (sources in code order).
__global__ __launch_bounds__(256) void k_bin_base(const int *ctl,
                                                  const int *counts,
                                                  const int *start, int *base,
                                                  TileGeom T) {
  const int ntiles = T.ntx * T.nty * T.ntz;
  const int d = blockIdx.x * 256 + threadIdx.x;
  if (d >= ntiles || !ctl[5]) return;
  int run = start[d];
  for (int o = 0; o < 27; o++) {
    int s;
    if (tile_src(T, d, o, s)) {
      base[s * 27 + o] = run;
      run += counts[s * 27 + o];
    }
  }
}

// One wavefront per source tile s: walk the particles the last push gave to
// s, in the order it took them (its partition rows, or its list), and write
// each index to the list of the tile its code names.  The order within one
// (source, code) pair is kept.  A count that does not match the push's (the
// partition or the particles changed in between) switches the lists off
// for this push.
__global__ __launch_bounds__(256) void k_bin_scatter(
    int *ctl, const int *__restrict__ part, int np,
    const int *__restrict__ lcur, const int *__restrict__ scur,
    const unsigned char *__restrict__ key, const int *__restrict__ counts,
    const int *__restrict__ base, int *__restrict__ lnext, GridDev G,
    TileGeom T) {
  if (!ctl[5]) return;
  const int lane = threadIdx.x & 63;
  const int s = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int ntiles = T.ntx * T.nty * T.ntz;
  if (s >= ntiles) return;
  const int b27 = lane < 27 ? lane : 0;
  int runv = base[s * 27 + b27];
  const int want = runv + (lane < 27 ? counts[s * 27 + b27] : 0);
  bool bad = false;
  const uint64_t lt = (1ull << lane) - 1ull;
  auto emit = [&](int n, bool v) {
    int o = v ? (int)key[n] : 0;
    if (o > 26) {
      bad = true;
      v = false;
    }
    uint64_t rem = __ballot(v);
    while (rem) {
      const int l = __ffsll((unsigned long long)rem) - 1;
      const int bo = __builtin_amdgcn_readlane(o, l);
      const uint64_t m = __ballot(v && o == bo);
      const int r0 = __builtin_amdgcn_readlane(runv, bo);
      if (v && o == bo) {
        const int pos = r0 + __popcll(m & lt);
        if ((unsigned)pos < (unsigned)np)
          lnext[pos] = n;
        else
          bad = true;
      }
      if (lane == bo) runv += __popcll(m);
      rem &= ~m;
    }
  };
  if (ctl[6]) {
    int ls = scur[s], le = scur[s + 1];
    ls = min(max(ls, 0), np);
    le = min(max(le, ls), np);
    for (int e0 = ls; e0 < le; e0 += 64) {
      const int e = e0 + lane;
      const bool v = e < le;
      const int n = v ? lcur[e] : 0;
      emit(n, v && (unsigned)n < (unsigned)np);
    }
  } else {
    const int btx = s % T.ntx, bty = (s / T.ntx) % T.nty,
              btz = s / (T.ntx * T.nty);
    const int x0 = 1 + btx * T.tx, y0 = 1 + bty * T.ty, z0 = 1 + btz * T.tz;
    const int x1 = min(x0 + T.tx - 1, G.nx), y1 = min(y0 + T.ty - 1, G.ny),
              z1 = min(z0 + T.tz - 1, G.nz);
    for (int z = z0; z <= z1; z++)
      for (int y = y0; y <= y1; y++) {
        int rs = part[vox(x0, y, z, G.nx, G.ny)];
        int re = part[vox(x1 + 1, y, z, G.nx, G.ny)];
        rs = min(max(rs, 0), np);
        re = min(max(re, rs), np);
        for (int e0 = rs; e0 < re; e0 += 64) {
          const int e = e0 + lane;
          emit(e, e < re);
        }
      }
  }
  if (__ballot(bad || (lane < 27 && runv != want)) && lane == 0) {
    ctl[5] = 0;
    ctl[8] = 1;
  }
}

// --------------------------------------------------------------------------
// Returned movers in particle order
// --------------------------------------------------------------------------

// Order the returned movers by particle index (bitonic sort of
// (index, slot) pairs in global memory by one workgroup), keep the first
// max_nm and roll the rest back.  dev_nm[0] = nm, dev_nm[1] = n_ignored,
// dev_nm[2] = particles skipped for an out-of-range voxel index.
__global__ __launch_bounds__(1024) void k_order_movers(
    vpic_particle_t *__restrict__ p, const vpic_particle_mover_t *rlist,
    int rcap, int2 *__restrict__ buf, vpic_particle_mover_t *__restrict__ pm,
    int max_nm, const int *__restrict__ cnt, int *__restrict__ dev_nm) {
  const int n = min(cnt[1], rcap);
  const int lost = cnt[2];
  const int t = threadIdx.x, T = blockDim.x;
  if (n == 0) {
    if (t == 0) {
      dev_nm[0] = 0;
      dev_nm[1] = lost;
      dev_nm[2] = cnt[3];
    }
    return;
  }
  int pad = 1;
  while (pad < n) pad <<= 1;
  for (int k = t; k < pad; k += T)
    buf[k] = k < n ? make_int2(rlist[k].i, k) : make_int2(0x7fffffff, k);
  __syncthreads();
  for (int size = 2; size <= pad; size <<= 1) {
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int k = t; k < pad; k += T) {
        const int l = k ^ stride;
        if (l > k) {
          const bool up = (k & size) == 0;
          const int2 a = buf[k], bb = buf[l];
          if ((a.x > bb.x) == up) {
            buf[k] = bb;
            buf[l] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  const int keep = max_nm < 0 ? 0 : min(n, max_nm);
  for (int k = t; k < n; k += T) {
    const vpic_particle_mover_t mv = rlist[buf[k].y];
    if (k < keep) {
      pm[k] = mv;
    } else {
      int *pi = reinterpret_cast<int *>(&p[mv.i].i);
      *pi = *pi >> 3;
    }
  }
  if (t == 0) {
    dev_nm[0] = keep;
    dev_nm[1] = (n - keep) + lost;
    dev_nm[2] = cnt[3];
  }
}

__global__ void k_move_p_serial(vpic_particle_t *p, vpic_particle_mover_t *pm,
                                int nm, float *acc, GridDev G, float q,
                                int *ret) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  const GlobalSink sink{acc};
  for (int k = 0; k < nm; k++) {
    vpic_particle_mover_t mv = pm[k];
    ret[k] = dev_move_p(p, mv, sink, G, q);
    pm[k] = mv;
  }
}

// The injection loop of boundary_p (boundary_p.cc:595-711) for one species,
// on one lane: the injectors last first, each appended to p with a mover
// for its remaining displacement and moved by move_p (depositing into acc);
// the mover stays in use when move_p returns 1.  st = {np, nm, particles
// dropped, movers dropped, injectors with a voxel off the grid}.
__global__ void k_inject_p(vpic_particle_t *p, int max_np,
                           vpic_particle_mover_t *pm, int max_nm,
                           const vpic_particle_injector_t *pi, int n,
                           float *acc, GridDev G, float q, int *st) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  int np = st[0], nm = st[1];
  const GlobalSink sink{acc};
  for (int k = n - 1; k >= 0; k--) {
    const vpic_particle_injector_t c = pi[k];
    if (np >= max_np) {
      st[2]++;
      continue;
    }
    // move_p deposits into the injector's voxel and looks up its faces
    if (c.i < 0 || c.i + G.sz + G.sy + 1 >= G.nv) {
      st[4]++;
      continue;
    }
    float4 *d = reinterpret_cast<float4 *>(p + np);
    d[0] = make_float4(c.dx, c.dy, c.dz, __int_as_float(c.i));
    d[1] = make_float4(c.ux, c.uy, c.uz, c.w);
    const int at = np++;
    if (nm >= max_nm) {
      st[3]++;
      continue;
    }
    vpic_particle_mover_t mv;
    mv.dispx = c.dispx;
    mv.dispy = c.dispy;
    mv.dispz = c.dispz;
    mv.i = at;
    const int r = dev_move_p(p, mv, sink, G, q);
    pm[nm] = mv;
    nm += r;
  }
  st[0] = np;
  st[1] = nm;
}

// Scratch of the push: counters, crosser list, returned-mover list and the
// ordering buffer, one set per device.
// Tile lists of one particle array (keyed by its address).
struct BinState {
  const void *p = nullptr;
  int np = -1, ntiles = 0;
  int64_t used = 0;
  int cur = 0;  // list[cur] / start[cur]: the lists the last push used
  DeviceScratch key, counts, base, total, sums, ctl, list[2], start[2];
};

struct PushScratch {
  // per species of one call: counters, crosser list, returned movers and
  // their ordering buffer
  DeviceScratch counters[TILE_MAXSP], xlist[TILE_MAXSP], rlist[TILE_MAXSP],
      sortbuf[TILE_MAXSP];
  DeviceScratch wbuf, trash, inject;
  DeviceScratch ocnt, orec;  // out-of-window buckets (OowBuckets)
  BinState bins[8];
  int64_t bin_clock = 0;
};

PushScratch &push_scratch() {
  static thread_local PushScratch s[16];
  int dev = 0;
  (void)hipGetDevice(&dev);
  return s[dev & 15];
}

int pow2_at_least(int64_t n) {
  int64_t p = 1;
  while (p < n) p <<= 1;
  return (int)p;
}

// 7x8x7 cells (9x10x9 window cells and their coefficients in LDS): the
// fastest of the shapes timed on the 256x128x128 thermal plasma over whole
// sort intervals (8x8x8, whose coefficients fit for the tile only, took 6%
// longer per push)
TileGeom g_tile = {7, 8, 7, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1};
int g_group = 16;  // lanes sharing one run of particles in the tiled push
int g_bin_mode = 0;  // tile lists: 0 off, 1 when drift warrants, 2 always

BinState &bin_state(PushScratch &S, const void *p) {
  BinState *hit = nullptr, *lru = &S.bins[0];
  for (BinState &b : S.bins) {
    if (b.p == p) hit = &b;
    if (b.used < lru->used) lru = &b;
  }
  if (!hit) {
    hit = lru;
    hit->p = p;
    hit->np = -1;
  }
  hit->used = ++S.bin_clock;
  return *hit;
}

typedef void (*TiledKernel)(TileSpeciesSet, const vpic_interpolator_t *,
                            float *, float *, int, vpic_particle_t *, GridDev,
                            TileGeom, BinArgs, OowBuckets);

int tiled_kernel_id(int gs) {
  int id = 0;
  while ((1 << id) < gs && id < 6) id++;
  return id;  // log2 of the group size (1 .. 64)
}

template <int MODE, int NSP>
TiledKernel tiled_kernel_mode(int gs) {
  switch (tiled_kernel_id(gs)) {
    case 0: return k_advance_p_tiled<1, MODE, NSP>;
    case 1: return k_advance_p_tiled<2, MODE, NSP>;
    case 2: return k_advance_p_tiled<4, MODE, NSP>;
    case 3: return k_advance_p_tiled<8, MODE, NSP>;
    case 4: return k_advance_p_tiled<16, MODE, NSP>;
    case 5: return k_advance_p_tiled<32, MODE, NSP>;
    default: return k_advance_p_tiled<64, MODE, NSP>;
  }
}

// The instance for lane groups of gs, binning state `mode` (one species)
// or nsp species in one launch (no binning).
TiledKernel tiled_kernel(int gs, int mode, int nsp = 1) {
  if (nsp == 2) return tiled_kernel_mode<0, 2>(gs);
  return mode == 0 ? tiled_kernel_mode<0, 1>(gs)
                   : mode == 1 ? tiled_kernel_mode<1, 1>(gs)
                               : tiled_kernel_mode<2, 1>(gs);
}

// ceil(2^32 / d) for fast_div.
unsigned div_magic(unsigned d) {
  return (unsigned)((((uint64_t)1 << 32) + d - 1) / d);
}

// Cells of the LDS window of a tx x ty x tz tile with a halo of h cells.
size_t tile_win_cells(int tx, int ty, int tz, int h) {
  return (size_t)(tx + 2 * h) * (ty + 2 * h) * (tz + 2 * h);
}

// The per-axis halos of T on grid g (0 along an axis of one cell) and the
// cells of its window and of its coefficient staging box (the tile and a
// ring of min(halo, 1) cells).
void tile_axis_halos(TileGeom &T, const vpic_grid_t *g) {
  T.hx = g->nx == 1 ? 0 : T.halo;
  T.hy = g->ny == 1 ? 0 : T.halo;
  T.hz = g->nz == 1 ? 0 : T.halo;
}
size_t tile_window(const TileGeom &T) {
  return (size_t)(T.tx + 2 * T.hx) * (T.ty + 2 * T.hy) * (T.tz + 2 * T.hz);
}
size_t tile_ring(const TileGeom &T) {
  const int rx = T.hx < 1 ? T.hx : 1, ry = T.hy < 1 ? T.hy : 1,
            rz = T.hz < 1 ? T.hz : 1;
  return (size_t)(T.tx + 2 * rx) * (T.ty + 2 * ry) * (T.tz + 2 * rz);
}

// The tile can fit the tiled push: rows and cells within the kernel's
// tables, and its window (96 B per cell) plus at least the tile's own
// staged coefficients (72 B per cell) within the dynamic LDS -- counting
// no halo along an axis of one tile cell, as on a grid one cell thick
// there (a launch on a thicker grid checks the window it needs).
bool tile_fits(int tx, int ty, int tz, int h) {
  if (!(tx >= 1 && ty >= 1 && tz >= 1 && (h == 1 || h == 2) &&
        ty * tz <= TILE_MAXROWS && tx * ty * tz <= TILE_MAXCELLS))
    return false;
  const size_t w = (size_t)(tx + (tx > 1 ? 2 * h : 0)) *
                   (ty + (ty > 1 ? 2 * h : 0)) * (tz + (tz > 1 ? 2 * h : 0));
  return 96 * w + 72 * (size_t)tx * ty * tz <= (size_t)TILE_DYN_BYTES;
}

}  // namespace

// Tile shape and window halo of the tiled push (the window of a tile is
// (tx+2h)(ty+2h)(tz+2h) cells of 96 B in LDS).
int push_set_tile(int tx, int ty, int tz, int halo) {
  if (!tile_fits(tx, ty, tz, halo)) return VPIC_EINVAL;
  g_tile.tx = tx;
  g_tile.ty = ty;
  g_tile.tz = tz;
  g_tile.halo = halo;
  return VPIC_OK;
}

void push_get_tile(int out[4]) {
  out[0] = g_tile.tx;
  out[1] = g_tile.ty;
  out[2] = g_tile.tz;
  out[3] = g_tile.halo;
}

int push_set_group(int gs) {
  if (gs < 1 || gs > 64 || (gs & (gs - 1))) return VPIC_EINVAL;
  g_group = gs;
  return VPIC_OK;
}

int push_set_bin_mode(int mode) {
  if (mode < 0 || mode > 2) return VPIC_EINVAL;
  g_bin_mode = mode;
  // lists resume only from codes written while binning was on
  PushScratch &S = push_scratch();
  for (BinState &b : S.bins) b.np = -1;
  return VPIC_OK;
}

void push_invalidate_bins(const void *p) {
  PushScratch &S = push_scratch();
  for (BinState &b : S.bins)
    if (b.p == p) b.np = -1;
}

int push_set_direct_runs(int runs) {
  if (runs < 2 || runs > 65) return VPIC_EINVAL;
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_direct_runs), &runs,
                                   sizeof(int));
  return e == hipSuccess ? VPIC_OK : vpic_check_hip(e, "hipMemcpyToSymbol");
}

namespace {

// advance_p_pipeline.cc:279-283, in float.
PushCoef push_coef(float q, float m, const vpic_grid_t *g) {
  PushCoef C;
  C.qdt_2mc = (q * g->dt) / (2 * m * g->cvac);
  C.cdt_dx = g->cvac * g->dt * g->rdx;
  C.cdt_dy = g->cvac * g->dt * g->rdy;
  C.cdt_dz = g->cvac * g->dt * g->rdz;
  C.qsp = q;
  return C;
}

// Crossers listed for move_p (the rare excess is moved inline by the
// push kernel) and movers returned to the caller: ~20% of the particles
// per step, at least 65536, at most np.
int crosser_capacity(int np) {
  int64_t xcap64 = np / 5;
  if (xcap64 < 65536) xcap64 = 65536;
  if (xcap64 > np) xcap64 = np < 1 ? 1 : np;
  return (int)xcap64;
}

// The flat push of particles [0, np) (k_advance_p), the listed crossers'
// move_p and the ordering of the returned movers, on scratch the caller
// provides: cnt (4 ints, cleared here by a kernel), xlist and rlist (xcap
// movers each), sortbuf (pow2_at_least(xcap) int2).
int launch_flat_push(vpic_particle_t *p, int np, float q, const PushCoef &C,
                     const GridDev &G, vpic_particle_mover_t *pm, int max_nm,
                     int *dev_nm, vpic_accumulator_t *a,
                     const vpic_interpolator_t *f, int *cnt,
                     vpic_particle_mover_t *xlist,
                     vpic_particle_mover_t *rlist, int2 *sortbuf, int xcap,
                     hipStream_t s) {
  hipLaunchKernelGGL(k_clear_counters, dim3(1), dim3(64), 0, s, cnt);
  VPIC_LAUNCHED("k_clear_counters");
  if (np > 0) {
    int64_t want = ((int64_t)np + 255) / 256;
    const int blocks = (int)(want < 16384 ? want : 16384);
    hipLaunchKernelGGL(k_advance_p, dim3(blocks), dim3(256), 0, s, p,
                       (const int *)nullptr, np, f, (float *)a, xlist, xcap,
                       rlist, cnt, C, G, (int *)nullptr);
    VPIC_LAUNCHED("k_advance_p");
    int mblocks = (xcap + 255) / 256;
    if (mblocks > 8192) mblocks = 8192;
    hipLaunchKernelGGL(k_move_p_list, dim3(mblocks), dim3(256), 0, s, p,
                       xlist, xcap, rlist, (float *)a, cnt, q, G,
                       (int *)nullptr);
    VPIC_LAUNCHED("k_move_p_list");
  }
  hipLaunchKernelGGL(k_order_movers, dim3(1), dim3(1024), 0, s, p, rlist,
                     xcap, sortbuf, pm, max_nm, cnt, dev_nm);
  VPIC_LAUNCHED("k_order_movers");
  return VPIC_OK;
}

}  // namespace

// Scratch of the flat push for callers that cannot allocate
// (vpic_hip_advance_p_noalloc): one array per GPU, held until the process
// ends, laid out as 4 counters (in a 256-B slot), the crosser list and the
// returned-mover list (NOALLOC_XCAP movers each) and the ordering buffer
// (NOALLOC_XCAP int2): 2.5 MiB.
constexpr int NOALLOC_XCAP = 65536;
constexpr size_t NOALLOC_CNT_BYTES = 256;
constexpr size_t NOALLOC_LIST_BYTES =
    sizeof(vpic_particle_mover_t) * (size_t)NOALLOC_XCAP;
constexpr size_t NOALLOC_BYTES = NOALLOC_CNT_BYTES + 2 * NOALLOC_LIST_BYTES +
                                 sizeof(int2) * (size_t)NOALLOC_XCAP;
static __device__ __align__(256) unsigned char scratch[NOALLOC_BYTES];

namespace {
// Device address of `scratch` per device (the host-side name is only a
// stand-in and must never reach a kernel).
constexpr int VPIC_MAX_DEVICES = 16;
std::mutex g_noalloc_mu;
void *g_noalloc_ptr[VPIC_MAX_DEVICES] = {nullptr};
}  // namespace

int push_noalloc_setup() {
  int dev = 0;
  VPIC_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= VPIC_MAX_DEVICES)
    return vpic_set_error(VPIC_EINVAL, "device %d beyond %d", dev,
                          VPIC_MAX_DEVICES);
  std::lock_guard<std::mutex> lk(g_noalloc_mu);
  if (g_noalloc_ptr[dev]) return VPIC_OK;
  void *ptr = nullptr;
  VPIC_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr)
    return vpic_set_error(VPIC_EHIP, "hipGetSymbolAddress(scratch) gave NULL");
  g_noalloc_ptr[dev] = ptr;
  return VPIC_OK;
}

int launch_advance_p_noalloc(vpic_particle_t *p, int np, float q, float m,
                             vpic_particle_mover_t *pm, int max_nm,
                             int *dev_nm, vpic_accumulator_t *a,
                             const vpic_interpolator_t *f,
                             const vpic_grid_t *g, hipStream_t s) {
  if (max_nm > NOALLOC_XCAP)
    return vpic_set_error(VPIC_EINVAL,
                          "advance_p_noalloc: max_nm %d exceeds the %d movers "
                          "its fixed scratch holds", max_nm, NOALLOC_XCAP);
  int rc = push_noalloc_setup();
  if (rc) return rc;
  int dev = 0;
  VPIC_HIP(hipGetDevice(&dev));
  unsigned char *base = (unsigned char *)g_noalloc_ptr[dev];
  int *cnt = (int *)base;
  vpic_particle_mover_t *xlist =
      (vpic_particle_mover_t *)(base + NOALLOC_CNT_BYTES);
  vpic_particle_mover_t *rlist =
      (vpic_particle_mover_t *)(base + NOALLOC_CNT_BYTES + NOALLOC_LIST_BYTES);
  int2 *sortbuf =
      (int2 *)(base + NOALLOC_CNT_BYTES + 2 * NOALLOC_LIST_BYTES);
  int xcap = crosser_capacity(np);
  if (xcap > NOALLOC_XCAP) xcap = NOALLOC_XCAP;
  // a flat push moves particles without recording their tiles
  push_invalidate_bins(p);
  return launch_flat_push(p, np, q, push_coef(q, m, g), make_grid_dev(g), pm,
                          max_nm, dev_nm, a, f, cnt, xlist, rlist, sortbuf,
                          xcap, s);
}

int launch_advance_p(vpic_particle_t *p, int np, float q, float m,
                     vpic_particle_mover_t *pm, int max_nm, int *dev_nm,
                     vpic_accumulator_t *a, const vpic_interpolator_t *f,
                     const vpic_grid_t *g, const int *partition, int *keys,
                     hipStream_t s) {
  const vpic_push_species_t one = {p, np, q, m, pm, max_nm, dev_nm,
                                   partition, keys};
  return launch_advance_p_set(1, &one, a, f, g, s);
}

// advance_p of nsp species into one accumulator.  Species that all have a
// sort-time partition share one tiled launch (the same windows, staged
// coefficients and merge, the species pushed in turn by every tile);
// otherwise each goes on its own.  The tail of each (particles past its
// partition, listed crossers, returned movers) is its own.
int launch_advance_p_set(int nsp, const vpic_push_species_t *sp,
                         vpic_accumulator_t *a, const vpic_interpolator_t *f,
                         const vpic_grid_t *g, hipStream_t s) {
  if (nsp < 1 || nsp > TILE_MAXSP)
    return vpic_set_error(VPIC_EINVAL, "advance_p: %d species in one call "
                          "(1 to %d)", nsp, TILE_MAXSP);
  if (nsp > 1) {
    bool together = g_bin_mode == 0;
    for (int k = 0; k < nsp; k++)
      together = together && sp[k].partition && sp[k].np > 0;
    if (!together) {
      for (int k = 0; k < nsp; k++) {
        const int rc = launch_advance_p_set(1, sp + k, a, f, g, s);
        if (rc) return rc;
      }
      return VPIC_OK;
    }
  }
  PushScratch &S = push_scratch();
  int xcap[TILE_MAXSP], *cnt[TILE_MAXSP];
  vpic_particle_mover_t *xlist[TILE_MAXSP], *rlist[TILE_MAXSP];
  PushCoef C[TILE_MAXSP];
  int64_t np_total = 0;
  int rc;
  for (int k = 0; k < nsp; k++) {
    xcap[k] = crosser_capacity(sp[k].np);
    if ((rc = scratch_reserve(S.counters[k], 64)) ||
        (rc = scratch_reserve(S.xlist[k],
                              sizeof(vpic_particle_mover_t) * xcap[k])) ||
        (rc = scratch_reserve(S.rlist[k],
                              sizeof(vpic_particle_mover_t) * xcap[k])) ||
        (rc = scratch_reserve(S.sortbuf[k],
                              sizeof(int2) * pow2_at_least(xcap[k]))))
      return rc;
    cnt[k] = (int *)S.counters[k].ptr;
    xlist[k] = (vpic_particle_mover_t *)S.xlist[k].ptr;
    rlist[k] = (vpic_particle_mover_t *)S.rlist[k].ptr;
    C[k] = push_coef(sp[k].q, sp[k].m, g);
    np_total += sp[k].np;
    hipLaunchKernelGGL(k_clear_counters, dim3(1), dim3(64), 0, s, cnt[k]);
    VPIC_LAUNCHED("k_clear_counters");
    // a flat push moves particles without recording their tiles
    if (!sp[k].partition) push_invalidate_bins(sp[k].p);
  }
  const GridDev G = make_grid_dev(g);
  // the timing brackets the main push kernel only: k_advance_p_tiled when
  // a partition is given, else k_advance_p
  const bool timed = push_timing_on();
  const int *flat_start[TILE_MAXSP] = {};
  if (sp[0].partition && np_total > 0) {
    // (one species here, or several that all have partitions)
    vpic_particle_t *p = sp[0].p;
    const int np = sp[0].np;
    const int *partition = sp[0].partition;
    TileGeom T = g_tile;
    // a tile no larger than the grid (a 2D grid's tiles are one cell
    // thick; the window then has no halo along that axis either)
    T.tx = T.tx < g->nx ? T.tx : g->nx;
    T.ty = T.ty < g->ny ? T.ty : g->ny;
    T.tz = T.tz < g->nz ? T.tz : g->nz;
    T.ntx = (g->nx + T.tx - 1) / T.tx;
    T.nty = (g->ny + T.ty - 1) / T.ty;
    T.ntz = (g->nz + T.tz - 1) / T.tz;
    const unsigned sx1 = g->nx + 2, sy1 = g->ny + 2;
    const bool magic_ok = (uint64_t)g->nv * sx1 < ((uint64_t)1 << 32);
    T.mx = magic_ok ? div_magic(sx1) : 0;
    T.my = magic_ok ? div_magic(sy1) : 0;
    tile_axis_halos(T, g);
    const size_t wcells = tile_window(T);
    const size_t win_bytes = sizeof(unsigned long long) * 12 * wcells;
    const size_t scells = tile_ring(T);
    if (win_bytes + 72 * (size_t)T.tx * T.ty * T.tz > (size_t)TILE_DYN_BYTES)
      return vpic_set_error(VPIC_EINVAL,
                            "advance_p: push tile %dx%dx%d with halo %d "
                            "needs %zu window cells on this grid, more "
                            "than the LDS holds", T.tx, T.ty, T.tz, T.halo,
                            wcells);
    T.stage_win = win_bytes + 72 * scells <= (size_t)TILE_DYN_BYTES;
    const size_t lds =
        win_bytes +
        72 * (T.stage_win ? scells : (size_t)T.tx * T.ty * T.tz);
    const int nblk = T.ntx * T.nty * T.ntz;
    const int wstride = (int)(12 * wcells);
    if ((rc = scratch_reserve(S.wbuf, sizeof(float) * (size_t)wstride *
                                          (size_t)nblk)))
      return rc;
    if ((rc = scratch_reserve(S.trash, sizeof(vpic_particle_t) * TILE_BS)))
      return rc;
    // Sub-buckets for out-of-window deposits: 27 per tile of 1/32 of a
    // tile's average particle count (all species) each (late in a sort
    // interval about a tenth of the electrons' deposits leave their
    // window, mostly across the six faces), 64 to 4096 records; a tile's
    // whole count when that takes at most 256 MiB (small grids, whose
    // narrow or periodic-edge tiles can send most of their deposits to
    // one neighbour).  A deposit that finds its sub-bucket full takes a
    // float atomic.  Every tiled push writes all 27 counts of every
    // tile, so they need no clearing.
    OowBuckets OB;
    {
      const int64_t per_tile = np_total / nblk + 1;
      int64_t cap = per_tile / 32;
      cap = cap < 64 ? 64 : cap > 4096 ? 4096 : cap;
      if (64 * 27 * per_tile * nblk <= ((int64_t)256 << 20) && per_tile > cap)
        cap = per_tile;
      if ((rc = scratch_reserve(S.ocnt, sizeof(int) * 54 * (size_t)nblk)) ||
          (rc = scratch_reserve(S.orec,
                                64 * 27 * (size_t)cap * (size_t)nblk)))
        return rc;
      OB.cnt = (int *)S.ocnt.ptr;
      OB.rec = (float4 *)S.orec.ptr;
      OB.cap = (int)cap;
    }
    static bool attr_set[TILE_MAXSP][7] = {};
    const int ak = tiled_kernel_id(g_group);
    if (!attr_set[nsp - 1][ak]) {
      for (int mode = 0; mode < (nsp == 1 ? 3 : 1); mode++)
        VPIC_HIP(hipFuncSetAttribute(
            reinterpret_cast<const void *>(tiled_kernel(g_group, mode, nsp)),
            hipFuncAttributeMaxDynamicSharedMemorySize, TILE_DYN_BYTES));
      attr_set[nsp - 1][ak] = true;
    }
    // Tile lists from the codes of the last tiled push of this array.
    BinArgs BA = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
    BinState *BSt = nullptr;
    if (g_bin_mode > 0 && nsp == 1) {
      BinState &B = bin_state(S, p);
      const int nt = nblk;
      const int nsum = (nt + 1 + 2047) / 2048 + 1;
      if ((rc = scratch_reserve(B.key, (size_t)np)) ||
          (rc = scratch_reserve(B.counts, sizeof(int) * 27 * (size_t)nt)) ||
          (rc = scratch_reserve(B.base, sizeof(int) * 27 * (size_t)nt)) ||
          (rc = scratch_reserve(B.total, sizeof(int) * (size_t)(nt + 1))) ||
          (rc = scratch_reserve(B.sums, sizeof(int) * (size_t)nsum)) ||
          (rc = scratch_reserve(B.ctl, sizeof(int) * BIN_CTL)) ||
          (rc = scratch_reserve(B.list[0], sizeof(int) * (size_t)np)) ||
          (rc = scratch_reserve(B.list[1], sizeof(int) * (size_t)np)) ||
          (rc = scratch_reserve(B.start[0], sizeof(int) * (size_t)(nt + 1))) ||
          (rc = scratch_reserve(B.start[1], sizeof(int) * (size_t)(nt + 1))))
        return rc;
      int *ctl = (int *)B.ctl.ptr;
      if (B.np != np || B.ntiles != nt) {
        hipLaunchKernelGGL(k_fill_int32, dim3(1), dim3(64), 0, s, ctl,
                           BIN_CTL, 0);
        VPIC_LAUNCHED("k_fill_int32");
        B.np = np;
        B.ntiles = nt;
        B.cur = 0;
      }
      const int nxt = B.cur ^ 1;
      int *counts = (int *)B.counts.ptr, *base = (int *)B.base.ptr;
      int *total = (int *)B.total.ptr;
      int *snext = (int *)B.start[nxt].ptr;
      hipLaunchKernelGGL(k_bin_decide, dim3(1), dim3(64), 0, s, ctl,
                         partition, g->nv, np, g_bin_mode);
      VPIC_LAUNCHED("k_bin_decide");
      hipLaunchKernelGGL(k_bin_total, dim3((nt + 1 + 255) / 256), dim3(256),
                         0, s, (const int *)ctl, (const int *)counts, total,
                         T);
      VPIC_LAUNCHED("k_bin_total");
      if ((rc = device_exclusive_scan(total, snext, nt + 1,
                                      (int *)B.sums.ptr, s)))
        return rc;
      hipLaunchKernelGGL(k_bin_base, dim3((nt + 255) / 256), dim3(256), 0,
                         s, (const int *)ctl, (const int *)counts,
                         (const int *)snext, base, T);
      VPIC_LAUNCHED("k_bin_base");
      hipLaunchKernelGGL(k_bin_scatter, dim3((nt + 3) / 4), dim3(256), 0, s,
                         ctl, partition, np,
                         (const int *)B.list[B.cur].ptr,
                         (const int *)B.start[B.cur].ptr,
                         (const unsigned char *)B.key.ptr,
                         (const int *)counts, (const int *)base,
                         (int *)B.list[nxt].ptr, G, T);
      VPIC_LAUNCHED("k_bin_scatter");
      BA.ctl = ctl;
      BA.list = (const int *)B.list[nxt].ptr;
      BA.lstart = snext;
      BA.key = (unsigned char *)B.key.ptr;
      BA.counts = counts;
      BA.oow = ctl + 4;
      BSt = &B;
    }
    if (timed) push_timing_record(s, np_total, true);
    // without binning only the mode-0 instance; with it all three, of
    // which the one matching the device's binning state runs
    TileSpeciesSet SS = {};
    for (int k = 0; k < nsp; k++)
      SS.s[k] = {sp[k].p, sp[k].partition, rlist[k], cnt[k], sp[k].keys,
                 C[k], sp[k].np, xcap[k]};
    for (int mode = 0; mode < (BSt ? 3 : 1); mode++) {
      hipLaunchKernelGGL(tiled_kernel(g_group, mode, nsp), dim3(nblk),
                         dim3(TILE_BS), lds, s, SS, f, (float *)a,
                         (float *)S.wbuf.ptr, wstride,
                         (vpic_particle_t *)S.trash.ptr, G, T, BA, OB);
      VPIC_LAUNCHED("k_advance_p_tiled");
    }
    if (timed) push_timing_record(s, np_total, false);
    if (BSt) {
      hipLaunchKernelGGL(k_bin_post, dim3(1), dim3(64), 0, s,
                         (int *)BSt->ctl.ptr, partition, g->nv, np);
      VPIC_LAUNCHED("k_bin_post");
      BSt->cur ^= 1;
    }
    hipLaunchKernelGGL(k_merge_windows, dim3(nblk), dim3(256), 0, s,
                       (float *)a, (const float *)S.wbuf.ptr, wstride, G, T,
                       OB);
    VPIC_LAUNCHED("k_merge_windows");
    // particles past the partition (added since the sort) go flat
    for (int k = 0; k < nsp; k++) flat_start[k] = sp[k].partition + g->nv;
  }
  for (int k = 0; k < nsp; k++) {
    const int np = sp[k].np;
    if (np > 0) {
      int64_t want = ((int64_t)np + 255) / 256;
      const int blocks = (int)(want < 16384 ? want : 16384);
      const bool flat_timed = timed && !sp[k].partition;
      if (flat_timed) push_timing_record(s, np, true);
      hipLaunchKernelGGL(k_advance_p, dim3(blocks), dim3(256), 0, s, sp[k].p,
                         flat_start[k], np, f, (float *)a, xlist[k], xcap[k],
                         rlist[k], cnt[k], C[k], G, sp[k].keys);
      VPIC_LAUNCHED("k_advance_p");
      if (flat_timed) push_timing_record(s, np, false);
      int mblocks = (xcap[k] + 255) / 256;
      if (mblocks > 8192) mblocks = 8192;
      hipLaunchKernelGGL(k_move_p_list, dim3(mblocks), dim3(256), 0, s,
                         sp[k].p, xlist[k], xcap[k], rlist[k], (float *)a,
                         cnt[k], sp[k].q, G, sp[k].keys);
      VPIC_LAUNCHED("k_move_p_list");
    }
    hipLaunchKernelGGL(k_order_movers, dim3(1), dim3(1024), 0, s, sp[k].p,
                       rlist[k], xcap[k], (int2 *)S.sortbuf[k].ptr, sp[k].pm,
                       sp[k].max_nm, cnt[k], sp[k].dev_nm);
    VPIC_LAUNCHED("k_order_movers");
  }
  return VPIC_OK;
}

// --------------------------------------------------------------------------
// The tiled push of a batched ensemble (k_ens_push_tiled + k_ens_merge):
// per-box tile geometry, window slots, out-of-window sub-buckets and the
// box-relative partitions the batched sort writes.
// --------------------------------------------------------------------------
struct EnsTiled {
  int nbox = 0, nitems = 0;
  size_t lds = 0;
  EnsTileBox *boxes = nullptr;
  EnsTileItem *items = nullptr;
  int *epart = nullptr;  // box b's partition at voff_b + b (nv_b + 1 ints)
  float *wbuf = nullptr;
  int *ocnt = nullptr;
  float4 *orec = nullptr;
  vpic_particle_t *trash = nullptr;
  vpic_particle_mover_t *rlist = nullptr;
  int *cnt = nullptr;  // the push's counters: [1] movers, [3] bad voxels
  ~EnsTiled() {
    void *bufs[] = {boxes, items, epart, wbuf, ocnt, orec, trash, rlist, cnt};
    for (void *b : bufs)
      if (b) (void)hipFree(b);
  }
};

namespace {
int ens_dev_alloc(void **p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes < 16 ? 16 : bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return vpic_set_error(VPIC_ENOMEM, "ensemble tiles: hipMalloc(%zu): %s",
                          bytes, hipGetErrorString(e));
  }
  return VPIC_OK;
}
}  // namespace

int ens_tiled_create(EnsTiled **out, int nbox, const vpic_grid_t *grids,
                     const int *np, const float *q, const float *m,
                     const int64_t *voff, const int64_t *poff,
                     vpic_particle_t *P, const vpic_interpolator_t *FI,
                     vpic_accumulator_t *A, int tx, int ty, int tz,
                     int halo) {
  *out = nullptr;
  if (tx < 1 || ty < 1 || tz < 1 || halo < 1 || halo > ENS_MAX_HALO)
    return vpic_set_error(VPIC_EINVAL, "ensemble tiles: bad tile shape");
  std::vector<EnsTileBox> hb(nbox);
  std::vector<EnsTileItem> hi;
  int64_t wtot = 0, otot = 0, rtot = 0, nvtot = 0;
  size_t lds = 0;
  for (int b = 0; b < nbox; b++) {
    const vpic_grid_t &g = grids[b];
    TileGeom T = {};
    T.tx = tx < g.nx ? tx : g.nx;
    T.ty = ty < g.ny ? ty : g.ny;
    T.tz = tz < g.nz ? tz : g.nz;
    T.halo = halo;
    tile_axis_halos(T, &g);
    if (T.ty * T.tz > TILE_MAXROWS || T.tx * T.ty * T.tz > TILE_MAXCELLS ||
        96 * tile_window(T) + 72 * (size_t)T.tx * T.ty * T.tz >
            (size_t)TILE_DYN_BYTES)
      return vpic_set_error(VPIC_EINVAL, "ensemble tiles: tile %dx%dx%d of "
                            "box %d does not fit the LDS window",
                            T.tx, T.ty, T.tz, b);
    T.ntx = (g.nx + T.tx - 1) / T.tx;
    T.nty = (g.ny + T.ty - 1) / T.ty;
    T.ntz = (g.nz + T.tz - 1) / T.tz;
    const unsigned sx1 = g.nx + 2, sy1 = g.ny + 2;
    if ((uint64_t)g.nv * sx1 >= ((uint64_t)1 << 32))
      return vpic_set_error(VPIC_EINVAL, "ensemble tiles: box %d too large",
                            b);
    T.mx = div_magic(sx1);
    T.my = div_magic(sy1);
    const size_t wcells = tile_window(T);
    const size_t win_bytes = sizeof(unsigned long long) * 12 * wcells;
    const size_t scells = tile_ring(T);
    T.stage_win = win_bytes + 72 * scells <= (size_t)TILE_DYN_BYTES;
    const size_t l = win_bytes +
                     72 * (T.stage_win ? scells : (size_t)T.tx * T.ty * T.tz);
    if (l > lds) lds = l;
    const int ntiles = T.ntx * T.nty * T.ntz;
    int64_t cap = (int64_t)np[b] / ntiles / 32;
    cap = cap < 64 ? 64 : cap > 4096 ? 4096 : cap;
    EnsTileBox &B = hb[b];
    B.p = P + poff[b];
    B.part = nullptr;  // set below (device addresses)
    B.f0 = FI + voff[b];
    B.acc = (float *)(A + voff[b]);
    B.wstride = (int)(12 * wcells);
    B.wbuf = reinterpret_cast<float *>((intptr_t)wtot);    // offset for now
    B.OB.cnt = reinterpret_cast<int *>((intptr_t)otot);
    B.OB.rec = reinterpret_cast<float4 *>((intptr_t)rtot);
    B.OB.cap = (int)cap;
    B.C = push_coef(q[b], m[b], &g);
    B.G = make_grid_dev(&g);
    B.T = T;
    B.np = np[b];
    B.part = reinterpret_cast<const int *>((intptr_t)(voff[b] + b));
    wtot += (int64_t)ntiles * B.wstride;
    otot += 54 * (int64_t)ntiles;
    rtot += 4 * 27 * cap * (int64_t)ntiles;
    nvtot += g.nv + 1;
    for (int t = 0; t < ntiles; t++) hi.push_back({b, t});
  }
  EnsTiled *et = new EnsTiled;
  et->nbox = nbox;
  et->nitems = (int)hi.size();
  et->lds = lds;
  int rc;
  if ((rc = ens_dev_alloc((void **)&et->boxes, sizeof(EnsTileBox) * nbox)) ||
      (rc = ens_dev_alloc((void **)&et->items,
                          sizeof(EnsTileItem) * hi.size())) ||
      (rc = ens_dev_alloc((void **)&et->epart, sizeof(int) * nvtot)) ||
      (rc = ens_dev_alloc((void **)&et->wbuf, sizeof(float) * wtot)) ||
      (rc = ens_dev_alloc((void **)&et->ocnt, sizeof(int) * otot)) ||
      (rc = ens_dev_alloc((void **)&et->orec, sizeof(float4) * rtot)) ||
      (rc = ens_dev_alloc((void **)&et->trash,
                          sizeof(vpic_particle_t) * TILE_BS)) ||
      (rc = ens_dev_alloc((void **)&et->rlist, sizeof(vpic_particle_mover_t))) ||
      (rc = ens_dev_alloc((void **)&et->cnt, sizeof(int) * 64))) {
    delete et;
    return rc;
  }
  for (EnsTileBox &B : hb) {
    B.wbuf = et->wbuf + (intptr_t)B.wbuf;
    B.OB.cnt = et->ocnt + (intptr_t)B.OB.cnt;
    B.OB.rec = et->orec + (intptr_t)B.OB.rec;
    B.part = et->epart + (intptr_t)B.part;
  }
  hipError_t he;
  if ((he = hipMemcpy(et->boxes, hb.data(), sizeof(EnsTileBox) * nbox,
                      hipMemcpyHostToDevice)) != hipSuccess ||
      (he = hipMemcpy(et->items, hi.data(), sizeof(EnsTileItem) * hi.size(),
                      hipMemcpyHostToDevice)) != hipSuccess ||
      (he = hipMemset(et->cnt, 0, sizeof(int) * 64)) != hipSuccess) {
    delete et;
    return vpic_check_hip(he, "ensemble tiles");
  }
  *out = et;
  return VPIC_OK;
}

void ens_tiled_destroy(EnsTiled *t) { delete t; }

int *ens_tiled_epart(EnsTiled *t) { return t ? t->epart : nullptr; }

int *ens_tiled_counters(EnsTiled *t) { return t ? t->cnt : nullptr; }

int launch_ens_push_tiled(EnsTiled *et, int *queue, int nblocks,
                          hipStream_t s) {
  static bool attr_set = false;
  if (!attr_set) {
    VPIC_HIP(hipFuncSetAttribute(
        reinterpret_cast<const void *>(k_ens_push_tiled),
        hipFuncAttributeMaxDynamicSharedMemorySize, TILE_DYN_BYTES));
    attr_set = true;
  }
  if (et->nitems <= 0) return VPIC_OK;
  const int nb = nblocks < et->nitems ? nblocks : et->nitems;
  hipLaunchKernelGGL(k_ens_push_tiled, dim3(nb), dim3(TILE_BS), et->lds, s,
                     (const EnsTileBox *)et->boxes,
                     (const EnsTileItem *)et->items, et->nitems, queue,
                     et->trash, et->rlist, et->cnt);
  VPIC_LAUNCHED("k_ens_push_tiled");
  hipLaunchKernelGGL(k_ens_merge, dim3(et->nitems), dim3(256), 0, s,
                     (const EnsTileBox *)et->boxes,
                     (const EnsTileItem *)et->items, et->nitems);
  VPIC_LAUNCHED("k_ens_merge");
  return VPIC_OK;
}

int launch_ens_push(vpic_particle_t *P, const vpic_interpolator_t *FI,
                    vpic_accumulator_t *A, const EnsPushBox *boxes,
                    const EnsItem *items, int nitems, int *queue, int *nbad,
                    int nblocks, int bs, hipStream_t s) {
  if (nitems <= 0) return VPIC_OK;
  if (bs < 64 || bs > 256 || (bs & 63))
    return vpic_set_error(VPIC_EINVAL, "ensemble push: block size %d", bs);
  hipLaunchKernelGGL(k_ens_push, dim3(nblocks < nitems ? nblocks : nitems),
                     dim3(bs), 0, s, P, FI, (float *)A, boxes, items, nitems,
                     queue, nbad);
  VPIC_LAUNCHED("k_ens_push");
  return VPIC_OK;
}

int launch_inject_p(vpic_particle_t *p, int *np, int max_np,
                    vpic_particle_mover_t *pm, int *nm, int max_nm,
                    const vpic_particle_injector_t *pi, int n,
                    vpic_accumulator_t *a, float q, const vpic_grid_t *g,
                    int *counts, hipStream_t s) {
  counts[0] = counts[1] = 0;
  if (n <= 0) return VPIC_OK;
  PushScratch &S = push_scratch();
  int rc;
  if ((rc = scratch_reserve(S.inject, 64))) return rc;
  int st[8] = {*np, *nm, 0, 0, 0, 0, 0, 0};
  int *dst = (int *)S.inject.ptr;
  VPIC_HIP(hipMemcpyAsync(dst, st, sizeof st, hipMemcpyHostToDevice, s));
  const GridDev G = make_grid_dev(g);
  hipLaunchKernelGGL(k_inject_p, dim3(1), dim3(64), 0, s, p, max_np, pm,
                     max_nm, pi, n, (float *)a, G, q, dst);
  VPIC_LAUNCHED("k_inject_p");
  VPIC_HIP(hipMemcpyAsync(st, dst, sizeof st, hipMemcpyDeviceToHost, s));
  VPIC_HIP(hipStreamSynchronize(s));
  *np = st[0];
  *nm = st[1];
  counts[0] = st[2];
  counts[1] = st[3];
  if (st[4])
    return vpic_set_error(VPIC_EINVAL,
                          "inject_p: %d injectors with a voxel index off the "
                          "grid were skipped", st[4]);
  return VPIC_OK;
}

int launch_move_p_list(vpic_particle_t *p, vpic_particle_mover_t *pm, int nm,
                       vpic_accumulator_t *a, const vpic_grid_t *g, float q,
                       int *dev_ret, hipStream_t s) {
  const GridDev G = make_grid_dev(g);
  hipLaunchKernelGGL(k_move_p_serial, dim3(1), dim3(64), 0, s, p, pm, nm,
                     (float *)a, G, q, dev_ret);
  VPIC_LAUNCHED("k_move_p_serial");
  return VPIC_OK;
}
