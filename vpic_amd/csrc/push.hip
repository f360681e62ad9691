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
//    window (the tile plus a one-cell halo, 48 B per cell) and the window is
//    added to the global accumulator once at the end.  Cell crossers are
//    queued in LDS and moved by the same workgroup, so their segments also
//    land in the window.  Deposits outside the window (fast particles,
//    periodic wrap) fall back to global atomics.
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

// A box of voxels [w0, w0+wn) per axis held in LDS, 12 floats per cell.
struct WindowSink {
  float *lds;  // window cells * 12 (LDS)
  float *acc;  // global accumulator
  int wx0, wy0, wz0, wnx, wny, wnz;
  int sx1, sy1;     // nx+2, ny+2
  unsigned mx, my;  // fast_div multipliers for sx1, sy1 (0: divide)
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
    if (s >= 0) {
      float *a = lds + 12 * s;
#pragma unroll
      for (int k = 0; k < 12; k++) atomicAdd(a + k, j[k]);
    } else {
      float *a = acc + 12 * (int64_t)voxel;
#pragma unroll
      for (int k = 0; k < 12; k++) atomic_add_f32(a + k, j[k]);
    }
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
                          const GridDev &G, float qsp) {
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
  mv.dispx = md[0];
  mv.dispy = md[1];
  mv.dispz = md[2];
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

__device__ __forceinline__ void push_compute(
    const float4 A, const float4 B, bool valid_slot,
    vpic_particle_t *__restrict__ p, int n,
    const vpic_interpolator_t *__restrict__ f0, const PushCoef &C, int nv,
    int *__restrict__ nbad, PushResult &R) {
  const float one = 1.0f, one_third = 1.0 / 3.0, two_fifteenths = 2.0 / 15.0;
  // A voxel index outside the grid (e.g. a mover left pending as
  // 8*voxel+face) is never dereferenced: the particle is left untouched and
  // counted, and the call reports an error.
  const int iraw = __float_as_int(A.w);
  const bool active = valid_slot && (unsigned)iraw < (unsigned)nv;
  if (valid_slot && !active) atomicAdd(nbad, 1);
  float dx = A.x, dy = A.y, dz = A.z;
  const int ii = active ? iraw : 0;

  // Interpolate E and B (advance_p_pipeline.cc:96-109).
  const float4 *fp = reinterpret_cast<const float4 *>(f0 + ii);
  const float4 fe0 = fp[0], fe1 = fp[1], fe2 = fp[2], fb0 = fp[3];
  const float2 fb1 = reinterpret_cast<const float2 *>(fp + 4)[0];
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
    if (active) pp[1] = Bnew;  // momentum stored, position left for move_p
  }
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
// Flat kernel: particles [n0, np), any order, grid-stride over whole
// wavefronts.  n0 = *start clamped to [0,np] (the end of the sort-time
// partition when the tiled kernel ran first), or 0 when start is NULL.
// --------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_advance_p(
    vpic_particle_t *__restrict__ p, const int *__restrict__ start, int np,
    const vpic_interpolator_t *__restrict__ f0, float *__restrict__ acc,
    vpic_particle_mover_t *__restrict__ xlist, int xcap,
    vpic_particle_mover_t *__restrict__ rlist, int *__restrict__ cnt,
    PushCoef C, GridDev G) {
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
        } else if (dev_move_p(p, mv, sink, G, C.qsp)) {
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
    float *__restrict__ acc, int *__restrict__ cnt, float qsp, GridDev G) {
  const int n = min(cnt[0], xcap);
  const GlobalSink sink{acc};
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n;
       k += gridDim.x * blockDim.x) {
    vpic_particle_mover_t mv = xlist[k];
    if (dev_move_p(p, mv, sink, G, qsp))
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
};

constexpr int TILE_BS = 512;            // threads per workgroup
constexpr int TILE_MAXROWS = 256;       // ty * tz
constexpr int TILE_QCAP = 2 * TILE_BS;  // crosser queue

// Particle index of slot k of the tile (rows of the sort-time partition);
// r0 is a row at or before k's.
__device__ __forceinline__ int tile_particle(const int *row_start,
                                             const int *row_pref, int nrow,
                                             int r0, int k) {
  int r = r0;
  while (r + 1 < nrow && row_pref[r + 1] <= k) r++;
  return row_start[r] + (k - row_pref[r]);
}

__global__ __launch_bounds__(TILE_BS) void k_advance_p_tiled(
    vpic_particle_t *__restrict__ p, int np, const int *__restrict__ part,
    const vpic_interpolator_t *__restrict__ f0, float *__restrict__ acc,
    vpic_particle_mover_t *__restrict__ rlist, int rcap, int *__restrict__ cnt,
    PushCoef C, GridDev G, TileGeom T) {
  extern __shared__ float lds_acc[];
  __shared__ vpic_particle_mover_t q[TILE_QCAP];
  __shared__ int row_start[TILE_MAXROWS];
  __shared__ int row_pref[TILE_MAXROWS + 1];
  __shared__ int qn;

  const int t = threadIdx.x;
  const int lane = lane_id();
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

  WindowSink W;
  W.lds = lds_acc;
  W.acc = acc;
  W.wx0 = x0 - 1;
  W.wy0 = y0 - 1;
  W.wz0 = z0 - 1;
  W.wnx = x1 - x0 + 3;
  W.wny = y1 - y0 + 3;
  W.wnz = z1 - z0 + 3;
  W.sx1 = G.nx + 2;
  W.sy1 = G.ny + 2;
  W.mx = T.mx;
  W.my = T.my;
  const int ncell = W.wnx * W.wny * W.wnz;

  // Rows of the tile in the sort-time partition, clamped to [0, np).
  for (int r = t; r < nrow; r += TILE_BS) {
    const int y = y0 + r % rows_y, z = z0 + r / rows_y;
    int s = part[vox(x0, y, z, G.nx, G.ny)];
    int e = part[vox(x1 + 1, y, z, G.nx, G.ny)];
    s = min(max(s, 0), np);
    e = min(max(e, s), np);
    row_start[r] = s;
    row_pref[r + 1] = e - s;
  }
  for (int k = t; k < 12 * ncell; k += TILE_BS) lds_acc[k] = 0.0f;
  if (t == 0) qn = 0;
  __syncthreads();
  if (t == 0) {
    row_pref[0] = 0;
    for (int r = 0; r < nrow; r++) row_pref[r + 1] += row_pref[r];
  }
  __syncthreads();
  const int total = row_pref[nrow];

  int r0 = 0;  // row holding particle slot `base` (same in every lane)
  int n = total > t ? tile_particle(row_start, row_pref, nrow, 0, t) : 0;
  float4 A, B;
  load_particle(p, n, t < total, A, B);
  for (int base = 0; base < total; base += TILE_BS) {
    const bool slot_ok = base + t < total;
    // prefetch the next round's particle
    const int kn = base + TILE_BS + t;
    int r1 = r0;
    while (r1 + 1 < nrow && row_pref[r1 + 1] <= base + TILE_BS) r1++;
    const int nn = kn < total ? tile_particle(row_start, row_pref, nrow, r1, kn)
                              : 0;
    float4 An, Bn;
    load_particle(p, nn, kn < total, An, Bn);

    PushResult R;
    push_compute(A, B, slot_ok, p, n, f0, C, G.nv, cnt + 3, R);
    wave_deposit(R.active ? R.voxel : -1, R.j, W);

    // queue this round's crossers
    const uint64_t cm = __ballot(R.cross);
    if (cm) {
      const int leader = __ffsll((unsigned long long)cm) - 1;
      int qb = 0;
      if (lane == leader) qb = atomicAdd(&qn, __popcll(cm));
      qb = __shfl(qb, leader);
      if (R.cross) {
        vpic_particle_mover_t mv;
        mv.dispx = R.dispx;
        mv.dispy = R.dispy;
        mv.dispz = R.dispz;
        mv.i = n;
        q[qb + __popcll(cm & ((1ull << lane) - 1ull))] = mv;
      }
    }
    __syncthreads();
    const int nq = qn;
    __syncthreads();  // every lane has read qn before anyone changes it
    if (nq >= TILE_BS) {  // uniform: move the top TILE_BS queued crossers
      vpic_particle_mover_t mv = q[nq - TILE_BS + t];
      if (t == 0) qn = nq - TILE_BS;
      if (dev_move_p(p, mv, W, G, C.qsp))
        push_returned_mover(p, mv, rlist, rcap, cnt);
      __syncthreads();
    }
    r0 = r1;
    n = nn;
    A = An;
    B = Bn;
  }
  {
    const int nq = qn;
    if (t < nq) {
      vpic_particle_mover_t mv = q[t];
      if (dev_move_p(p, mv, W, G, C.qsp))
        push_returned_mover(p, mv, rlist, rcap, cnt);
    }
  }
  __syncthreads();
  // Add the window to the global accumulator (untouched cells skipped).
  for (int k = t; k < 12 * ncell; k += TILE_BS) {
    const float v = lds_acc[k];
    if (v != 0.0f) {
      const int c = k / 12, comp = k - 12 * c;
      const int cx = c % W.wnx, cr = c / W.wnx;
      const int cy = cr % W.wny, cz = cr / W.wny;
      const int vv = vox(W.wx0 + cx, W.wy0 + cy, W.wz0 + cz, G.nx, G.ny);
      atomic_add_f32(acc + 12 * (int64_t)vv + comp, v);
    }
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

// Scratch of the push: counters, crosser list, returned-mover list and the
// ordering buffer, one set per device.
struct PushScratch {
  DeviceScratch counters, xlist, rlist, sortbuf;
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

TileGeom g_tile = {8, 8, 4, 0, 0, 0, 0, 0};

// ceil(2^32 / d) for fast_div.
unsigned div_magic(unsigned d) {
  return (unsigned)((((uint64_t)1 << 32) + d - 1) / d);
}

}  // namespace

// Tile shape of the tiled push (set by the tuner; the window of a tile is
// (tx+2)(ty+2)(tz+2) cells of 48 B in LDS).
int push_set_tile(int tx, int ty, int tz) {
  if (tx < 1 || ty < 1 || tz < 1 || ty * tz > TILE_MAXROWS) return VPIC_EINVAL;
  if ((size_t)(tx + 2) * (ty + 2) * (tz + 2) * 48 > 96 * 1024)
    return VPIC_EINVAL;
  g_tile.tx = tx;
  g_tile.ty = ty;
  g_tile.tz = tz;
  return VPIC_OK;
}

int push_set_direct_runs(int runs) {
  if (runs < 2 || runs > 65) return VPIC_EINVAL;
  hipError_t e = hipMemcpyToSymbol(HIP_SYMBOL(c_direct_runs), &runs,
                                   sizeof(int));
  return e == hipSuccess ? VPIC_OK : vpic_check_hip(e, "hipMemcpyToSymbol");
}

int launch_advance_p(vpic_particle_t *p, int np, float q, float m,
                     vpic_particle_mover_t *pm, int max_nm, int *dev_nm,
                     vpic_accumulator_t *a, const vpic_interpolator_t *f,
                     const vpic_grid_t *g, const int *partition,
                     hipStream_t s) {
  PushScratch &S = push_scratch();
  // Crosser list sized for ~20% of the particles per step; the rare
  // excess is moved inline by the push kernel.
  int64_t xcap64 = np / 5;
  if (xcap64 < 65536) xcap64 = 65536;
  if (xcap64 > np) xcap64 = np < 1 ? 1 : np;
  const int xcap = (int)xcap64;
  int rc;
  if ((rc = scratch_reserve(S.counters, 64)) ||
      (rc = scratch_reserve(S.xlist, sizeof(vpic_particle_mover_t) * xcap)) ||
      (rc = scratch_reserve(S.rlist, sizeof(vpic_particle_mover_t) * xcap)) ||
      (rc = scratch_reserve(S.sortbuf, sizeof(int2) * pow2_at_least(xcap))))
    return rc;
  int *cnt = (int *)S.counters.ptr;
  vpic_particle_mover_t *xlist = (vpic_particle_mover_t *)S.xlist.ptr;
  vpic_particle_mover_t *rlist = (vpic_particle_mover_t *)S.rlist.ptr;

  // advance_p_pipeline.cc:279-283, in float.
  PushCoef C;
  C.qdt_2mc = (q * g->dt) / (2 * m * g->cvac);
  C.cdt_dx = g->cvac * g->dt * g->rdx;
  C.cdt_dy = g->cvac * g->dt * g->rdy;
  C.cdt_dz = g->cvac * g->dt * g->rdz;
  C.qsp = q;
  const GridDev G = make_grid_dev(g);

  hipLaunchKernelGGL(k_clear_counters, dim3(1), dim3(64), 0, s, cnt);
  VPIC_LAUNCHED("k_clear_counters");
  if (np > 0) {
    const bool timed = push_timing_on();
    if (timed) push_timing_record(s, np, true);
    const int *flat_start = nullptr;
    if (partition) {
      TileGeom T = g_tile;
      T.ntx = (g->nx + T.tx - 1) / T.tx;
      T.nty = (g->ny + T.ty - 1) / T.ty;
      T.ntz = (g->nz + T.tz - 1) / T.tz;
      const unsigned sx1 = g->nx + 2, sy1 = g->ny + 2;
      const bool magic_ok = (uint64_t)g->nv * sx1 < ((uint64_t)1 << 32);
      T.mx = magic_ok ? div_magic(sx1) : 0;
      T.my = magic_ok ? div_magic(sy1) : 0;
      const size_t lds = sizeof(float) * 12 * (size_t)(T.tx + 2) *
                         (T.ty + 2) * (T.tz + 2);
      static bool attr_set = false;
      if (!attr_set) {
        VPIC_HIP(hipFuncSetAttribute(
            reinterpret_cast<const void *>(k_advance_p_tiled),
            hipFuncAttributeMaxDynamicSharedMemorySize, 96 * 1024));
        attr_set = true;
      }
      const int nblk = T.ntx * T.nty * T.ntz;
      hipLaunchKernelGGL(k_advance_p_tiled, dim3(nblk), dim3(TILE_BS), lds, s,
                         p, np, partition, f, (float *)a, rlist, xcap, cnt, C,
                         G, T);
      VPIC_LAUNCHED("k_advance_p_tiled");
      // particles past the partition (added since the sort) go flat
      flat_start = partition + g->nv;
    }
    {
      int64_t want = ((int64_t)np + 255) / 256;
      const int blocks = (int)(want < 16384 ? want : 16384);
      hipLaunchKernelGGL(k_advance_p, dim3(blocks), dim3(256), 0, s, p,
                         flat_start, np, f, (float *)a, xlist, xcap, rlist,
                         cnt, C, G);
      VPIC_LAUNCHED("k_advance_p");
    }
    if (timed) push_timing_record(s, np, false);
    int mblocks = (xcap + 255) / 256;
    if (mblocks > 8192) mblocks = 8192;
    hipLaunchKernelGGL(k_move_p_list, dim3(mblocks), dim3(256), 0, s, p,
                       xlist, xcap, rlist, (float *)a, cnt, q, G);
    VPIC_LAUNCHED("k_move_p_list");
  }
  hipLaunchKernelGGL(k_order_movers, dim3(1), dim3(1024), 0, s, p, rlist,
                     xcap, (int2 *)S.sortbuf.ptr, pm, max_nm, cnt, dev_nm);
  VPIC_LAUNCHED("k_order_movers");
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
