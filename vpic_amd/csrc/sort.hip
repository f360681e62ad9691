// sort.hip -- sort_p on gfx950: stable sort of the particles by voxel, and
// the voxel partition.
//
// Reference: src/species_advance/standard/pipeline/sort_p_pipeline.cc
// (:143-214 fine subsort, :221-371 driver).  The reference's result is the
// stable sort by p.i (the coarse/fine split never reorders equal keys), and
// partition[v] = index of the first particle of voxel v, 0 below the first
// interior voxel and np above the last (:335-340, :354-360).
//
// Main path (every axis of the grid within RS_MAXB - 2 cells): three stable
// counting passes over (key, index) pairs, x then y then z digit of the
// voxel (least significant first; the voxel index is monotone in (z, y, x)):
//   k_rs_hist    per block of RS_BLK elements, the count of each digit
//                (LDS counters, runs of equal digits over consecutive lanes
//                add once); the first pass reads the keys from the particles
//   scan         exclusive scan of the (digit, block) counts
//   k_rs_scatter every element to its digit's next slot of its block, in
//                index order (ballot match on the digit: stable, no global
//                atomics)
//   k_gather_part  out[j] = p[perm[j]], and the partition from the sorted
//                keys
// Fallback for larger grids (k_count, scan, k_scatter, k_fixup, k_gather):
// a counting sort with one global atomic per run of equal keys, whose
// equal-key segments are then sorted by index in LDS.
// Batched sort of an ensemble's boxes (k_ens_count, k_ens_scatter): one
// counting sort over packed arrays, no order within a voxel.
#include "common.h"

#include <atomic>

namespace {

__device__ __forceinline__ void run_info(int key, int lane, int &seg_start,
                                         int &run_tail, bool &is_tail) {
  const int kprev = __shfl_up(key, 1);
  const int knext = __shfl_down(key, 1);
  const bool head = (lane == 0) || (kprev != key);
  is_tail = (lane == 63) || (knext != key);
  const uint64_t heads = __ballot(head);
  const uint64_t tails = __ballot(is_tail);
  const uint64_t upto = (lane == 63) ? ~0ull : ((2ull << lane) - 1ull);
  seg_start = 63 - __clzll(heads & upto);
  run_tail = __ffsll((unsigned long long)(tails & ~((1ull << lane) - 1ull))) - 1;
}

__global__ __launch_bounds__(256) void k_fill_int(int *a, int64_t n, int v) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n;
       k += (int64_t)gridDim.x * blockDim.x)
    a[k] = v;
}

__global__ __launch_bounds__(256) void k_count(
    const vpic_particle_t *__restrict__ p, int np, int nv,
    int *__restrict__ keys, int *__restrict__ count, int *__restrict__ nbad) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = lane_id();
  int key = k < np ? p[k].i : -1;
  if (k < np && (unsigned)key >= (unsigned)nv) {
    atomicAdd(nbad, 1);  // never used as an index; the sort reports it
    key = -1;
  }
  if (k < np) keys[k] = key;
  int s0, tail_lane;
  bool tail;
  run_info(key, lane, s0, tail_lane, tail);
  if (tail && key >= 0) atomicAdd(count + key, lane - s0 + 1);
}

__global__ __launch_bounds__(256) void k_scatter(const int *__restrict__ keys,
                                                 int np, int *__restrict__ cursor,
                                                 int *__restrict__ perm) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  const int lane = lane_id();
  const int key = k < np ? keys[k] : -1;
  int s0, tail_lane;
  bool tail;
  run_info(key, lane, s0, tail_lane, tail);
  int base = 0;
  if (tail && key >= 0) base = atomicAdd(cursor + key, lane - s0 + 1);
  base = __shfl(base, tail_lane);
  if (key >= 0) perm[base + (lane - s0)] = k;
}

// Block-wise exclusive scan: 256 threads x 8 items.
constexpr int SCAN_ITEMS = 2048;

__device__ int block_exclusive_scan(int v, int *sh) {
  // inclusive scan of 256 values in shared memory, returns exclusive
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int d = 1; d < 256; d <<= 1) {
    int x = t >= d ? sh[t - d] : 0;
    __syncthreads();
    sh[t] += x;
    __syncthreads();
  }
  int incl = sh[t];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(256) void k_scan_local(const int *__restrict__ in,
                                                    int *__restrict__ out,
                                                    int64_t n,
                                                    int *__restrict__ sums) {
  __shared__ int sh[256];
  const int64_t base = (int64_t)blockIdx.x * SCAN_ITEMS + threadIdx.x * 8;
  int v[8], s = 0;
#pragma unroll
  for (int i = 0; i < 8; i++) {
    v[i] = (base + i < n) ? in[base + i] : 0;
    s += v[i];
  }
  int off = block_exclusive_scan(s, sh);
#pragma unroll
  for (int i = 0; i < 8; i++) {
    if (base + i < n) out[base + i] = off;
    off += v[i];
  }
  if (threadIdx.x == 255) sums[blockIdx.x] = off;
}

__global__ __launch_bounds__(256) void k_scan_sums(int *sums, int nb) {
  __shared__ int sh[256];
  int carry = 0;
  for (int c = 0; c < nb; c += 256) {
    const int i = c + threadIdx.x;
    const int v = i < nb ? sums[i] : 0;
    const int ex = block_exclusive_scan(v, sh);
    if (i < nb) sums[i] = ex + carry;
    __syncthreads();
    // total of this chunk
    sh[threadIdx.x] = v;
    __syncthreads();
    if (threadIdx.x == 0) {
      int t = 0;
      for (int j = 0; j < 256; j++) t += sh[j];
      sh[0] = t;
    }
    __syncthreads();
    carry += sh[0];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_scan_add(int *out, int64_t n,
                                                  const int *sums) {
  const int64_t base = (int64_t)blockIdx.x * SCAN_ITEMS;
  const int add = sums[blockIdx.x];
  for (int i = threadIdx.x; i < SCAN_ITEMS; i += 256)
    if (base + i < n) out[base + i] += add;
}

// One single-wavefront workgroup per voxel segment (so every barrier below
// is uniform); segments longer than FIX_CAP go to a list handled by
// k_fixup_big.
constexpr int FIX_CAP = 1024;

__global__ __launch_bounds__(64) void k_fixup(const int *__restrict__ part,
                                              int nv, int *__restrict__ perm,
                                              int *__restrict__ big,
                                              int *__restrict__ nbig) {
  __shared__ int s[FIX_CAP];
  const int lane = threadIdx.x;
  const int v = blockIdx.x;
  if (v >= nv) return;
  const int b = part[v], c = part[v + 1] - b;
  if (c <= 1) return;
  if (c > FIX_CAP) {
    if (lane == 0) {
      big[atomicAdd(nbig, 1)] = v;
      atomicMax(nbig + 1, c);
    }
    return;
  }
  // already ascending?
  bool ok = true;
  for (int k = lane; k + 1 < c; k += 64) ok &= perm[b + k] < perm[b + k + 1];
  if (__all(ok)) return;
  int pad = 1;
  while (pad < c) pad <<= 1;
  for (int k = lane; k < pad; k += 64) s[k] = k < c ? perm[b + k] : 0x7fffffff;
  __syncthreads();
  for (int size = 2; size <= pad; size <<= 1)
    for (int stride = size >> 1; stride > 0; stride >>= 1) {
      for (int k = lane; k < pad; k += 64) {
        const int l = k ^ stride;
        if (l > k) {
          const bool up = (k & size) == 0;
          const int x = s[k], y = s[l];
          if ((x > y) == up) {
            s[k] = y;
            s[l] = x;
          }
        }
      }
      __syncthreads();
    }
  for (int k = lane; k < c; k += 64) perm[b + k] = s[k];
}

// Long segments: one workgroup each, bitonic in global memory on a padded
// copy held in `tmp` (sized for the longest segment).
__global__ __launch_bounds__(1024) void k_fixup_big(const int *__restrict__ part,
                                                    int *__restrict__ perm,
                                                    const int *__restrict__ big,
                                                    const int *__restrict__ nbig,
                                                    int *__restrict__ tmp,
                                                    int64_t tmp_stride) {
  for (int j = blockIdx.x; j < *nbig; j += gridDim.x) {
    const int v = big[j];
    const int b = part[v], c = part[v + 1] - b;
    int *s = tmp + (int64_t)blockIdx.x * tmp_stride;
    int pad = 1;
    while (pad < c) pad <<= 1;
    for (int k = threadIdx.x; k < pad; k += blockDim.x)
      s[k] = k < c ? perm[b + k] : 0x7fffffff;
    __syncthreads();
    for (int size = 2; size <= pad; size <<= 1)
      for (int stride = size >> 1; stride > 0; stride >>= 1) {
        for (int k = threadIdx.x; k < pad; k += blockDim.x) {
          const int l = k ^ stride;
          if (l > k) {
            const bool up = (k & size) == 0;
            const int a = s[k], bb = s[l];
            if ((a > bb) == up) {
              s[k] = bb;
              s[l] = a;
            }
          }
        }
        __syncthreads();
      }
    for (int k = threadIdx.x; k < c; k += blockDim.x) perm[b + k] = s[k];
    __syncthreads();
  }
}

__global__ __launch_bounds__(256) void k_gather(
    const vpic_particle_t *__restrict__ src, const int *__restrict__ perm,
    int np, vpic_particle_t *__restrict__ dst) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= np) return;
  const float4 *s = reinterpret_cast<const float4 *>(src + perm[j]);
  float4 *d = reinterpret_cast<float4 *>(dst + j);
  const float4 a = s[0], b = s[1];
  d[0] = a;
  d[1] = b;
}

// --------------------------------------------------------------------------
// Axis passes: the stable sort by voxel v = x + sx(y + sy z) as three stable
// counting sorts of (key, index) pairs, by x, then y, then z (least
// significant first).  Particles sorted at the last sort_p and moved a few
// cells since are still nearly in voxel order, so each pass sees long runs
// of equal digits: one single-wavefront block counts and places them with
// LDS counters, no global atomics, and lanes of a run write consecutive
// slots.
// --------------------------------------------------------------------------
constexpr int RS_BLK = 16384;  // elements per block (one wavefront)
constexpr int RS_MAXB = 2048;  // digit values per pass (LDS counters)

struct SortDig {
  unsigned sx1, sy1;  // nx + 2, ny + 2
  int nv;
};

__device__ __forceinline__ int axis_digit(int v, int axis, const SortDig &D) {
  const unsigned t = (unsigned)v / D.sx1;
  if (axis == 0) return v - (int)(t * D.sx1);
  const unsigned z = t / D.sy1;
  if (axis == 1) return (int)(t - z * D.sy1);
  return (int)z;
}

// Per block, the count of each digit: hist[d * nblocks + block].  SRC 0:
// the voxel indices read from the particles (checked, kept in keys_out);
// 1: keys_in; 2: keys_in as given by the caller (checked like SRC 0, an
// index outside [0, nv) counted in nbad and taken as 0).
template <int SRC>
__global__ __launch_bounds__(64) void k_rs_hist(
    const vpic_particle_t *__restrict__ p, const int *__restrict__ keys_in,
    int *__restrict__ keys_out, int np, SortDig D, int axis, int nbins,
    int *__restrict__ hist, int nblocks, int *__restrict__ nbad) {
  __shared__ int h[RS_MAXB];
  const int lane = threadIdx.x;
  const int b = blockIdx.x;
  for (int d = lane; d < nbins; d += 64) h[d] = 0;
  __syncthreads();
  const int base = b * RS_BLK, end = min(base + RS_BLK, np);
  int bad = 0;
  // U rounds of 64 keys per batch; the next batch is loaded while this one
  // is counted
  constexpr int U = 4;
  int kk[U];
  // (indices clamped into the block, so every load is issued and the
  // wait for a batch does not wait for the next)
  auto fetch = [&](int i0, int (&k)[U]) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = min(i0 + u * 64 + lane, end - 1);
      k[u] = SRC == 0 ? p[i].i : keys_in[i];
    }
  };
  fetch(base, kk);
  for (int i0 = base; i0 < end; i0 += 64 * U) {
    int kn[U];
    fetch(i0 + 64 * U, kn);
#pragma unroll
    for (int u = 0; u < U; u++) {
    const int i = i0 + u * 64 + lane;
    const bool valid = i < end;
    int key = kk[u];
    if (valid && SRC != 1) {
      if ((unsigned)key >= (unsigned)D.nv) {
        bad++;  // never used as an index; the sort reports it
        key = 0;
      }
      if (SRC == 0) keys_out[i] = key;
    }
    const int d = valid ? axis_digit(key, axis, D) : -1;
    // runs of equal digits over consecutive lanes: the last lane of a run
    // adds its length
    const int dn = __shfl_down(d, 1);
    const int dp = __shfl_up(d, 1);
    const bool head = lane == 0 || dp != d;
    const bool tail = lane == 63 || dn != d;
    const uint64_t heads = __ballot(head);
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const int h0 = 63 - __clzll(heads & upto);
    if (tail && d >= 0) atomicAdd(&h[d], lane - h0 + 1);
    }
#pragma unroll
    for (int u = 0; u < U; u++) kk[u] = kn[u];
  }
  if (SRC != 1 && bad) atomicAdd(nbad, bad);
  __syncthreads();
  for (int d = lane; d < nbins; d += 64) hist[(int64_t)d * nblocks + b] = h[d];
}

// Place every element at offs[d * nblocks + block] + (its rank among the
// block's elements of digit d), in index order: stable.
template <bool FIRST>
__global__ __launch_bounds__(64) void k_rs_scatter(
    const int *__restrict__ keys_in, const int *__restrict__ idx_in, int np,
    SortDig D, int axis, int nbins, int nbits, const int *__restrict__ offs,
    int nblocks, int *__restrict__ keys_out, int *__restrict__ idx_out) {
  __shared__ int pos[RS_MAXB];
  const int lane = threadIdx.x;
  const int b = blockIdx.x;
  for (int d = lane; d < nbins; d += 64) pos[d] = offs[(int64_t)d * nblocks + b];
  __syncthreads();
  const uint64_t lt = (1ull << lane) - 1ull;
  const int base = b * RS_BLK, end = min(base + RS_BLK, np);
  // U rounds of 64 elements per batch; the next batch is loaded while this
  // one is placed
  constexpr int U = 4;
  int kk[U], ii[U];
  // (indices clamped into the block, so every load is issued and the
  // wait for a batch does not wait for the next)
  auto fetch = [&](int i0, int (&k)[U], int (&x)[U]) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int i = min(i0 + u * 64 + lane, end - 1);
      k[u] = keys_in[i];
      x[u] = FIRST ? i : idx_in[i];
    }
  };
  fetch(base, kk, ii);
  for (int i0 = base; i0 < end; i0 += 64 * U) {
    int kn[U], xn[U];
    fetch(i0 + 64 * U, kn, xn);
#pragma unroll
    for (int u = 0; u < U; u++) {
    const int i = i0 + u * 64 + lane;
    const bool valid = i < end;
    // the first pass may read a caller's keys: one outside [0, nv) is
    // counted by the histogram pass and sorted as 0 here
    const int key =
        FIRST ? ((unsigned)kk[u] < (unsigned)D.nv ? kk[u] : 0) : kk[u];
    const int idx = ii[u];
    const int d = valid ? axis_digit(key, axis, D) : 0;
    // lanes holding the same digit
    uint64_t m = __ballot(valid);
    for (int bit = 0; bit < nbits; bit++) {
      const bool on = (d >> bit) & 1;
      const uint64_t bm = __ballot(on);
      m &= on ? bm : ~bm;
    }
    const int rank = __popcll(m & lt);
    const int leader = __ffsll((unsigned long long)m) - 1;
    int p0 = 0;
    if (valid && lane == leader) {
      p0 = pos[d];
      pos[d] = p0 + __popcll(m);
    }
    p0 = __shfl(p0, leader < 0 ? 0 : leader);
    if (valid) {
      keys_out[p0 + rank] = key;
      idx_out[p0 + rank] = idx;
    }
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
      kk[u] = kn[u];
      ii[u] = xn[u];
    }
  }
}

// out[j] = p[perm[j]] for the 128 * GU slots of a block, each particle
// moved by a pair of lanes (16 B each, so every load and store instruction
// covers consecutive 16-B pieces), GU particles in flight per lane, and
// partition[v] = first slot whose key is >= v (partition[nv] = np), from
// the sorted keys (written by the even lane of the pair).  With `stale`
// given, every particle's voxel index is compared with the key it was
// sorted by, and the differences are counted there.  np >= 1.
constexpr int GATHER_GU = 4;
typedef float gather_f4 __attribute__((ext_vector_type(4)));

template <int GU>
__global__ __launch_bounds__(256) void k_gather_part(
    const vpic_particle_t *__restrict__ src, const int *__restrict__ perm,
    const int *__restrict__ keys, int np, int nv,
    vpic_particle_t *__restrict__ dst, int *__restrict__ part,
    int *__restrict__ stale) {
  const int half = threadIdx.x & 1;
  const int64_t j0 = (int64_t)blockIdx.x * (128 * GU) + (threadIdx.x >> 1);
  // indices clamped into the arrays, so that every load is issued at once
  int q[GU], kc[GU], kp[GU];
#pragma unroll
  for (int u = 0; u < GU; u++) {
    const int64_t j = j0 + 128 * u;
    const int jc = (int)(j < np ? j : np - 1);
    const int jp = (int)(j - 1 < np ? (j > 0 ? j - 1 : 0) : np - 1);
    q[u] = perm[jc];
    kc[u] = keys[jc];
    kp[u] = keys[jp];
  }
  gather_f4 a[GU];
#pragma unroll
  for (int u = 0; u < GU; u++)
    a[u] = reinterpret_cast<const gather_f4 *>(src + q[u])[half];
  // every particle load issued before the first store waits for one (the
  // compiler would otherwise move each load down to its store)
#pragma unroll
  for (int u = 0; u < GU; u++) asm volatile("" : "+v"(a[u]));
  int nstale = 0;
#pragma unroll
  for (int u = 0; u < GU; u++) {
    const int64_t j = j0 + 128 * u;
    const int kcu = j < np ? kc[u] : nv;
    const int kpu = j == 0 ? -1 : (j <= np ? kp[u] : nv);
    if (j < np) reinterpret_cast<gather_f4 *>(dst + j)[half] = a[u];
    if (stale && !half && j < np) nstale += __float_as_int(a[u][3]) != kcu;
    if (!half && j <= np)
      for (int v = kpu + 1; v <= kcu && v <= nv; v++) part[v] = (int)j;
  }
  if (nstale) atomicAdd(stale, nstale);
}

struct SortScratch {
  DeviceScratch keys, perm, count, sums, big, nbig, tmp, keys2, idx2, hist;
};
SortScratch &sort_scratch() {
  static thread_local SortScratch s[16];
  int dev = 0;
  (void)hipGetDevice(&dev);
  return s[dev & 15];
}

inline int blocks_for(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

}  // namespace

// Exclusive scan of n ints (in may equal out).  sums: ceil(n/2048) ints.
int device_exclusive_scan(const int *in, int *out, int64_t n, int *sums,
                          hipStream_t s) {
  const int nb = blocks_for(n, SCAN_ITEMS);
  hipLaunchKernelGGL(k_scan_local, dim3(nb), dim3(256), 0, s, in, out, n,
                     sums);
  VPIC_LAUNCHED("k_scan_local");
  hipLaunchKernelGGL(k_scan_sums, dim3(1), dim3(256), 0, s, sums, nb);
  VPIC_LAUNCHED("k_scan_sums");
  hipLaunchKernelGGL(k_scan_add, dim3(nb), dim3(256), 0, s, out, n,
                     (const int *)sums);
  VPIC_LAUNCHED("k_scan_add");
  return VPIC_OK;
}

namespace {

int nbits_for(int nbins) {
  int b = 0;
  while ((1 << b) < nbins) b++;
  return b;
}

// sort_p_axes' answer when the caller's keys were not the particles' own
// (nothing of p was changed; the caller sorts again without them).
constexpr int SORT_KEYS_STALE = 1;

std::atomic<long long> g_sort_keys_stale{0};
void note_sort_keys_stale() { g_sort_keys_stale.fetch_add(1); }

// The axis-pass sort (grids up to RS_MAXB - 2 cells per axis).  With
// `hint` (the caller's copy of p[k].i, k < np) the first pass reads the
// keys from it instead of from the particles, and the gather checks every
// particle's voxel index against the key it was sorted by; a difference,
// or a key outside [0, nv), returns SORT_KEYS_STALE before p is touched.
int sort_p_axes(vpic_particle_t *p, vpic_particle_t *aux, int np,
                int *partition, const vpic_grid_t *g, int in_place,
                const int *hint, hipStream_t s) {
  SortScratch &S = sort_scratch();
  const int nv = g->nv;
  const int nblocks = blocks_for(np, RS_BLK);
  const int bins[3] = {g->nx + 2, g->ny + 2, g->nz + 2};
  const int maxb = bins[0] > bins[1] ? (bins[0] > bins[2] ? bins[0] : bins[2])
                                     : (bins[1] > bins[2] ? bins[1] : bins[2]);
  const int64_t nh = (int64_t)maxb * nblocks;
  int rc;
  if ((rc = scratch_reserve(S.keys, sizeof(int) * (size_t)np)) ||
      (rc = scratch_reserve(S.perm, sizeof(int) * (size_t)np)) ||
      (rc = scratch_reserve(S.keys2, sizeof(int) * (size_t)np)) ||
      (rc = scratch_reserve(S.idx2, sizeof(int) * (size_t)np)) ||
      (rc = scratch_reserve(S.hist, sizeof(int) * (size_t)nh)) ||
      (rc = scratch_reserve(S.sums,
                            sizeof(int) * (size_t)(blocks_for(nh, SCAN_ITEMS) + 1))) ||
      (rc = scratch_reserve(S.nbig, 64)))
    return rc;
  int *ka = (int *)S.keys.ptr, *ia = (int *)S.perm.ptr;
  int *kb = (int *)S.keys2.ptr, *ib = (int *)S.idx2.ptr;
  int *hist = (int *)S.hist.ptr, *sums = (int *)S.sums.ptr;
  int *nbad = (int *)S.nbig.ptr;
  SortDig D;
  D.sx1 = (unsigned)(g->nx + 2);
  D.sy1 = (unsigned)(g->ny + 2);
  D.nv = nv;
  // nbad[0]: keys outside [0, nv); nbad[1]: particles whose voxel index is
  // not the caller's key
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, nbad, (int64_t)2, 0);
  VPIC_LAUNCHED("k_fill_int");
  // pass x: keys from the particles into kb (or the caller's keys), then
  // (keys, index) -> (ka, ia)
  if (hint) {
    hipLaunchKernelGGL(k_rs_hist<2>, dim3(nblocks), dim3(64), 0, s,
                       (const vpic_particle_t *)nullptr, hint, (int *)nullptr,
                       np, D, 0, bins[0], hist, nblocks, nbad);
  } else {
    hipLaunchKernelGGL(k_rs_hist<0>, dim3(nblocks), dim3(64), 0, s, p,
                       (const int *)nullptr, kb, np, D, 0, bins[0], hist,
                       nblocks, nbad);
    VPIC_LAUNCHED("k_rs_hist");
    int h_nbad = 0;
    VPIC_HIP(hipMemcpyAsync(&h_nbad, nbad, sizeof(int), hipMemcpyDeviceToHost,
                            s));
    VPIC_HIP(hipStreamSynchronize(s));
    if (h_nbad > 0)
      return vpic_set_error(VPIC_EINVAL,
                            "sort_p: %d particles have a voxel index outside "
                            "[0,nv); particles left unsorted", h_nbad);
  }
  VPIC_LAUNCHED("k_rs_hist");
  if ((rc = device_exclusive_scan(hist, hist, (int64_t)bins[0] * nblocks,
                                  sums, s)))
    return rc;
  hipLaunchKernelGGL(k_rs_scatter<true>, dim3(nblocks), dim3(64), 0, s,
                     hint ? hint : (const int *)kb, (const int *)nullptr, np,
                     D, 0, bins[0], nbits_for(bins[0]), (const int *)hist,
                     nblocks, ka, ia);
  VPIC_LAUNCHED("k_rs_scatter");
  // passes y and z: (ka, ia) -> (kb, ib) -> (ka, ia)
  for (int axis = 1; axis < 3; axis++) {
    const int *kin = axis == 1 ? ka : kb, *iin = axis == 1 ? ia : ib;
    int *kout = axis == 1 ? kb : ka, *iout = axis == 1 ? ib : ia;
    hipLaunchKernelGGL(k_rs_hist<1>, dim3(nblocks), dim3(64), 0, s,
                       (const vpic_particle_t *)nullptr, kin, (int *)nullptr,
                       np, D, axis, bins[axis], hist, nblocks, nbad);
    VPIC_LAUNCHED("k_rs_hist");
    if ((rc = device_exclusive_scan(hist, hist,
                                    (int64_t)bins[axis] * nblocks, sums, s)))
      return rc;
    hipLaunchKernelGGL(k_rs_scatter<false>, dim3(nblocks), dim3(64), 0, s,
                       kin, iin, np, D, axis, bins[axis], nbits_for(bins[axis]),
                       (const int *)hist, nblocks, kout, iout);
    VPIC_LAUNCHED("k_rs_scatter");
  }
  hipLaunchKernelGGL(k_gather_part<GATHER_GU>,
                     dim3(blocks_for((int64_t)np + 1, 128 * GATHER_GU)),
                     dim3(256), 0, s, p, (const int *)ia, (const int *)ka, np,
                     nv, aux, partition, hint ? nbad + 1 : (int *)nullptr);
  VPIC_LAUNCHED("k_gather_part");
  if (hint) {
    int h[2] = {0, 0};
    VPIC_HIP(hipMemcpyAsync(h, nbad, sizeof(h), hipMemcpyDeviceToHost, s));
    VPIC_HIP(hipStreamSynchronize(s));
    if (h[0] > 0 || h[1] > 0) return SORT_KEYS_STALE;
  }
  if (in_place)
    VPIC_HIP(hipMemcpyAsync(p, aux, sizeof(vpic_particle_t) * (size_t)np,
                            hipMemcpyDeviceToDevice, s));
  return VPIC_OK;
}

// The counting sort with global atomics (any grid).
int sort_p_counting(vpic_particle_t *p, vpic_particle_t *aux, int np,
                    int *partition, const vpic_grid_t *g, int in_place,
                    hipStream_t s) {
  SortScratch &S = sort_scratch();
  const int nv = g->nv;
  const int64_t nslots = (int64_t)nv + 1;
  int rc;
  if ((rc = scratch_reserve(S.keys, sizeof(int) * (size_t)(np > 0 ? np : 1))) ||
      (rc = scratch_reserve(S.perm, sizeof(int) * (size_t)(np > 0 ? np : 1))) ||
      (rc = scratch_reserve(S.count, sizeof(int) * (size_t)nslots)) ||
      (rc = scratch_reserve(S.sums,
                            sizeof(int) * (size_t)(blocks_for(nslots, SCAN_ITEMS) + 1))) ||
      (rc = scratch_reserve(S.big, sizeof(int) * (size_t)nv)) ||
      (rc = scratch_reserve(S.nbig, 64)))
    return rc;
  int *keys = (int *)S.keys.ptr, *perm = (int *)S.perm.ptr;
  int *count = (int *)S.count.ptr, *sums = (int *)S.sums.ptr;
  int *big = (int *)S.big.ptr, *nbig = (int *)S.nbig.ptr;

  int fb = blocks_for(nslots, 256);
  if (fb > 65536) fb = 65536;
  hipLaunchKernelGGL(k_fill_int, dim3(fb), dim3(256), 0, s, count, nslots, 0);
  VPIC_LAUNCHED("k_fill_int");
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, nbig, (int64_t)3, 0);
  VPIC_LAUNCHED("k_fill_int");
  if (np > 0) {
    hipLaunchKernelGGL(k_count, dim3(blocks_for(np, 256)), dim3(256), 0, s, p,
                       np, nv, keys, count, nbig + 2);
    VPIC_LAUNCHED("k_count");
  }
  // partition = exclusive scan of the counts over nv+1 slots; count[nv] = 0
  // so partition[nv] = np.
  if ((rc = device_exclusive_scan(count, partition, nslots, sums, s)))
    return rc;
  if (np == 0) return VPIC_OK;
  // cursor = partition (reuse the count array)
  VPIC_HIP(hipMemcpyAsync(count, partition, sizeof(int) * (size_t)nslots,
                          hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_scatter, dim3(blocks_for(np, 256)), dim3(256), 0, s,
                     keys, np, count, perm);
  VPIC_LAUNCHED("k_scatter");
  hipLaunchKernelGGL(k_fixup, dim3(nv), dim3(64), 0, s,
                     (const int *)partition, nv, perm, big, nbig);
  VPIC_LAUNCHED("k_fixup");
  // Long segments (rare): size the padded scratch for the longest possible
  // segment only when some exist.
  int h_nbig[3] = {0, 0, 0};
  VPIC_HIP(hipMemcpyAsync(h_nbig, nbig, sizeof(h_nbig), hipMemcpyDeviceToHost,
                          s));
  VPIC_HIP(hipStreamSynchronize(s));
  if (h_nbig[2] > 0)
    return vpic_set_error(VPIC_EINVAL,
                          "sort_p: %d particles have a voxel index outside "
                          "[0,nv); particles left unsorted", h_nbig[2]);
  if (h_nbig[0] > 0) {
    int64_t pad = 1;
    while (pad < h_nbig[1]) pad <<= 1;
    const int nblk = h_nbig[0] < 64 ? h_nbig[0] : 64;
    if ((rc = scratch_reserve(S.tmp, sizeof(int) * (size_t)(pad * nblk))))
      return rc;
    hipLaunchKernelGGL(k_fixup_big, dim3(nblk), dim3(1024), 0, s,
                       (const int *)partition, perm, (const int *)big,
                       (const int *)nbig, (int *)S.tmp.ptr, pad);
    VPIC_LAUNCHED("k_fixup_big");
  }
  hipLaunchKernelGGL(k_gather, dim3(blocks_for(np, 256)), dim3(256), 0, s, p,
                     perm, np, aux);
  VPIC_LAUNCHED("k_gather");
  if (in_place)
    VPIC_HIP(hipMemcpyAsync(p, aux, sizeof(vpic_particle_t) * (size_t)np,
                            hipMemcpyDeviceToDevice, s));
  return VPIC_OK;
}

}  // namespace

// --------------------------------------------------------------------------
// Batched sort of an ensemble's boxes (no host synchronisation, so a graph
// can hold it).  The key of a particle is its box's voxel offset plus its
// voxel, so one counting sort over the packed arrays sorts every box within
// its own particle range (the keys of box b lie between those of boxes b-1
// and b+1, and every box keeps its particle count).  A voxel outside the
// box is counted in nbad and sorted as the box's voxel 0.
// --------------------------------------------------------------------------
namespace {

__device__ __forceinline__ int ens_key(const vpic_particle_t &p,
                                       const EnsPushBox &b, int *nbad) {
  const int v = p.i;
  if ((unsigned)v >= (unsigned)b.G.nv) {
    atomicAdd(nbad, 1);
    return (int)b.voff;
  }
  return (int)b.voff + v;
}

// Runs of equal keys over consecutive lanes: `len` for the run's first
// lane (0 elsewhere) and each lane's rank in its run.
__device__ __forceinline__ void key_runs(int key, bool valid, int lane,
                                         int &len, int &rank) {
  const int kp = __shfl_up(key, 1);
  const bool head = valid && (lane == 0 || kp != key);
  const uint64_t heads = __ballot(head);
  const uint64_t act = __ballot(valid);
  const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
  const int h0 = 63 - __clzll(heads & upto);  // this lane's run head
  rank = lane - h0;
  // the next head after this lane (or the end of the valid lanes)
  const uint64_t later = heads & ~upto;
  const int nxt = later ? __ffsll((unsigned long long)later) - 1
                        : 64 - __clzll(act);
  len = head ? nxt - lane : 0;
}

__global__ __launch_bounds__(256) void k_ens_count(
    const vpic_particle_t *__restrict__ P, const EnsPushBox *__restrict__ boxes,
    const EnsItem *__restrict__ items, int *__restrict__ keys,
    int *__restrict__ count, int *__restrict__ nbad) {
  const EnsItem I = items[blockIdx.x];
  const EnsPushBox b = boxes[I.box];
  const int lane = lane_id();
  const int end = I.first + I.count;
  for (int base = I.first; base < end; base += blockDim.x) {
    const int n = base + threadIdx.x;
    const bool valid = n < end;
    int key = -1;
    if (valid) {
      const int64_t g = b.poff + n;
      key = ens_key(P[g], b, nbad);
      keys[g] = key;
    }
    int len, rank;
    key_runs(key, valid, lane, len, rank);
    if (len) atomicAdd(count + key, len);
  }
}

__global__ __launch_bounds__(256) void k_ens_scatter(
    const vpic_particle_t *__restrict__ P, const EnsPushBox *__restrict__ boxes,
    const EnsItem *__restrict__ items, const int *__restrict__ keys,
    int *__restrict__ cursor, vpic_particle_t *__restrict__ aux) {
  const EnsItem I = items[blockIdx.x];
  const EnsPushBox b = boxes[I.box];
  const int lane = lane_id();
  const int end = I.first + I.count;
  for (int base = I.first; base < end; base += blockDim.x) {
    const int n = base + threadIdx.x;
    const bool valid = n < end;
    const int64_t g = b.poff + n;
    const int key = valid ? keys[g] : -1;
    int len, rank;
    key_runs(key, valid, lane, len, rank);
    int at = 0;
    if (len) at = atomicAdd(cursor + key, len);
    // the run head's slot, to every lane of the run
    const int kp = __shfl_up(key, 1);
    const bool head = valid && (lane == 0 || kp != key);
    const uint64_t heads = __ballot(head);
    const uint64_t upto = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    const int h0 = 63 - __clzll(heads & upto);
    at = __shfl(at, h0 < 0 ? 0 : h0);
    if (valid) {
      const float4 *src = reinterpret_cast<const float4 *>(P + g);
      float4 *dst = reinterpret_cast<float4 *>(aux + at + rank);
      dst[0] = src[0];
      dst[1] = src[1];
    }
  }
}

// Box b's partition, box-relative, at epart + voff_b + b (nv_b + 1 ints):
// the offsets of its voxels in the packed order less the box's first
// particle (start holds the exclusive scan over the packed keys, whose entry
// after box b's last voxel is box b + 1's first particle).
__global__ __launch_bounds__(256) void k_ens_part(
    const EnsPushBox *__restrict__ boxes, const int *__restrict__ start,
    int *__restrict__ epart) {
  const int b = blockIdx.x;
  const EnsPushBox B = boxes[b];
  const int base = (int)B.voff;
  const int first = (int)B.poff;
  for (int v = threadIdx.x; v <= B.G.nv; v += blockDim.x)
    epart[base + b + v] = start[base + v] - first;
}

}  // namespace

int launch_ens_sort(vpic_particle_t *P, vpic_particle_t *aux,
                    const EnsPushBox *boxes, const EnsItem *items, int nitems,
                    int nbox, int64_t np_total, int64_t nv_total, int *keys,
                    int *count, int *start, int *sums, int *nbad, int *epart,
                    hipStream_t s) {
  if (nitems <= 0 || np_total <= 0) return VPIC_OK;
  const int64_t nslots = nv_total + 1;
  int fb = blocks_for(nslots, 256);
  if (fb > 65536) fb = 65536;
  hipLaunchKernelGGL(k_fill_int, dim3(fb), dim3(256), 0, s, count, nslots, 0);
  VPIC_LAUNCHED("k_fill_int");
  hipLaunchKernelGGL(k_ens_count, dim3(nitems), dim3(256), 0, s, P, boxes,
                     items, keys, count, nbad);
  VPIC_LAUNCHED("k_ens_count");
  int rc;
  if ((rc = device_exclusive_scan(count, start, nslots, sums, s))) return rc;
  if (epart) {
    hipLaunchKernelGGL(k_ens_part, dim3(nbox), dim3(256), 0, s, boxes,
                       (const int *)start, epart);
    VPIC_LAUNCHED("k_ens_part");
  }
  hipLaunchKernelGGL(k_ens_scatter, dim3(nitems), dim3(256), 0, s, P, boxes,
                     items, (const int *)keys, start, aux);
  VPIC_LAUNCHED("k_ens_scatter");
  VPIC_HIP(hipMemcpyAsync(P, aux, sizeof(vpic_particle_t) * (size_t)np_total,
                          hipMemcpyDeviceToDevice, s));
  return VPIC_OK;
}

// sort_p: the axis passes when every axis fits their LDS counters, else the
// counting sort.
int launch_sort_p(vpic_particle_t *p, vpic_particle_t *aux, int np,
                  int *partition, const vpic_grid_t *g, int in_place,
                  hipStream_t s) {
  push_invalidate_bins(p);
  push_invalidate_bins(aux);
  if (np > 0 && g->nx + 2 <= RS_MAXB && g->ny + 2 <= RS_MAXB &&
      g->nz + 2 <= RS_MAXB)
    return sort_p_axes(p, aux, np, partition, g, in_place, nullptr, s);
  return sort_p_counting(p, aux, np, partition, g, in_place, s);
}

// sort_p with the caller's copy of the voxel keys (keys[k] = p[k].i, e.g.
// written by the push before): the same result, from one pass fewer over
// the particle records when the keys are right, and from the particles
// when any is not.
int launch_sort_p_keys(vpic_particle_t *p, vpic_particle_t *aux, int np,
                       int *partition, const int *keys, const vpic_grid_t *g,
                       int in_place, hipStream_t s) {
  if (!keys || !(np > 0 && g->nx + 2 <= RS_MAXB && g->ny + 2 <= RS_MAXB &&
                 g->nz + 2 <= RS_MAXB))
    return launch_sort_p(p, aux, np, partition, g, in_place, s);
  push_invalidate_bins(p);
  push_invalidate_bins(aux);
  const int rc = sort_p_axes(p, aux, np, partition, g, in_place, keys, s);
  if (rc != SORT_KEYS_STALE) return rc;
  note_sort_keys_stale();
  return sort_p_axes(p, aux, np, partition, g, in_place, nullptr, s);
}

long long sort_keys_stale_count() { return g_sort_keys_stale.load(); }

