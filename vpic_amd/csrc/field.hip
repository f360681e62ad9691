// field.hip -- voxel kernels of the step on gfx950: load_interpolator,
// clear/unload of the accumulator, clear_jf / synchronize_jf, advance_b,
// vacuum_advance_e (with its tangential-B ghosts and local adjustments) and
// the FP64 field energy.
//
// field_t (80 B) is read and written as five 16-B vectors:
//   [0] ex ey ez div_e_err   [1] cbx cby cbz div_b_err
//   [2] tcax tcay tcaz rhob  [3] jfx jfy jfz rhof   [4] material ids
// One lane per voxel, consecutive lanes on consecutive x, so every vector
// access of a wavefront is a contiguous 5 KiB span.
//
// Each stencil is the reference's expression with the same association, so
// the results are bit-identical:
//   load_interpolator  interpolator_array_pipeline.cc:66-121
//   unload_accumulator unload_accumulator_pipeline.cc:65-67,137-139
//   advance_b          advance_b_pipeline.h:26-28,57-59
//   vacuum_advance_e   vacuum_advance_e_pipeline.h:27-33,56-69
//   ghosts / adjusts   standard/local.cc:50-122,224-296,336-368,
//                      standard/remote.cc:61-134,418-507
//   energy_f           vacuum_energy_f_pipeline.h:17-75, .cc:67-96
#include "common.h"

namespace {

struct Box {
  int x0, y0, z0, nxr, nyr, nzr;
};

__device__ __forceinline__ bool box_coords(const Box &b, int64_t l, int &x,
                                           int &y, int &z) {
  if (l >= (int64_t)b.nxr * b.nyr * b.nzr) return false;
  int t = (int)(l / b.nxr);
  x = b.x0 + (int)(l - (int64_t)t * b.nxr);
  z = t / b.nyr;
  y = b.y0 + (t - z * b.nyr);
  z += b.z0;
  return true;
}

inline int blocks_for(int64_t n, int bs) { return (int)((n + bs - 1) / bs); }

// Block order of the field-advance kernels (advance_b, advance_e), which
// read the E or B of the +/-y and +/-z neighbours.  Workgroups go to the
// eight XCDs in turn (workgroup i to XCD i mod 8), each XCD with its own L2;
// numbered as below, XCD k runs the k-th contiguous eighth of the blocks,
// so the neighbour rows and planes a block reads are mostly in the L2 of
// the XCD that processes them (advance_b 20% and advance_e 9% faster on a
// 256x128x128 grid; load_interpolator and unload_accumulator, measured the
// same way, gained nothing and keep the plain order).  The grid is
// xcd_grid(n) blocks, a multiple of eight.
constexpr int NXCD = 8;
inline int xcd_grid(int64_t n, int bs) {
  return (blocks_for(n, bs) + NXCD - 1) / NXCD * NXCD;
}
__device__ __forceinline__ int64_t xcd_index() {
  const int per = gridDim.x / NXCD;
  const int blk = (int)(blockIdx.x % NXCD) * per + (int)(blockIdx.x / NXCD);
  return (int64_t)blk * blockDim.x + threadIdx.x;
}

__device__ __forceinline__ float4 *F4(vpic_field_t *f, int v) {
  return reinterpret_cast<float4 *>(f + v);
}
__device__ __forceinline__ const float4 *F4(const vpic_field_t *f, int v) {
  return reinterpret_cast<const float4 *>(f + v);
}

// ---------------------------------------------------------------------------
// load_interpolator_array
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_interpolator_at(
    vpic_interpolator_t *__restrict__ fi, const vpic_field_t *__restrict__ f,
    int x, int y, int z, int nx, int ny) {
  const float fourth = 0.25f, half = 0.50f;
  const int v = vox(x, y, z, nx, ny);
  const int sy = nx + 2, sz = (ny + 2) * sy;
  const float4 e0 = F4(f, v)[0], b0 = F4(f, v)[1];
  const float4 ex_ = F4(f, v + 1)[0], bx_ = F4(f, v + 1)[1];
  const float4 ey_ = F4(f, v + sy)[0], by_ = F4(f, v + sy)[1];
  const float4 ez_ = F4(f, v + sz)[0], bz_ = F4(f, v + sz)[1];
  const float e_yz = F4(f, v + sy + sz)[0].x;  // pfyz->ex
  const float e_zx = F4(f, v + sz + 1)[0].y;   // pfzx->ey
  const float e_xy = F4(f, v + 1 + sy)[0].z;   // pfxy->ez
  float w0, w1, w2, w3;
  float4 o0, o1, o2, o3;
  w0 = e0.x; w1 = ey_.x; w2 = ez_.x; w3 = e_yz;
  o0.x = fourth * ((w3 + w0) + (w1 + w2));
  o0.y = fourth * ((w3 - w0) + (w1 - w2));
  o0.z = fourth * ((w3 - w0) - (w1 - w2));
  o0.w = fourth * ((w3 + w0) - (w1 + w2));
  w0 = e0.y; w1 = ez_.y; w2 = ex_.y; w3 = e_zx;
  o1.x = fourth * ((w3 + w0) + (w1 + w2));
  o1.y = fourth * ((w3 - w0) + (w1 - w2));
  o1.z = fourth * ((w3 - w0) - (w1 - w2));
  o1.w = fourth * ((w3 + w0) - (w1 + w2));
  w0 = e0.z; w1 = ex_.z; w2 = ey_.z; w3 = e_xy;
  o2.x = fourth * ((w3 + w0) + (w1 + w2));
  o2.y = fourth * ((w3 - w0) + (w1 - w2));
  o2.z = fourth * ((w3 - w0) - (w1 - w2));
  o2.w = fourth * ((w3 + w0) - (w1 + w2));
  w0 = b0.x; w1 = bx_.x;
  o3.x = half * (w1 + w0);
  o3.y = half * (w1 - w0);
  w0 = b0.y; w1 = by_.y;
  o3.z = half * (w1 + w0);
  o3.w = half * (w1 - w0);
  w0 = b0.z; w1 = bz_.z;
  float4 *out = reinterpret_cast<float4 *>(fi + v);
  out[0] = o0;
  out[1] = o1;
  out[2] = o2;
  out[3] = o3;
  reinterpret_cast<float2 *>(out + 4)[0] =
      make_float2(half * (w1 + w0), half * (w1 - w0));
}

__global__ __launch_bounds__(256) void k_load_interpolator(
    vpic_interpolator_t *__restrict__ fi, const vpic_field_t *__restrict__ f,
    Box b, int nx, int ny) {
  int x, y, z;
  if (!box_coords(b, blockIdx.x * (int64_t)blockDim.x + threadIdx.x, x, y, z))
    return;
  load_interpolator_at(fi, f, x, y, z, nx, ny);
}

// ---------------------------------------------------------------------------
// accumulators and jf
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_fill_zero4(float4 *a, int64_t n4) {
  for (int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; k < n4;
       k += (int64_t)gridDim.x * blockDim.x)
    a[k] = make_float4(0.f, 0.f, 0.f, 0.f);
}

__global__ __launch_bounds__(256) void k_clear_jf(vpic_field_t *f, int nv,
                                                  int rhof) {
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nv;
       v += gridDim.x * blockDim.x) {
    if (rhof) {
      f[v].rhof = 0.0f;
    } else {
      f[v].jfx = 0.0f;
      f[v].jfy = 0.0f;
      f[v].jfz = 0.0f;
    }
  }
}

// jf of voxel v (interior or on the upper faces) plus its quarter-face
// currents (unload_accumulator_pipeline.cc:18-82).
__device__ __forceinline__ float4 unload_add(
    float4 jf, const vpic_accumulator_t *__restrict__ a, int v, int sy,
    int sz, float cx, float cy, float cz) {
  const vpic_accumulator_t *a0 = a + v, *ax = a + v - 1, *ay = a + v - sy,
                           *az = a + v - sz, *ayz = a + v - sy - sz,
                           *azx = a + v - 1 - sz, *axy = a + v - 1 - sy;
  jf.x += cx * (a0->jx[0] + ay->jx[1] + az->jx[2] + ayz->jx[3]);
  jf.y += cy * (a0->jy[0] + az->jy[1] + ax->jy[2] + azx->jy[3]);
  jf.z += cz * (a0->jz[0] + ax->jz[1] + ay->jz[2] + axy->jz[3]);
  return jf;
}

__device__ __forceinline__ void unload_accumulator_at(
    vpic_field_t *__restrict__ f, const vpic_accumulator_t *__restrict__ a,
    int x, int y, int z, int nx, int ny, float cx, float cy, float cz) {
  const int v = vox(x, y, z, nx, ny);
  const int sy = nx + 2, sz = (ny + 2) * sy;
  F4(f, v)[3] = unload_add(F4(f, v)[3], a, v, sy, sz, cx, cy, cz);
}

// clear_jf followed by unload_accumulator_array in one pass over the field
// records: jf of every voxel set to 0, then, on voxels 1..n+1 per axis,
// 0 + the unload's sum (the += of the unload on a cleared jf, so the two
// calls' result bit for bit), rhof kept.
__global__ __launch_bounds__(256) void k_clear_unload_accumulator(
    vpic_field_t *__restrict__ f, const vpic_accumulator_t *__restrict__ a,
    int nv, int nx, int ny, int nz, float cx, float cy, float cz) {
  const int sy = nx + 2, sz = (ny + 2) * sy;
  for (int v = blockIdx.x * blockDim.x + threadIdx.x; v < nv;
       v += gridDim.x * blockDim.x) {
    const int x = v % sy, t = v / sy, y = t % (ny + 2), z = t / (ny + 2);
    float4 jf = F4(f, v)[3];
    jf.x = 0.0f;
    jf.y = 0.0f;
    jf.z = 0.0f;
    if (x >= 1 && x <= nx + 1 && y >= 1 && y <= ny + 1 && z >= 1 &&
        z <= nz + 1)
      jf = unload_add(jf, a, v, sy, sz, cx, cy, cz);
    F4(f, v)[3] = jf;
  }
}

__global__ __launch_bounds__(256) void k_unload_accumulator(
    vpic_field_t *__restrict__ f, const vpic_accumulator_t *__restrict__ a,
    Box b, int nx, int ny, float cx, float cy, float cz) {
  int x, y, z;
  if (!box_coords(b, blockIdx.x * (int64_t)blockDim.x + threadIdx.x, x, y, z))
    return;
  unload_accumulator_at(f, a, x, y, z, nx, ny, cx, cy, cz);
}

// Rotated face addressing: normal axis X, tangential Y=(X+1)%3, Z=(X+2)%3.
struct FaceGeom {
  int n[3];  // nx, ny, nz
  int s[3];  // strides 1, sy, sz
};

__device__ __forceinline__ int face_vox(const FaceGeom &G, int X, int plane,
                                        int yy, int zz) {
  const int Y = (X + 1) % 3, Z = (X + 2) % 3;
  return plane * G.s[X] + yy * G.s[Y] + zz * G.s[Z];
}

__device__ __forceinline__ float &fcomp(vpic_field_t *f, int v, int block,
                                        int axis) {
  return reinterpret_cast<float *>(f + v)[4 * block + axis];
}

// local_adjust_jf (local.cc:336-368) on one non-periodic face k (order
// -x,-y,-z,+x,+y,+z).  Faces are launched one after another, as the
// reference applies them, because edges shared by two faces are scaled by
// both.  mode: 1 = zero (anti-symmetric), 2 = double (the other local BCs).
__global__ void k_adjust_jf(vpic_field_t *f, FaceGeom G, int k, int mode) {
  const int X = k % 3, Y = (X + 1) % 3, Z = (X + 2) % 3;
  const int plane = k >= 3 ? G.n[X] + 1 : 1;
  const int wy = G.n[Y] + 1, wz = G.n[Z] + 1;
  for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < wy * wz;
       l += gridDim.x * blockDim.x) {
    const int yy = 1 + l % wy, zz = 1 + l / wy;
    const int v = face_vox(G, X, plane, yy, zz);
    if (yy <= G.n[Y]) {  // Y##Z## edges: Y in 1..nY, Z in 1..nZ+1
      float &j = fcomp(f, v, 3, Y);
      if (mode == 1) j = 0; else j *= 2.f;
    }
    if (zz <= G.n[Z]) {  // Z##Y## edges: Y in 1..nY+1, Z in 1..nZ
      float &j = fcomp(f, v, 3, Z);
      if (mode == 1) j = 0; else j *= 2.f;
    }
  }
}

// One axis of the periodic jf fold (remote.cc:474-502 with lw = rw = 1).
__device__ __forceinline__ void fold_jf_at(vpic_field_t *f, const FaceGeom &G,
                                           int X, int l) {
  const int Y = (X + 1) % 3, Z = (X + 2) % 3;
  const int wy = G.n[Y] + 1;
  const float lw = 1.0f, rw = 1.0f;
  {
    const int yy = 1 + l % wy, zz = 1 + l / wy;
    const int lo = face_vox(G, X, 1, yy, zz);
    const int hi = face_vox(G, X, G.n[X] + 1, yy, zz);
    if (yy <= G.n[Y]) {
      float &a = fcomp(f, lo, 3, Y), &b = fcomp(f, hi, 3, Y);
      const float sa = a, sb = b;
      b = lw * sb + rw * sa;
      a = lw * sa + rw * sb;
    }
    if (zz <= G.n[Z]) {
      float &a = fcomp(f, lo, 3, Z), &b = fcomp(f, hi, 3, Z);
      const float sa = a, sb = b;
      b = lw * sb + rw * sa;
      a = lw * sa + rw * sb;
    }
  }
}

__global__ void k_fold_jf(vpic_field_t *f, FaceGeom G, int X) {
  const int Y = (X + 1) % 3, Z = (X + 2) % 3;
  const int wy = G.n[Y] + 1, wz = G.n[Z] + 1;
  for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < wy * wz;
       l += gridDim.x * blockDim.x)
    fold_jf_at(f, G, X, l);
}

// local_adjust_rhof + local_adjust_rhob (local.cc:376-440) on one local face:
// mode 1 (anti-symmetric) zeroes rhof and rhob, mode 2 doubles rhof.
__global__ void k_adjust_rho(vpic_field_t *f, FaceGeom G, int k, int mode) {
  const int X = k % 3, Y = (X + 1) % 3, Z = (X + 2) % 3;
  const int plane = k >= 3 ? G.n[X] + 1 : 1;
  const int wy = G.n[Y] + 1, wz = G.n[Z] + 1;
  for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < wy * wz;
       l += gridDim.x * blockDim.x) {
    const int yy = 1 + l % wy, zz = 1 + l / wy;
    vpic_field_t &c = f[face_vox(G, X, plane, yy, zz)];
    if (mode == 1) {
      c.rhof = 0;
      c.rhob = 0;
    } else {
      c.rhof *= 2;
    }
  }
}

// One axis of the periodic node fold of synchronize_rho (remote.cc:534-625):
// rhof = lw*rhof + rw*remote with lw = rw = 1, rhob = hlw*rhob + hrw*remote
// with hlw = hrw = 1/2 (equal cell sizes).
__global__ void k_fold_rho(vpic_field_t *f, FaceGeom G, int X) {
  const int Y = (X + 1) % 3, Z = (X + 2) % 3;
  const int wy = G.n[Y] + 1, wz = G.n[Z] + 1;
  const float hlw = 0.5f, hrw = 0.5f, lw = hlw + hlw, rw = hrw + hrw;
  for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < wy * wz;
       l += gridDim.x * blockDim.x) {
    const int yy = 1 + l % wy, zz = 1 + l / wy;
    vpic_field_t &a = f[face_vox(G, X, 1, yy, zz)];
    vpic_field_t &b = f[face_vox(G, X, G.n[X] + 1, yy, zz)];
    const float af = a.rhof, bf = b.rhof, ab = a.rhob, bb = b.rhob;
    b.rhof = lw * bf + rw * af;
    b.rhob = hlw * bb + hrw * ab;
    a.rhof = lw * af + rw * bf;
    a.rhob = hlw * ab + hrw * bb;
  }
}

// ---------------------------------------------------------------------------
// advance_b
// ---------------------------------------------------------------------------
__device__ __forceinline__ void advance_b_at(vpic_field_t *__restrict__ f,
                                             int x, int y, int z, int nx,
                                             int ny, int nz, float px,
                                             float py, float pz) {
  const int v = vox(x, y, z, nx, ny);
  const int sy = nx + 2, sz = (ny + 2) * sy;
  const float4 e0 = F4(f, v)[0];
  float4 b0 = F4(f, v)[1];
  const bool ux = x <= nx, uy = y <= ny, uz = z <= nz;
  // cbx on x in 1..nx+1 (y<=ny, z<=nz), cby on y in 1..ny+1, cbz on z in
  // 1..nz+1; each neighbour is read only where its component is updated.
  float4 ex1 = e0, ey1 = e0, ez1 = e0;
  if (ux && (uy || uz)) ex1 = F4(f, v + 1)[0];
  if (uy && (ux || uz)) ey1 = F4(f, v + sy)[0];
  if (uz && (ux || uy)) ez1 = F4(f, v + sz)[0];
  if (uy && uz) b0.x -= (py * (ey1.z - e0.z) - pz * (ez1.y - e0.y));
  if (ux && uz) b0.y -= (pz * (ez1.x - e0.x) - px * (ex1.z - e0.z));
  if (ux && uy) b0.z -= (px * (ex1.y - e0.y) - py * (ey1.x - e0.x));
  if ((uy && uz) || (ux && uz) || (ux && uy)) F4(f, v)[1] = b0;
}

__global__ __launch_bounds__(256) void k_advance_b(vpic_field_t *__restrict__ f,
                                                   Box b, int nx, int ny,
                                                   int nz, float px, float py,
                                                   float pz) {
  int x, y, z;
  if (!box_coords(b, xcd_index(), x, y, z)) return;
  advance_b_at(f, x, y, z, nx, ny, nz, px, py, pz);
}

// local_adjust_norm_b (local.cc:266-296): symmetric faces zero cb normal.
__global__ void k_adjust_norm_b(vpic_field_t *f, FaceGeom G, int4 m0,
                                int2 m1) {
  const int k = blockIdx.y;
  const int mode = k == 0 ? m0.x : k == 1 ? m0.y : k == 2 ? m0.z
                   : k == 3 ? m0.w : k == 4 ? m1.x : m1.y;
  if (mode != VPIC_BC_SYMMETRIC_FIELDS) return;
  const int X = k % 3, Y = (X + 1) % 3, Z = (X + 2) % 3;
  const int plane = k >= 3 ? G.n[X] + 1 : 1;
  for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < G.n[Y] * G.n[Z];
       l += gridDim.x * blockDim.x) {
    const int yy = 1 + l % G.n[Y], zz = 1 + l / G.n[Y];
    fcomp(f, face_vox(G, X, plane, yy, zz), 1, X) = 0;
  }
}

// ---------------------------------------------------------------------------
// vacuum_advance_e
// ---------------------------------------------------------------------------

// Tangential-B ghosts of one normal axis X (both sides): periodic partners
// take the opposite interior face with rw = 1, lw = 0 (remote.cc:105-116 for
// equal cell sizes); local faces copy (anti-symmetric) or negate
// (symmetric/pmc) the first interior plane (local.cc:67-81).
__device__ __forceinline__ void ghost_tang_b_at(vpic_field_t *f,
                                                const FaceGeom &G, int X,
                                                int mlo, int mhi, int l) {
  const int Y = (X + 1) % 3, Z = (X + 2) % 3;
  const int nX = G.n[X];
  const int wy = G.n[Y] + 1;
  const float rw = 1.0f, lw = 0.0f;
  {
    const int yy = 1 + l % wy, zz = 1 + l / wy;
    const int g0 = face_vox(G, X, 0, yy, zz);
    const int g1 = face_vox(G, X, nX + 1, yy, zz);
    const int i1 = face_vox(G, X, 1, yy, zz);
    const int iN = face_vox(G, X, nX, yy, zz);
#pragma unroll
    for (int c = 0; c < 2; c++) {
      const int comp = c == 0 ? Y : Z;
      // cb_Y on (Y in 1..nY+1, Z in 1..nZ), cb_Z on (Y in 1..nY, Z in 1..nZ+1)
      if (c == 0 ? (zz > G.n[Z]) : (yy > G.n[Y])) continue;
      const float a1 = fcomp(f, i1, 1, comp), aN = fcomp(f, iN, 1, comp);
      // low ghost (plane 0)
      if (mlo == VPIC_BC_PERIODIC)
        fcomp(f, g0, 1, comp) = rw * aN + lw * a1;
      else if (mlo == VPIC_BC_ANTI_SYMMETRIC_FIELDS)
        fcomp(f, g0, 1, comp) = a1;
      else
        fcomp(f, g0, 1, comp) = -a1;
      // high ghost (plane nX+1)
      if (mhi == VPIC_BC_PERIODIC)
        fcomp(f, g1, 1, comp) = rw * a1 + lw * aN;
      else if (mhi == VPIC_BC_ANTI_SYMMETRIC_FIELDS)
        fcomp(f, g1, 1, comp) = aN;
      else
        fcomp(f, g1, 1, comp) = -aN;
    }
  }
}

__global__ void k_ghost_tang_b(vpic_field_t *f, FaceGeom G, int X, int mlo,
                               int mhi) {
  const int Y = (X + 1) % 3, Z = (X + 2) % 3;
  const int wy = G.n[Y] + 1, wz = G.n[Z] + 1;
  for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < wy * wz;
       l += gridDim.x * blockDim.x)
    ghost_tang_b_at(f, G, X, mlo, mhi, l);
}


__device__ __forceinline__ void advance_e_at(vpic_field_t *__restrict__ f,
                                             int x, int y, int z, int nx,
                                             int ny, int nz, const ECoef &c) {
  const int v = vox(x, y, z, nx, ny);
  const int sy = nx + 2, sz = (ny + 2) * sy;
  float4 e0 = F4(f, v)[0];
  const float4 b0 = F4(f, v)[1];
  float4 t0 = F4(f, v)[2];
  const float4 j0 = F4(f, v)[3];
  const float4 bx = F4(f, v - 1)[1], by = F4(f, v - sy)[1],
               bz = F4(f, v - sz)[1];
  if (x <= nx) {
    t0.x = (c.py_muz * (b0.z - by.z) - c.pz_muy * (b0.y - bz.y)) -
           c.damp * t0.x;
    e0.x = c.decayx * e0.x + c.drivex * (t0.x - c.cj * j0.x);
  }
  if (y <= ny) {
    t0.y = (c.pz_mux * (b0.x - bz.x) - c.px_muz * (b0.z - bx.z)) -
           c.damp * t0.y;
    e0.y = c.decayy * e0.y + c.drivey * (t0.y - c.cj * j0.y);
  }
  if (z <= nz) {
    t0.z = (c.px_muy * (b0.y - bx.y) - c.py_mux * (b0.x - by.x)) -
           c.damp * t0.z;
    e0.z = c.decayz * e0.z + c.drivez * (t0.z - c.cj * j0.z);
  }
  F4(f, v)[0] = e0;
  F4(f, v)[2] = t0;
}

__global__ __launch_bounds__(256) void k_advance_e(vpic_field_t *__restrict__ f,
                                                   Box b, int nx, int ny,
                                                   int nz, ECoef c) {
  int x, y, z;
  if (!box_coords(b, xcd_index(), x, y, z)) return;
  advance_e_at(f, x, y, z, nx, ny, nz, c);
}

// local_adjust_tang_e (local.cc:224-264): PEC faces zero e and tca tangential.
__global__ void k_adjust_tang_e(vpic_field_t *f, FaceGeom G, int4 m0,
                                int2 m1) {
  const int k = blockIdx.y;
  const int mode = k == 0 ? m0.x : k == 1 ? m0.y : k == 2 ? m0.z
                   : k == 3 ? m0.w : k == 4 ? m1.x : m1.y;
  if (mode != VPIC_BC_ANTI_SYMMETRIC_FIELDS) return;
  const int X = k % 3, Y = (X + 1) % 3, Z = (X + 2) % 3;
  const int plane = k >= 3 ? G.n[X] + 1 : 1;
  const int wy = G.n[Y] + 1, wz = G.n[Z] + 1;
  for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < wy * wz;
       l += gridDim.x * blockDim.x) {
    const int yy = 1 + l % wy, zz = 1 + l / wy;
    const int v = face_vox(G, X, plane, yy, zz);
    if (yy <= G.n[Y]) {
      fcomp(f, v, 0, Y) = 0;
      fcomp(f, v, 2, Y) = 0;
    }
    if (zz <= G.n[Z]) {
      fcomp(f, v, 0, Z) = 0;
      fcomp(f, v, 2, Z) = 0;
    }
  }
}

// ---------------------------------------------------------------------------
// energy_f: per-voxel FP32 terms, FP64 sums (per-block partials, then one
// block folds them in a fixed order).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_energy_f(const vpic_field_t *__restrict__ f,
                                                  Box b, int nx, int ny,
                                                  float4 q, float4 h,
                                                  double *__restrict__ part) {
  __shared__ double red[6][256];
  int x, y, z;
  double en[6] = {0, 0, 0, 0, 0, 0};
  if (box_coords(b, blockIdx.x * (int64_t)blockDim.x + threadIdx.x, x, y, z)) {
    const int v = vox(x, y, z, nx, ny);
    const int sy = nx + 2, sz = (ny + 2) * sy;
    const float4 e0 = F4(f, v)[0], b0 = F4(f, v)[1];
    const float4 ex_ = F4(f, v + 1)[0], bx_ = F4(f, v + 1)[1];
    const float4 ey_ = F4(f, v + sy)[0], by_ = F4(f, v + sy)[1];
    const float4 ez_ = F4(f, v + sz)[0], bz_ = F4(f, v + sz)[1];
    const float eyz = F4(f, v + sy + sz)[0].x;
    const float ezx = F4(f, v + sz + 1)[0].y;
    const float exy = F4(f, v + 1 + sy)[0].z;
    en[0] = (double)(q.x * (e0.x * e0.x + ey_.x * ey_.x + ez_.x * ez_.x +
                            eyz * eyz));
    en[1] = (double)(q.y * (e0.y * e0.y + ez_.y * ez_.y + ex_.y * ex_.y +
                            ezx * ezx));
    en[2] = (double)(q.z * (e0.z * e0.z + ex_.z * ex_.z + ey_.z * ey_.z +
                            exy * exy));
    en[3] = (double)(h.x * (b0.x * b0.x + bx_.x * bx_.x));
    en[4] = (double)(h.y * (b0.y * b0.y + by_.y * by_.y));
    en[5] = (double)(h.z * (b0.z * b0.z + bz_.z * bz_.z));
  }
  for (int k = 0; k < 6; k++) red[k][threadIdx.x] = en[k];
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s)
      for (int k = 0; k < 6; k++)
        red[k][threadIdx.x] += red[k][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 6) part[6 * (int64_t)blockIdx.x + threadIdx.x] =
      red[threadIdx.x][0];
}

__global__ __launch_bounds__(256) void k_fold_partials(const double *part,
                                                       int nblk, int width,
                                                       double *out) {
  __shared__ double red[256];
  for (int k = 0; k < width; k++) {
    double s = 0;
    for (int i = threadIdx.x; i < nblk; i += blockDim.x)
      s += part[(int64_t)width * i + k];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
      if (threadIdx.x < st) red[threadIdx.x] += red[threadIdx.x + st];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[k] = red[0];
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Divergence cleaning and initialization kernels of the vacuum field table
// (sfa.cc:203-210).  Same expressions and association as the reference, so
// the field results are bit-identical; the FP64 sums differ only in order.
// ---------------------------------------------------------------------------

// Normal-E ghosts of axis X (both sides) on the node planes 0 and nX+1
// (Y in 1..nY+1, Z in 1..nZ+1): periodic faces take the opposite interior
// plane with rw = 1, lw = 0 (remote.cc:137-207); local faces copy
// (anti-symmetric) or negate (symmetric, pmc) e and tca of the adjacent
// plane (local.cc:129-179).
__global__ void k_ghost_norm_e(vpic_field_t *f, FaceGeom G, int X, int mlo,
                               int mhi) {
  const int Y = (X + 1) % 3, Z = (X + 2) % 3;
  const int nX = G.n[X];
  const int wy = G.n[Y] + 1, wz = G.n[Z] + 1;
  const float rw = 1.0f, lw = 0.0f;
  for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < wy * wz;
       l += gridDim.x * blockDim.x) {
    const int yy = 1 + l % wy, zz = 1 + l / wy;
    const int g0 = face_vox(G, X, 0, yy, zz);
    const int g1 = face_vox(G, X, nX + 1, yy, zz);
    const int i1 = face_vox(G, X, 1, yy, zz);
    const int iN = face_vox(G, X, nX, yy, zz);
    const float e1 = fcomp(f, i1, 0, X), eN = fcomp(f, iN, 0, X);
    const float t1 = fcomp(f, i1, 2, X), tN = fcomp(f, iN, 2, X);
    if (mlo == VPIC_BC_PERIODIC) {
      fcomp(f, g0, 0, X) = rw * eN + lw * e1;
    } else if (mlo == VPIC_BC_ANTI_SYMMETRIC_FIELDS) {
      fcomp(f, g0, 0, X) = e1;
      fcomp(f, g0, 2, X) = t1;
    } else {
      fcomp(f, g0, 0, X) = -e1;
      fcomp(f, g0, 2, X) = -t1;
    }
    if (mhi == VPIC_BC_PERIODIC) {
      fcomp(f, g1, 0, X) = rw * e1 + lw * eN;
    } else if (mhi == VPIC_BC_ANTI_SYMMETRIC_FIELDS) {
      fcomp(f, g1, 0, X) = eN;
      fcomp(f, g1, 2, X) = tN;
    } else {
      fcomp(f, g1, 0, X) = -eN;
      fcomp(f, g1, 2, X) = -tN;
    }
  }
}

// div_b_err ghosts of axis X (cell planes 0 and nX+1, Y in 1..nY, Z in
// 1..nZ): periodic as above (remote.cc:209-281); local faces copy (anti-
// symmetric) or negate (symmetric, pmc) (local.cc:182-221).
__global__ void k_ghost_div_b(vpic_field_t *f, FaceGeom G, int X, int mlo,
                              int mhi) {
  const int Y = (X + 1) % 3, Z = (X + 2) % 3;
  const int nX = G.n[X];
  const int wy = G.n[Y], wz = G.n[Z];
  const float rw = 1.0f, lw = 0.0f;
  for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < wy * wz;
       l += gridDim.x * blockDim.x) {
    const int yy = 1 + l % wy, zz = 1 + l / wy;
    vpic_field_t &g0 = f[face_vox(G, X, 0, yy, zz)];
    vpic_field_t &g1 = f[face_vox(G, X, nX + 1, yy, zz)];
    const float d1 = f[face_vox(G, X, 1, yy, zz)].div_b_err;
    const float dN = f[face_vox(G, X, nX, yy, zz)].div_b_err;
    g0.div_b_err = mlo == VPIC_BC_PERIODIC ? rw * dN + lw * d1
                   : mlo == VPIC_BC_ANTI_SYMMETRIC_FIELDS ? d1
                                                          : -d1;
    g1.div_b_err = mhi == VPIC_BC_PERIODIC ? rw * d1 + lw * dN
                   : mhi == VPIC_BC_ANTI_SYMMETRIC_FIELDS ? dN
                                                          : -dN;
  }
}

// Zero float slot `slot` of field_t on the node plane of face k
// (local_adjust_div_e, local.cc:299-334; local_adjust_rhob, 415-440).
__global__ void k_zero_node_face(vpic_field_t *f, FaceGeom G, int k,
                                 int slot) {
  const int X = k % 3, Y = (X + 1) % 3, Z = (X + 2) % 3;
  const int plane = k >= 3 ? G.n[X] + 1 : 1;
  const int wy = G.n[Y] + 1, wz = G.n[Z] + 1;
  for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < wy * wz;
       l += gridDim.x * blockDim.x) {
    const int yy = 1 + l % wy, zz = 1 + l / wy;
    reinterpret_cast<float *>(f + face_vox(G, X, plane, yy, zz))[slot] = 0;
  }
}

// Nodes 1..n+1: div_e_err (RHOB false: vacuum_compute_div_e_err_pipeline.h:
// 17-50) or rhob (RHOB true: vacuum_compute_rhob_pipeline.h:17-52).
template <bool RHOB>
__global__ __launch_bounds__(256) void k_div_e_nodes(vpic_field_t *__restrict__ f,
                                                     Box b, int nx, int ny,
                                                     float nc, float px,
                                                     float py, float pz,
                                                     float cj) {
  int x, y, z;
  if (!box_coords(b, blockIdx.x * (int64_t)blockDim.x + threadIdx.x, x, y, z))
    return;
  const int v = vox(x, y, z, nx, ny);
  const int sy = nx + 2, sz = (ny + 2) * sy;
  const float4 e0 = F4(f, v)[0];
  const float4 t0 = F4(f, v)[2];
  const float4 j0 = F4(f, v)[3];
  const float exm = F4(f, v - 1)[0].x, eym = F4(f, v - sy)[0].y,
              ezm = F4(f, v - sz)[0].z;
  if (RHOB) {
    reinterpret_cast<float *>(f + v)[11] =
        nc * (px * (e0.x - exm) + py * (e0.y - eym) + pz * (e0.z - ezm) -
              j0.w);
  } else {
    reinterpret_cast<float *>(f + v)[3] =
        nc * (px * (e0.x - exm) + py * (e0.y - eym) + pz * (e0.z - ezm) -
              cj * (j0.w + t0.w));
  }
}

// Edges: E += p (div_e_err(+1) - div_e_err) (vacuum_clean_div_e_pipeline.h:
// 45-47 over every edge of vacuum_clean_div_e_pipeline.cc).
__global__ __launch_bounds__(256) void k_clean_div_e(vpic_field_t *__restrict__ f,
                                                     Box b, int nx, int ny,
                                                     int nz, float px, float py,
                                                     float pz) {
  int x, y, z;
  if (!box_coords(b, blockIdx.x * (int64_t)blockDim.x + threadIdx.x, x, y, z))
    return;
  const int v = vox(x, y, z, nx, ny);
  const int sy = nx + 2, sz = (ny + 2) * sy;
  float4 e0 = F4(f, v)[0];
  const float d0 = e0.w;
  if (x <= nx) e0.x += px * (F4(f, v + 1)[0].w - d0);
  if (y <= ny) e0.y += py * (F4(f, v + sy)[0].w - d0);
  if (z <= nz) e0.z += pz * (F4(f, v + sz)[0].w - d0);
  F4(f, v)[0] = e0;
}

// Cells 1..n: div_b_err (compute_div_b_err_pipeline.cc:44-46).
__global__ __launch_bounds__(256) void k_div_b_cells(vpic_field_t *__restrict__ f,
                                                     Box b, int nx, int ny,
                                                     float px, float py,
                                                     float pz) {
  int x, y, z;
  if (!box_coords(b, blockIdx.x * (int64_t)blockDim.x + threadIdx.x, x, y, z))
    return;
  const int v = vox(x, y, z, nx, ny);
  const int sy = nx + 2, sz = (ny + 2) * sy;
  const float4 b0 = F4(f, v)[1];
  reinterpret_cast<float *>(f + v)[7] =
      px * (F4(f, v + 1)[1].x - b0.x) + py * (F4(f, v + sy)[1].y - b0.y) +
      pz * (F4(f, v + sz)[1].z - b0.z);
}

// Faces: cb += p (div_b_err - div_b_err(-1)) (clean_div_b_pipeline.h:
// 19-21 over every face of clean_div_b_pipeline.cc).
__global__ __launch_bounds__(256) void k_clean_div_b(vpic_field_t *__restrict__ f,
                                                     Box b, int nx, int ny,
                                                     int nz, float px, float py,
                                                     float pz) {
  int x, y, z;
  if (!box_coords(b, blockIdx.x * (int64_t)blockDim.x + threadIdx.x, x, y, z))
    return;
  const int v = vox(x, y, z, nx, ny);
  const int sy = nx + 2, sz = (ny + 2) * sy;
  float4 b0 = F4(f, v)[1];
  const float d0 = b0.w;
  if (y <= ny && z <= nz) b0.x += px * (d0 - F4(f, v - 1)[1].w);
  if (x <= nx && z <= nz) b0.y += py * (d0 - F4(f, v - sy)[1].w);
  if (x <= nx && y <= ny) b0.z += pz * (d0 - F4(f, v - sz)[1].w);
  F4(f, v)[1] = b0;
}

// Edges: tca = curl B (vacuum_compute_curl_b_pipeline.h:52-57).
__global__ __launch_bounds__(256) void k_curl_b(vpic_field_t *__restrict__ f,
                                                Box b, int nx, int ny, int nz,
                                                ECoef c) {
  int x, y, z;
  if (!box_coords(b, blockIdx.x * (int64_t)blockDim.x + threadIdx.x, x, y, z))
    return;
  const int v = vox(x, y, z, nx, ny);
  const int sy = nx + 2, sz = (ny + 2) * sy;
  const float4 b0 = F4(f, v)[1];
  float4 t0 = F4(f, v)[2];
  const float4 bx = F4(f, v - 1)[1], by = F4(f, v - sy)[1],
               bz = F4(f, v - sz)[1];
  if (x <= nx) t0.x = (c.py_muz * (b0.z - by.z) - c.pz_muy * (b0.y - bz.y));
  if (y <= ny) t0.y = (c.pz_mux * (b0.x - bz.x) - c.px_muz * (b0.z - bx.z));
  if (z <= nz) t0.z = (c.px_muy * (b0.y - bx.y) - c.py_mux * (b0.x - by.x));
  F4(f, v)[2] = t0;
}

// Per-block FP64 partial sums of the squared divergence errors: div_e_err
// on nodes 1..n+1 (interior products in FP32, boundary nodes in FP64 with
// weights 1/2, 1/4, 1/8; compute_rms_div_e_err_pipeline.cc:29-172) or
// div_b_err on cells 1..n (compute_rms_div_b_err_pipeline.cc:29-45).
template <bool DIVB>
__global__ __launch_bounds__(256) void k_sq_partials(const vpic_field_t *__restrict__ f,
                                                     Box b, int nx, int ny,
                                                     int nz,
                                                     double *__restrict__ part) {
  __shared__ double red[256];
  int x, y, z;
  double e = 0;
  if (box_coords(b, blockIdx.x * (int64_t)blockDim.x + threadIdx.x, x, y, z)) {
    const float d = reinterpret_cast<const float *>(
        f + vox(x, y, z, nx, ny))[DIVB ? 7 : 3];
    if (DIVB) {
      e = d * d;
    } else {
      const int nb = (x == 1 || x == nx + 1) + (y == 1 || y == ny + 1) +
                     (z == 1 || z == nz + 1);
      if (nb == 0) {
        e = d * d;
      } else {
        const double w = nb == 1 ? 0.5 : nb == 2 ? 0.25 : 0.125;
        e = w * (double)d * (double)d;
      }
    }
  }
  red[threadIdx.x] = e;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// synchronize_tang_e_norm_b on the periodic face pair of axis X
// (remote.cc:298-416): both faces take 0.5*(w1+w2) in FP64 of cb_X, e and
// tca tangential; the per-block sums of 2(w1-w2)^2 over cb and e (once
// from each face) go to part.
__global__ __launch_bounds__(256) void k_sync_tang_e_norm_b(vpic_field_t *f,
                                                            FaceGeom G, int X,
                                                            double *part) {
  __shared__ double red[256];
  const int Y = (X + 1) % 3, Z = (X + 2) % 3;
  const int wy = G.n[Y] + 1, wz = G.n[Z] + 1;
  const int hi = G.n[X] + 1;
  double err = 0;
  const int l = blockIdx.x * blockDim.x + threadIdx.x;
  if (l < wy * wz) {
    const int yy = 1 + l % wy, zz = 1 + l / wy;
    const int va = face_vox(G, X, 1, yy, zz), vb = face_vox(G, X, hi, yy, zz);
    // (block, component, valid on this edge/face, counts in err)
    const bool onY = yy <= G.n[Y], onZ = zz <= G.n[Z];
    const int blk[5] = {1, 0, 2, 0, 2};
    const int cmp[5] = {X, Y, Y, Z, Z};
    const bool ok[5] = {onY && onZ, onY, onY, onZ, onZ};
    const bool cnt[5] = {true, true, false, true, false};
#pragma unroll
    for (int q = 0; q < 5; q++) {
      if (!ok[q]) continue;
      float &a = fcomp(f, va, blk[q], cmp[q]);
      float &bb = fcomp(f, vb, blk[q], cmp[q]);
      const double w1 = a, w2 = bb;
      a = 0.5 * (w1 + w2);
      bb = 0.5 * (w2 + w1);
      if (cnt[q]) err += 2 * ((w1 - w2) * (w1 - w2));
    }
  }
  red[threadIdx.x] = err;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

FaceGeom face_geom(const vpic_grid_t *g) {
  FaceGeom G;
  G.n[0] = g->nx;
  G.n[1] = g->ny;
  G.n[2] = g->nz;
  G.s[0] = 1;
  G.s[1] = g->sy;
  G.s[2] = g->sz;
  return G;
}

int face_blocks(const vpic_grid_t *g) {
  int m = g->nx > g->ny ? g->nx : g->ny;
  m = m > g->nz ? m : g->nz;
  return blocks_for((int64_t)(m + 1) * (m + 1), 256) > 1024
             ? 1024
             : blocks_for((int64_t)(m + 1) * (m + 1), 256);
}

struct EnergyScratch {
  DeviceScratch part, out;
};
EnergyScratch &energy_scratch() {
  static thread_local EnergyScratch s[16];
  int dev = 0;
  (void)hipGetDevice(&dev);
  return s[dev & 15];
}

// ---------------------------------------------------------------------------
// Batched boxes: one persistent launch per phase over a work queue of
// (box, range) items; each box's voxels run the single-grid kernels' bodies
// (same arithmetic), at the box's offset in the packed arrays.
// ---------------------------------------------------------------------------
template <int PHASE>
__global__ __launch_bounds__(512) void k_ens_field(
    vpic_field_t *__restrict__ F, vpic_interpolator_t *__restrict__ FI,
    const vpic_accumulator_t *__restrict__ A,
    const EnsFieldBox *__restrict__ boxes, const EnsItem *__restrict__ items,
    int nitems, int *queue) {
  __shared__ int s_item;
  for (;;) {
    if (threadIdx.x == 0) s_item = atomicAdd(queue, 1);
    __syncthreads();
    const int it = s_item;
    __syncthreads();
    if (it >= nitems) break;
    const EnsItem I = items[it];
    const EnsFieldBox b = boxes[I.box];
    vpic_field_t *f = F + b.voff;
    const int nx = b.nx, ny = b.ny, nz = b.nz;
    for (int l = I.first + threadIdx.x; l < I.first + I.count;
         l += blockDim.x) {
      if (PHASE == ENS_GHOST || PHASE == ENS_FOLD) {
        FaceGeom G;
        G.n[0] = nx;
        G.n[1] = ny;
        G.n[2] = nz;
        G.s[0] = 1;
        G.s[1] = nx + 2;
        G.s[2] = (nx + 2) * (ny + 2);
        if (PHASE == ENS_GHOST)
          ghost_tang_b_at(f, G, I.axis, VPIC_BC_PERIODIC, VPIC_BC_PERIODIC, l);
        else
          fold_jf_at(f, G, I.axis, l);
        continue;
      }
      // voxel phases: l runs over (1..nx+1) x (1..ny+1) x (1..nz+1)
      const int wx = nx + 1, wy = ny + 1;
      const int x = 1 + l % wx, t = l / wx, y = 1 + t % wy, z = 1 + t / wy;
      if (PHASE == ENS_UNLOAD) {
        float4 *jf = reinterpret_cast<float4 *>(f + vox(x, y, z, nx, ny)) + 3;
        const float4 j0 = *jf;
        *jf = make_float4(0.0f, 0.0f, 0.0f, j0.w);  // clear_jf
        unload_accumulator_at(f, A + b.voff, x, y, z, nx, ny, b.uc[0], b.uc[1],
                              b.uc[2]);
      } else if (PHASE == ENS_ADVB) {
        advance_b_at(f, x, y, z, nx, ny, nz, b.bp[0], b.bp[1], b.bp[2]);
      } else if (PHASE == ENS_ADVE) {
        advance_e_at(f, x, y, z, nx, ny, nz, b.ec);
      } else if (PHASE == ENS_LOAD) {
        if (x <= nx && y <= ny && z <= nz)
          load_interpolator_at(FI + b.voff, f, x, y, z, nx, ny);
      }
    }
  }
}

__global__ void k_ens_reset(int *queues, int n) {
  for (int k = threadIdx.x; k < n; k += blockDim.x) queues[k] = 0;
}

}  // namespace

// The vacuum_advance_e couplings of grid g (vacuum_advance_e_pipeline.h:
// 27-33, in float, as launch_vacuum_advance_e forms them).
static ECoef make_ecoef(const vpic_grid_t *g, const vpic_field_params_t *m) {
  ECoef c;
  const float damp = m->damp;
  c.decayx = m->decayx;
  c.drivex = m->drivex;
  c.decayy = m->decayy;
  c.drivey = m->drivey;
  c.decayz = m->decayz;
  c.drivez = m->drivez;
  c.damp = damp;
  c.px_muz = ((g->nx > 1) ? (1 + damp) * g->cvac * g->dt * g->rdx : 0) * m->rmuz;
  c.px_muy = ((g->nx > 1) ? (1 + damp) * g->cvac * g->dt * g->rdx : 0) * m->rmuy;
  c.py_mux = ((g->ny > 1) ? (1 + damp) * g->cvac * g->dt * g->rdy : 0) * m->rmux;
  c.py_muz = ((g->ny > 1) ? (1 + damp) * g->cvac * g->dt * g->rdy : 0) * m->rmuz;
  c.pz_muy = ((g->nz > 1) ? (1 + damp) * g->cvac * g->dt * g->rdz : 0) * m->rmuy;
  c.pz_mux = ((g->nz > 1) ? (1 + damp) * g->cvac * g->dt * g->rdz : 0) * m->rmux;
  c.cj = g->dt / g->eps0;
  return c;
}

void ens_make_field_box(EnsFieldBox *b, const vpic_grid_t *g,
                        const vpic_field_params_t *m, int64_t voff) {
  b->nx = g->nx;
  b->ny = g->ny;
  b->nz = g->nz;
  b->voff = voff;
  const float frac = 0.5f;
  b->bp[0] = (g->nx > 1) ? frac * g->cvac * g->dt * g->rdx : 0;
  b->bp[1] = (g->ny > 1) ? frac * g->cvac * g->dt * g->rdy : 0;
  b->bp[2] = (g->nz > 1) ? frac * g->cvac * g->dt * g->rdz : 0;
  b->uc[0] = 0.25 * g->rdy * g->rdz / g->dt;
  b->uc[1] = 0.25 * g->rdz * g->rdx / g->dt;
  b->uc[2] = 0.25 * g->rdx * g->rdy / g->dt;
  b->ec = make_ecoef(g, m);
}

int launch_ens_field(int phase, vpic_field_t *F, vpic_interpolator_t *FI,
                     const vpic_accumulator_t *A, const EnsFieldBox *boxes,
                     const EnsItem *items, int nitems, int *queue,
                     int nblocks, int bs, hipStream_t s) {
  if (nitems <= 0) return VPIC_OK;
  const dim3 grid(nblocks < nitems ? nblocks : nitems), block(bs);
  switch (phase) {
    case ENS_UNLOAD:
      hipLaunchKernelGGL(k_ens_field<ENS_UNLOAD>, grid, block, 0, s, F, FI, A,
                         boxes, items, nitems, queue);
      break;
    case ENS_ADVB:
      hipLaunchKernelGGL(k_ens_field<ENS_ADVB>, grid, block, 0, s, F, FI, A,
                         boxes, items, nitems, queue);
      break;
    case ENS_ADVE:
      hipLaunchKernelGGL(k_ens_field<ENS_ADVE>, grid, block, 0, s, F, FI, A,
                         boxes, items, nitems, queue);
      break;
    case ENS_LOAD:
      hipLaunchKernelGGL(k_ens_field<ENS_LOAD>, grid, block, 0, s, F, FI, A,
                         boxes, items, nitems, queue);
      break;
    case ENS_GHOST:
      hipLaunchKernelGGL(k_ens_field<ENS_GHOST>, grid, block, 0, s, F, FI, A,
                         boxes, items, nitems, queue);
      break;
    default:
      hipLaunchKernelGGL(k_ens_field<ENS_FOLD>, grid, block, 0, s, F, FI, A,
                         boxes, items, nitems, queue);
      break;
  }
  VPIC_LAUNCHED("k_ens_field");
  return VPIC_OK;
}

int launch_ens_reset(int *queues, int n, hipStream_t s) {
  hipLaunchKernelGGL(k_ens_reset, dim3(1), dim3(64), 0, s, queues, n);
  VPIC_LAUNCHED("k_ens_reset");
  return VPIC_OK;
}

int launch_load_interpolator(vpic_interpolator_t *fi, const vpic_field_t *f,
                             const vpic_grid_t *g, hipStream_t s) {
  Box b = {1, 1, 1, g->nx, g->ny, g->nz};
  int64_t n = (int64_t)g->nx * g->ny * g->nz;
  hipLaunchKernelGGL(k_load_interpolator, dim3(xcd_grid(n, 256)), dim3(256),
                     0, s, fi, f, b, g->nx, g->ny);
  VPIC_LAUNCHED("k_load_interpolator");
  return VPIC_OK;
}

int launch_clear_accumulator(vpic_accumulator_t *a, const vpic_grid_t *g,
                             hipStream_t s) {
  int64_t n4 = (int64_t)g->nv * 3;
  int blk = blocks_for(n4, 256);
  if (blk > 65536) blk = 65536;
  hipLaunchKernelGGL(k_fill_zero4, dim3(blk), dim3(256), 0, s, (float4 *)a,
                     n4);
  VPIC_LAUNCHED("k_fill_zero4");
  return VPIC_OK;
}

int launch_clear_accumulator_n(vpic_accumulator_t *a, int64_t nv,
                               hipStream_t s) {
  const int64_t n4 = nv * 3;
  int blk = blocks_for(n4, 256);
  if (blk > 65536) blk = 65536;
  hipLaunchKernelGGL(k_fill_zero4, dim3(blk), dim3(256), 0, s, (float4 *)a,
                     n4);
  VPIC_LAUNCHED("k_fill_zero4");
  return VPIC_OK;
}

int launch_unload_accumulator(vpic_field_t *f, const vpic_accumulator_t *a,
                              const vpic_grid_t *g, hipStream_t s) {
  Box b = {1, 1, 1, g->nx + 1, g->ny + 1, g->nz + 1};
  int64_t n = (int64_t)(g->nx + 1) * (g->ny + 1) * (g->nz + 1);
  // unload_accumulator_pipeline.cc:137-139: double, stored as float.
  const float cx = 0.25 * g->rdy * g->rdz / g->dt;
  const float cy = 0.25 * g->rdz * g->rdx / g->dt;
  const float cz = 0.25 * g->rdx * g->rdy / g->dt;
  hipLaunchKernelGGL(k_unload_accumulator, dim3(xcd_grid(n, 256)),
                     dim3(256), 0, s, f, a, b, g->nx, g->ny, cx, cy, cz);
  VPIC_LAUNCHED("k_unload_accumulator");
  return VPIC_OK;
}

int launch_clear_unload_accumulator(vpic_field_t *f,
                                    const vpic_accumulator_t *a,
                                    const vpic_grid_t *g, hipStream_t s) {
  // the unload's coefficients, as launch_unload_accumulator computes them
  const float cx = 0.25 * g->rdy * g->rdz / g->dt;
  const float cy = 0.25 * g->rdz * g->rdx / g->dt;
  const float cz = 0.25 * g->rdx * g->rdy / g->dt;
  int blk = blocks_for(g->nv, 256);
  if (blk > 65536) blk = 65536;
  hipLaunchKernelGGL(k_clear_unload_accumulator, dim3(blk), dim3(256), 0, s,
                     f, a, g->nv, g->nx, g->ny, g->nz, cx, cy, cz);
  VPIC_LAUNCHED("k_clear_unload_accumulator");
  return VPIC_OK;
}

int launch_clear_jf(vpic_field_t *f, const vpic_grid_t *g, hipStream_t s) {
  int blk = blocks_for(g->nv, 256);
  if (blk > 65536) blk = 65536;
  hipLaunchKernelGGL(k_clear_jf, dim3(blk), dim3(256), 0, s, f, g->nv, 0);
  VPIC_LAUNCHED("k_clear_jf");
  return VPIC_OK;
}

int launch_clear_rhof(vpic_field_t *f, const vpic_grid_t *g, hipStream_t s) {
  int blk = blocks_for(g->nv, 256);
  if (blk > 65536) blk = 65536;
  hipLaunchKernelGGL(k_clear_jf, dim3(blk), dim3(256), 0, s, f, g->nv, 1);
  VPIC_LAUNCHED("k_clear_jf");
  return VPIC_OK;
}

static void face_modes(const vpic_grid_t *g, int4 &m0, int2 &m1) {
  m0 = make_int4(g->fbc[0], g->fbc[1], g->fbc[2], g->fbc[3]);
  m1 = make_int2(g->fbc[4], g->fbc[5]);
}

int launch_synchronize_jf(vpic_field_t *f, const vpic_grid_t *g,
                          hipStream_t s) {
  FaceGeom G = face_geom(g);
  const int fb = face_blocks(g);
  for (int k = 0; k < 6; k++) {
    if (g->fbc[k] == VPIC_BC_PERIODIC) continue;
    const int mode = g->fbc[k] == VPIC_BC_ANTI_SYMMETRIC_FIELDS ? 1 : 2;
    hipLaunchKernelGGL(k_adjust_jf, dim3(fb), dim3(256), 0, s, f, G, k, mode);
    VPIC_LAUNCHED("k_adjust_jf");
  }
  for (int X = 0; X < 3; X++) {
    if (g->fbc[X] != VPIC_BC_PERIODIC) continue;
    hipLaunchKernelGGL(k_fold_jf, dim3(fb), dim3(256), 0, s, f, G, X);
    VPIC_LAUNCHED("k_fold_jf");
  }
  return VPIC_OK;
}

int launch_synchronize_rho(vpic_field_t *f, const vpic_grid_t *g,
                           hipStream_t s) {
  FaceGeom G = face_geom(g);
  const int fb = face_blocks(g);
  for (int k = 0; k < 6; k++) {
    if (g->fbc[k] == VPIC_BC_PERIODIC) continue;
    const int mode = g->fbc[k] == VPIC_BC_ANTI_SYMMETRIC_FIELDS ? 1 : 2;
    hipLaunchKernelGGL(k_adjust_rho, dim3(fb), dim3(256), 0, s, f, G, k, mode);
    VPIC_LAUNCHED("k_adjust_rho");
  }
  for (int X = 0; X < 3; X++) {
    if (g->fbc[X] != VPIC_BC_PERIODIC) continue;
    hipLaunchKernelGGL(k_fold_rho, dim3(fb), dim3(256), 0, s, f, G, X);
    VPIC_LAUNCHED("k_fold_rho");
  }
  return VPIC_OK;
}

int launch_advance_b(vpic_field_t *f, const vpic_grid_t *g, float frac,
                     hipStream_t s) {
  // advance_b_pipeline.h:26-28, in float.
  const float px = (g->nx > 1) ? frac * g->cvac * g->dt * g->rdx : 0;
  const float py = (g->ny > 1) ? frac * g->cvac * g->dt * g->rdy : 0;
  const float pz = (g->nz > 1) ? frac * g->cvac * g->dt * g->rdz : 0;
  Box b = {1, 1, 1, g->nx + 1, g->ny + 1, g->nz + 1};
  int64_t n = (int64_t)(g->nx + 1) * (g->ny + 1) * (g->nz + 1);
  hipLaunchKernelGGL(k_advance_b, dim3(xcd_grid(n, 256)), dim3(256), 0, s, f,
                     b, g->nx, g->ny, g->nz, px, py, pz);
  VPIC_LAUNCHED("k_advance_b");
  bool sym = false;
  for (int k = 0; k < 6; k++) sym |= g->fbc[k] == VPIC_BC_SYMMETRIC_FIELDS;
  if (sym) {
    int4 m0;
    int2 m1;
    face_modes(g, m0, m1);
    hipLaunchKernelGGL(k_adjust_norm_b, dim3(face_blocks(g), 6), dim3(256), 0,
                       s, f, face_geom(g), m0, m1);
    VPIC_LAUNCHED("k_adjust_norm_b");
  }
  return VPIC_OK;
}

int launch_vacuum_advance_e(vpic_field_t *f, const vpic_grid_t *g,
                            const vpic_field_params_t *m, hipStream_t s) {
  FaceGeom G = face_geom(g);
  const int fb = face_blocks(g);
  for (int X = 0; X < 3; X++) {
    hipLaunchKernelGGL(k_ghost_tang_b, dim3(fb), dim3(256), 0, s, f, G, X,
                       g->fbc[X], g->fbc[X + 3]);
    VPIC_LAUNCHED("k_ghost_tang_b");
  }
  // vacuum_advance_e_pipeline.h:23-33, in float.
  const ECoef c = make_ecoef(g, m);
  Box b = {1, 1, 1, g->nx + 1, g->ny + 1, g->nz + 1};
  int64_t n = (int64_t)(g->nx + 1) * (g->ny + 1) * (g->nz + 1);
  hipLaunchKernelGGL(k_advance_e, dim3(xcd_grid(n, 256)), dim3(256), 0, s, f,
                     b, g->nx, g->ny, g->nz, c);
  VPIC_LAUNCHED("k_advance_e");
  bool pec = false;
  for (int k = 0; k < 6; k++) pec |= g->fbc[k] == VPIC_BC_ANTI_SYMMETRIC_FIELDS;
  if (pec) {
    int4 m0;
    int2 m1;
    face_modes(g, m0, m1);
    hipLaunchKernelGGL(k_adjust_tang_e, dim3(fb, 6), dim3(256), 0, s, f, G, m0,
                       m1);
    VPIC_LAUNCHED("k_adjust_tang_e");
  }
  return VPIC_OK;
}

int launch_energy_f(double en[6], const vpic_field_t *f, const vpic_grid_t *g,
                    const vpic_field_params_t *m, hipStream_t s) {
  EnergyScratch &S = energy_scratch();
  Box b = {1, 1, 1, g->nx, g->ny, g->nz};
  int64_t n = (int64_t)g->nx * g->ny * g->nz;
  int nblk = blocks_for(n, 256);
  int rc;
  if ((rc = scratch_reserve(S.part, sizeof(double) * 6 * (size_t)nblk)) ||
      (rc = scratch_reserve(S.out, sizeof(double) * 8)))
    return rc;
  // vacuum_energy_f_pipeline.h:24-29: float coefficients.
  float4 q = make_float4(0.25 * m->epsx, 0.25 * m->epsy, 0.25 * m->epsz, 0);
  float4 h = make_float4(0.50 * m->rmux, 0.50 * m->rmuy, 0.50 * m->rmuz, 0);
  hipLaunchKernelGGL(k_energy_f, dim3(nblk), dim3(256), 0, s, f, b, g->nx,
                     g->ny, q, h, (double *)S.part.ptr);
  VPIC_LAUNCHED("k_energy_f");
  hipLaunchKernelGGL(k_fold_partials, dim3(1), dim3(256), 0, s,
                     (const double *)S.part.ptr, nblk, 6, (double *)S.out.ptr);
  VPIC_LAUNCHED("k_fold_partials");
  double out[6];
  VPIC_HIP(hipMemcpyAsync(out, S.out.ptr, sizeof(out), hipMemcpyDeviceToHost,
                          s));
  VPIC_HIP(hipStreamSynchronize(s));
  const double v0 = 0.5 * g->eps0 * g->dV;
  for (int k = 0; k < 6; k++) en[k] = out[k] * v0;
  return VPIC_OK;
}

// ---------------------------------------------------------------------------
// Divergence cleaning and initialization (launchers)
// ---------------------------------------------------------------------------
namespace {

int launch_adjust_tang_e_if(vpic_field_t *f, const vpic_grid_t *g,
                            hipStream_t s) {
  bool pec = false;
  for (int k = 0; k < 6; k++) pec |= g->fbc[k] == VPIC_BC_ANTI_SYMMETRIC_FIELDS;
  if (!pec) return VPIC_OK;
  int4 m0;
  int2 m1;
  face_modes(g, m0, m1);
  hipLaunchKernelGGL(k_adjust_tang_e, dim3(face_blocks(g), 6), dim3(256), 0,
                     s, f, face_geom(g), m0, m1);
  VPIC_LAUNCHED("k_adjust_tang_e");
  return VPIC_OK;
}

int launch_adjust_norm_b_if(vpic_field_t *f, const vpic_grid_t *g,
                            hipStream_t s) {
  bool sym = false;
  for (int k = 0; k < 6; k++) sym |= g->fbc[k] == VPIC_BC_SYMMETRIC_FIELDS;
  if (!sym) return VPIC_OK;
  int4 m0;
  int2 m1;
  face_modes(g, m0, m1);
  hipLaunchKernelGGL(k_adjust_norm_b, dim3(face_blocks(g), 6), dim3(256), 0,
                     s, f, face_geom(g), m0, m1);
  VPIC_LAUNCHED("k_adjust_norm_b");
  return VPIC_OK;
}

Box node_box(const vpic_grid_t *g) {
  Box b = {1, 1, 1, g->nx + 1, g->ny + 1, g->nz + 1};
  return b;
}

int64_t node_count(const vpic_grid_t *g) {
  return (int64_t)(g->nx + 1) * (g->ny + 1) * (g->nz + 1);
}

// FP64 sum of per-block partials (one fold block, fixed order) to the host.
int fold_to_host(double *out, const double *part, int nblk, hipStream_t s) {
  EnergyScratch &S = energy_scratch();
  int rc;
  if ((rc = scratch_reserve(S.out, sizeof(double) * 8))) return rc;
  hipLaunchKernelGGL(k_fold_partials, dim3(1), dim3(256), 0, s, part, nblk, 1,
                     (double *)S.out.ptr);
  VPIC_LAUNCHED("k_fold_partials");
  VPIC_HIP(hipMemcpyAsync(out, S.out.ptr, sizeof(double),
                          hipMemcpyDeviceToHost, s));
  VPIC_HIP(hipStreamSynchronize(s));
  return VPIC_OK;
}

}  // namespace

int launch_compute_div_e_err(vpic_field_t *f, const vpic_grid_t *g,
                             const vpic_field_params_t *m, int rhob,
                             hipStream_t s) {
  const FaceGeom G = face_geom(g);
  const int fb = face_blocks(g);
  for (int X = 0; X < 3; X++) {
    hipLaunchKernelGGL(k_ghost_norm_e, dim3(fb), dim3(256), 0, s, f, G, X,
                       g->fbc[X], g->fbc[X + 3]);
    VPIC_LAUNCHED("k_ghost_norm_e");
  }
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const Box b = node_box(g);
  const int nblk = blocks_for(node_count(g), 256);
  if (rhob) {
    // vacuum_compute_rhob_pipeline.h:23-26
    const float px = (nx > 1) ? g->eps0 * m->epsx * g->rdx : 0;
    const float py = (ny > 1) ? g->eps0 * m->epsy * g->rdy : 0;
    const float pz = (nz > 1) ? g->eps0 * m->epsz * g->rdz : 0;
    hipLaunchKernelGGL(k_div_e_nodes<true>, dim3(nblk), dim3(256), 0, s, f, b,
                       nx, ny, m->nonconductive, px, py, pz, 0.0f);
  } else {
    // vacuum_compute_div_e_err_pipeline.h:23-27
    const float px = ((nx > 1) ? g->rdx : 0) * m->epsx;
    const float py = ((ny > 1) ? g->rdy : 0) * m->epsy;
    const float pz = ((nz > 1) ? g->rdz : 0) * m->epsz;
    const float cj = 1. / g->eps0;
    hipLaunchKernelGGL(k_div_e_nodes<false>, dim3(nblk), dim3(256), 0, s, f, b,
                       nx, ny, m->nonconductive, px, py, pz, cj);
  }
  VPIC_LAUNCHED("k_div_e_nodes");
  for (int k = 0; k < 6; k++) {
    if (g->fbc[k] != VPIC_BC_ANTI_SYMMETRIC_FIELDS) continue;
    hipLaunchKernelGGL(k_zero_node_face, dim3(fb), dim3(256), 0, s, f, G, k,
                       rhob ? 11 : 3);
    VPIC_LAUNCHED("k_zero_node_face");
  }
  return VPIC_OK;
}

int launch_compute_rms_div_err(double *out, const vpic_field_t *f,
                               const vpic_grid_t *g, int divb, hipStream_t s) {
  EnergyScratch &S = energy_scratch();
  const Box b = divb ? Box{1, 1, 1, g->nx, g->ny, g->nz} : node_box(g);
  const int64_t n = divb ? (int64_t)g->nx * g->ny * g->nz : node_count(g);
  const int nblk = blocks_for(n, 256);
  int rc;
  if ((rc = scratch_reserve(S.part, sizeof(double) * (size_t)nblk))) return rc;
  if (divb)
    hipLaunchKernelGGL(k_sq_partials<true>, dim3(nblk), dim3(256), 0, s, f, b,
                       g->nx, g->ny, g->nz, (double *)S.part.ptr);
  else
    hipLaunchKernelGGL(k_sq_partials<false>, dim3(nblk), dim3(256), 0, s, f, b,
                       g->nx, g->ny, g->nz, (double *)S.part.ptr);
  VPIC_LAUNCHED("k_sq_partials");
  double err = 0;
  if ((rc = fold_to_host(&err, (const double *)S.part.ptr, nblk, s))) return rc;
  // compute_rms_div_e_err_pipeline.cc:177-182 (the volume in FP32 as there)
  const double local0 = err * g->dV;
  const double local1 = (g->nx * g->ny * g->nz) * g->dV;
  *out = g->eps0 * sqrt(local0 / local1);
  return VPIC_OK;
}

int launch_clean_div_e(vpic_field_t *f, const vpic_grid_t *g,
                       const vpic_field_params_t *m, hipStream_t s) {
  // vacuum_clean_div_e_pipeline.h:20-27
  const int nx = g->nx, ny = g->ny, nz = g->nz;
  const float _rdx = (nx > 1) ? g->rdx : 0;
  const float _rdy = (ny > 1) ? g->rdy : 0;
  const float _rdz = (nz > 1) ? g->rdz : 0;
  const float alphadt = 0.3888889 / (_rdx * _rdx + _rdy * _rdy + _rdz * _rdz);
  const float px = (alphadt * _rdx) * m->drivex;
  const float py = (alphadt * _rdy) * m->drivey;
  const float pz = (alphadt * _rdz) * m->drivez;
  hipLaunchKernelGGL(k_clean_div_e, dim3(blocks_for(node_count(g), 256)),
                     dim3(256), 0, s, f, node_box(g), nx, ny, nz, px, py, pz);
  VPIC_LAUNCHED("k_clean_div_e");
  return launch_adjust_tang_e_if(f, g, s);
}

int launch_compute_div_b_err(vpic_field_t *f, const vpic_grid_t *g,
                             hipStream_t s) {
  // compute_div_b_err_pipeline.cc:27-29
  const float px = (g->nx > 1) ? g->rdx : 0;
  const float py = (g->ny > 1) ? g->rdy : 0;
  const float pz = (g->nz > 1) ? g->rdz : 0;
  const Box b = {1, 1, 1, g->nx, g->ny, g->nz};
  const int64_t n = (int64_t)g->nx * g->ny * g->nz;
  hipLaunchKernelGGL(k_div_b_cells, dim3(blocks_for(n, 256)), dim3(256), 0, s,
                     f, b, g->nx, g->ny, px, py, pz);
  VPIC_LAUNCHED("k_div_b_cells");
  return VPIC_OK;
}

int launch_clean_div_b(vpic_field_t *f, const vpic_grid_t *g, hipStream_t s) {
  // clean_div_b_pipeline.cc:245-251
  float px = (g->nx > 1) ? g->rdx : 0;
  float py = (g->ny > 1) ? g->rdy : 0;
  float pz = (g->nz > 1) ? g->rdz : 0;
  const float alphadt = 0.3888889 / (px * px + py * py + pz * pz);
  px *= alphadt;
  py *= alphadt;
  pz *= alphadt;
  const FaceGeom G = face_geom(g);
  const int fb = face_blocks(g);
  for (int X = 0; X < 3; X++) {
    hipLaunchKernelGGL(k_ghost_div_b, dim3(fb), dim3(256), 0, s, f, G, X,
                       g->fbc[X], g->fbc[X + 3]);
    VPIC_LAUNCHED("k_ghost_div_b");
  }
  hipLaunchKernelGGL(k_clean_div_b, dim3(blocks_for(node_count(g), 256)),
                     dim3(256), 0, s, f, node_box(g), g->nx, g->ny, g->nz, px,
                     py, pz);
  VPIC_LAUNCHED("k_clean_div_b");
  return launch_adjust_norm_b_if(f, g, s);
}

int launch_compute_curl_b(vpic_field_t *f, const vpic_grid_t *g,
                          const vpic_field_params_t *m, hipStream_t s) {
  const FaceGeom G = face_geom(g);
  const int fb = face_blocks(g);
  for (int X = 0; X < 3; X++) {
    hipLaunchKernelGGL(k_ghost_tang_b, dim3(fb), dim3(256), 0, s, f, G, X,
                       g->fbc[X], g->fbc[X + 3]);
    VPIC_LAUNCHED("k_ghost_tang_b");
  }
  // vacuum_compute_curl_b_pipeline.h:22-27
  ECoef c = {};
  c.px_muz = ((g->nx > 1) ? g->cvac * g->dt * g->rdx : 0) * m->rmuz;
  c.px_muy = ((g->nx > 1) ? g->cvac * g->dt * g->rdx : 0) * m->rmuy;
  c.py_mux = ((g->ny > 1) ? g->cvac * g->dt * g->rdy : 0) * m->rmux;
  c.py_muz = ((g->ny > 1) ? g->cvac * g->dt * g->rdy : 0) * m->rmuz;
  c.pz_muy = ((g->nz > 1) ? g->cvac * g->dt * g->rdz : 0) * m->rmuy;
  c.pz_mux = ((g->nz > 1) ? g->cvac * g->dt * g->rdz : 0) * m->rmux;
  hipLaunchKernelGGL(k_curl_b, dim3(blocks_for(node_count(g), 256)), dim3(256),
                     0, s, f, node_box(g), g->nx, g->ny, g->nz, c);
  VPIC_LAUNCHED("k_curl_b");
  return launch_adjust_tang_e_if(f, g, s);
}

int launch_synchronize_tang_e_norm_b(double *out, vpic_field_t *f,
                                     const vpic_grid_t *g, hipStream_t s) {
  EnergyScratch &S = energy_scratch();
  int rc;
  if ((rc = launch_adjust_tang_e_if(f, g, s)) ||
      (rc = launch_adjust_norm_b_if(f, g, s)))
    return rc;
  const FaceGeom G = face_geom(g);
  int nb[3], total = 0;
  for (int X = 0; X < 3; X++) {
    const int Y = (X + 1) % 3, Z = (X + 2) % 3;
    const bool per = g->fbc[X] == VPIC_BC_PERIODIC &&
                     g->fbc[X + 3] == VPIC_BC_PERIODIC;
    nb[X] = per ? blocks_for((int64_t)(G.n[Y] + 1) * (G.n[Z] + 1), 256) : 0;
    total += nb[X];
  }
  if ((rc = scratch_reserve(S.part, sizeof(double) * (size_t)(total + 1))))
    return rc;
  double *part = (double *)S.part.ptr;
  int off = 0;
  for (int X = 0; X < 3; X++) {
    if (!nb[X]) continue;
    hipLaunchKernelGGL(k_sync_tang_e_norm_b, dim3(nb[X]), dim3(256), 0, s, f,
                       G, X, part + off);
    VPIC_LAUNCHED("k_sync_tang_e_norm_b");
    off += nb[X];
  }
  if (total == 0) {
    *out = 0;
    return VPIC_OK;
  }
  return fold_to_host(out, part, total, s);
}
