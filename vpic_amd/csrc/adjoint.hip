// adjoint.hip -- the differentiable run: reverse-mode gradients of the
// magnetic field energy after N steps of the hot path, with respect to the
// initial particle momenta and positions and the initial fields.
//
// The step is vpic_simulation::advance (src/vpic/advance.cc:25-185) on a
// periodic box with one pushed species: clear_accumulator_array, advance_p
// (advance_p_pipeline.cc:96-238 with move_p.cc:233-375), unload_accumulator
// (unload_accumulator_pipeline.cc:65-67), synchronize_jf (remote.cc:418-507),
// advance_b(0.5) (advance_b_pipeline.h:57-59), the tangential-B ghost copy
// and vacuum_advance_e (vacuum_advance_e_pipeline.h:56-69, damp 0),
// advance_b(0.5), load_interpolator_array (interpolator_array_pipeline.cc:
// 67-121).  The observable is the magnetic part of energy_f
// (vacuum_energy_f_pipeline.h:44-75).
//
// Everything is written once over the scalar type R: R = float is the
// product run (its forward is the reference's arithmetic operation for
// operation, so positions, momenta and cell indices equal the main push's);
// R = double is the same run in FP64, the reference for finite-difference
// checks.
//
// The field operations are linear, so their adjoints are their exact
// transposes over the array elements (ghost planes and the duplicated
// periodic face planes included), each written as a gather (no atomics).
// The particle push is differentiated with its branch held fixed: the cell
// sequence of a crossing particle, the face of each segment and the in-cell
// test are taken from the recomputed forward.  A crossing sets the crossed
// coordinate to exactly +-1, so that coordinate carries no derivative; the
// other quantities of a segment are smooth.  The reverse kernel recomputes
// each particle's forward from its state at the start of the step and the
// step's interpolators, then runs the derivatives backwards; a crossing
// particle's segments are replayed from the start for each segment it
// reverses (the segment count is small), so no per-segment storage is
// needed.  Interpolator adjoints are summed over runs of lanes in the same
// cell (a segmented wavefront reduction) before the global atomics.
//
// Checkpointing: the forward stores the state (particles and fields) every
// `ckpt` steps; the reverse takes the segments last first, recomputes the
// segment's steps from its checkpoint storing each step's starting state,
// and reverses them.
#include "common.h"

#include <vector>

namespace {

template <class R>
struct Part {
  R dx, dy, dz;
  int32_t i;
  R ux, uy, uz, w;
};
static_assert(sizeof(Part<float>) == 32, "particle_t layout");
static_assert(sizeof(Part<double>) == 64, "FP64 particle layout");

// The six field components the step carries between steps plus the current
// the step forms (tca is not carried: with no damping advance_e overwrites
// it before it is read).
template <class R>
struct Fld {
  R ex, ey, ez, cbx, cby, cbz, jfx, jfy, jfz;
};

constexpr int NI = 18;  // interpolator coefficients ex .. dcbzdz

struct Box {
  int nx, ny, nz, sy, sz, nv;
};

template <class R>
struct Coef {
  R qdt_2mc, cdt[3], qsp;
  R bp[3];   // advance_b(0.5): frac c dt / d per axis (0 for one cell)
  R ep[3];   // vacuum_advance_e: c dt / d per axis (0 for one cell)
  R cj;      // dt / eps0
  R uc[3];   // unload: 0.25 rd_y rd_z / dt, cyclic
};

__device__ inline float rsqrt_div(float a, float b) { return a / sqrtf(b); }
__device__ inline double rsqrt_div(double a, double b) { return a / sqrt(b); }
// move_p.cc:277: q*sx*sy*sz promoted to double by (1.0/3.0)
__device__ inline float streak_third(float x) {
  return (float)((double)x * (1.0 / 3.0));
}
__device__ inline double streak_third(double x) { return x * (1.0 / 3.0); }
__device__ inline void atomic_add_r(float *a, float v) { atomic_add_f32(a, v); }
__device__ inline void atomic_add_r(double *a, double v) {
  atomic_add_f64(a, v);
}

__device__ inline void coords(const Box &B, int v, int &x, int &y, int &z) {
  x = v % (B.nx + 2);
  const int t = v / (B.nx + 2);
  y = t % (B.ny + 2);
  z = t / (B.ny + 2);
}

// Sum `vals` over the contiguous run of lanes holding the same key; the
// first lane of each run adds the run's sums to base + key*N + k.  Lanes with
// key < 0 take no part.
template <class R, int N>
__device__ inline void wave_run_add(R *base, int key, R (&vals)[N]) {
  const int lane = lane_id();
  const int prev = __shfl_up(key, 1);
  const bool head = key >= 0 && (lane == 0 || prev != key);
  const unsigned long long heads = __ballot(head || key < 0);
  // the last lane of this lane's run
  const unsigned long long after =
      lane == 63 ? 0ull : heads & (~0ull << (lane + 1));
  const int tail = after ? __ffsll((long long)after) - 2 : 63;
#pragma unroll
  for (int k = 0; k < N; k++) {
    R v = vals[k];
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const R o = __shfl_down(v, off);
      if (lane + off <= tail) v += o;
    }
    vals[k] = v;
  }
  if (head) {
#pragma unroll
    for (int k = 0; k < N; k++)
      if (vals[k] != R(0)) atomic_add_r(base + (size_t)key * N + k, vals[k]);
  }
}

// The twelve updates of one streak in ACCUMULATE_J's order
// (advance_p_pipeline.cc:186-208, move_p.cc:281-303) into out[12].
template <class R>
__device__ inline void streak(R (&out)[12], R qx, R qy, R qz, R mx, R my,
                              R mz, R v5) {
  const R qd[3] = {qx, qy, qz}, m[3] = {mx, my, mz};
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const R mY = m[(c + 1) % 3], mZ = m[(c + 2) % 3];
    R v4 = qd[c], v1 = v4 * mY, v0 = v4 - v1;
    v1 += v4;
    v4 = 1 + mZ;
    R v2 = v0 * v4, v3 = v1 * v4;
    v4 = 1 - mZ;
    v0 *= v4;
    v1 *= v4;
    v0 += v5;
    v1 -= v5;
    v2 -= v5;
    v3 += v5;
    out[4 * c + 0] = v0;
    out[4 * c + 1] = v1;
    out[4 * c + 2] = v2;
    out[4 * c + 3] = v3;
  }
}

// One move_p segment (move_p.cc:240-275): the fraction v3 of the remaining
// displacement that reaches the first face (or all of it), that face's axis
// (3 for none) and direction.
template <class R>
struct Seg {
  R f, dir;
  int axis;
};

template <class R>
__device__ inline Seg<R> seg_geom(const R (&d)[3], const R (&disp)[3]) {
  R sdir[3], t[3];
#pragma unroll
  for (int k = 0; k < 3; k++) {
    sdir[k] = (disp[k] > R(0)) ? R(1) : R(-1);
    t[k] = (disp[k] == R(0)) ? R(3.4e38f) : (sdir[k] - d[k]) / disp[k];
  }
  Seg<R> s;
  R v3 = 2;
  s.axis = 3;
  if (t[0] < v3) v3 = t[0], s.axis = 0;
  if (t[1] < v3) v3 = t[1], s.axis = 1;
  if (t[2] < v3) v3 = t[2], s.axis = 2;
  s.f = v3 * R(0.5);
  s.dir = s.axis < 3 ? sdir[s.axis] : R(0);
  return s;
}

// ---------------------------------------------------------------------------
// forward push: advance_p_pipeline.cc:96-238 with the scalar move_p for a
// periodic box; currents by run-summed atomics
// ---------------------------------------------------------------------------
template <class R>
__global__ void __launch_bounds__(256)
    k_push_fwd(Part<R> *__restrict__ P, int np, const R *__restrict__ fi,
               R *__restrict__ acc, Coef<R> C, GridDev G, int *err) {
  const R one = 1, one_third = R(1.0 / 3.0), two_fifteenths = R(2.0 / 15.0);
  const int stride = gridDim.x * blockDim.x;
  const int n_end = (np + 63) & ~63;  // whole wavefronts take the run sums
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < n_end; n += stride) {
    const bool live = n < np;
    R dep[12];
    int key = -1;
#pragma unroll
    for (int k = 0; k < 12; k++) dep[k] = 0;
    if (live) {
      Part<R> p = P[n];
      const int ii = p.i;
      const R *f = fi + (size_t)NI * ii;
      R dx = p.dx, dy = p.dy, dz = p.dz;
      const R qc = C.qdt_2mc;
      const R hax = qc * ((f[0] + dy * f[1]) + dz * (f[2] + dy * f[3]));
      const R hay = qc * ((f[4] + dz * f[5]) + dx * (f[6] + dz * f[7]));
      const R haz = qc * ((f[8] + dx * f[9]) + dy * (f[10] + dx * f[11]));
      const R cbx = f[12] + dx * f[13];
      const R cby = f[14] + dy * f[15];
      const R cbz = f[16] + dz * f[17];
      R ux = p.ux + hax, uy = p.uy + hay, uz = p.uz + haz;
      R v0 = rsqrt_div(qc, one + (ux * ux + (uy * uy + uz * uz)));
      R v1 = cbx * cbx + (cby * cby + cbz * cbz);
      R v2 = (v0 * v0) * v1;
      R v3 = v0 * (one + v2 * (one_third + v2 * two_fifteenths));
      R v4 = v3 / (one + v1 * (v3 * v3));
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
      p.ux = ux;
      p.uy = uy;
      p.uz = uz;
      v0 = rsqrt_div(one, one + (ux * ux + (uy * uy + uz * uz)));
      ux *= C.cdt[0];
      uy *= C.cdt[1];
      uz *= C.cdt[2];
      ux *= v0;
      uy *= v0;
      uz *= v0;
      v0 = dx + ux;
      v1 = dy + uy;
      v2 = dz + uz;
      v3 = v0 + ux;
      v4 = v1 + uy;
      R v5 = v2 + uz;
      const R q = p.w * C.qsp;
      if (v3 <= one && v4 <= one && v5 <= one && -v3 <= one && -v4 <= one &&
          -v5 <= one) {
        p.dx = v3;
        p.dy = v4;
        p.dz = v5;
        v5 = q * ux * uy * uz * one_third;
        streak(dep, q * ux, q * uy, q * uz, v0, v1, v2, v5);
        key = ii;
      } else {
        // move_p (move_p.cc:233-375) on a periodic box
        R d[3] = {dx, dy, dz}, disp[3] = {ux, uy, uz};
        int cell = ii;
        for (int guard = 0; guard < 64; guard++) {
          const Seg<R> s = seg_geom(d, disp);
          R sd[3], mid[3];
#pragma unroll
          for (int k = 0; k < 3; k++) {
            sd[k] = disp[k] * s.f;
            mid[k] = d[k] + sd[k];
          }
          const R w5 = streak_third(q * sd[0] * sd[1] * sd[2]);
          R st[12];
          streak(st, q * sd[0], q * sd[1], q * sd[2], mid[0], mid[1], mid[2],
                 w5);
#pragma unroll
          for (int k = 0; k < 12; k++)
            atomic_add_r(acc + (size_t)12 * cell + k, st[k]);
#pragma unroll
          for (int k = 0; k < 3; k++) {
            disp[k] -= sd[k];
            d[k] += sd[k] + sd[k];
          }
          if (s.axis == 3) break;
          d[s.axis] = s.dir;
          const int face = s.axis + (s.dir > R(0) ? 3 : 0);
          const int64_t nb = grid_neighbor(G, cell, face);
          if (nb < G.rangel || nb > G.rangeh) {
            atomicAdd(err, 1);  // only periodic boxes are differentiated
            break;
          }
          cell = (int)(nb - G.rangel);
          d[s.axis] = -s.dir;
        }
        p.dx = d[0];
        p.dy = d[1];
        p.dz = d[2];
        p.i = cell;
      }
      P[n] = p;
    }
    wave_run_add<R, 12>(acc, key, dep);
  }
}

// ---------------------------------------------------------------------------
// reverse push
// ---------------------------------------------------------------------------

// Adjoint of one streak's twelve updates: adds d/d(sd) and d/d(mid) of
// sum_k ab[k] * update_k to sdb and mb.
template <class R>
__device__ inline void streak_rev(const R *ab, R qw, const R (&sd)[3],
                                  const R (&m)[3], R (&sdb)[3], R (&mb)[3]) {
  R v5b = 0;
#pragma unroll
  for (int c = 0; c < 3; c++) {
    const int Y = (c + 1) % 3, Z = (c + 2) % 3;
    const R a0 = ab[4 * c], a1 = ab[4 * c + 1], a2 = ab[4 * c + 2],
            a3 = ab[4 * c + 3];
    const R mY = m[Y], mZ = m[Z];
    const R qd = qw * sd[c];
    const R qdb = a0 * (1 - mY) * (1 - mZ) + a1 * (1 + mY) * (1 - mZ) +
                  a2 * (1 - mY) * (1 + mZ) + a3 * (1 + mY) * (1 + mZ);
    mb[Y] += qd * ((1 - mZ) * (a1 - a0) + (1 + mZ) * (a3 - a2));
    mb[Z] += qd * ((1 - mY) * (a2 - a0) + (1 + mY) * (a3 - a1));
    v5b += (a0 - a1) - (a2 - a3);
    sdb[c] += qw * qdb;
  }
  const R t = v5b * qw * R(1.0 / 3.0);
  sdb[0] += t * sd[1] * sd[2];
  sdb[1] += t * sd[0] * sd[2];
  sdb[2] += t * sd[0] * sd[1];
}

// pb: per particle (d-bar[3], u-bar[3]) of the state after the step on
// entry, of the state before the step on return.  accb: the accumulator
// adjoint; fib: the interpolator adjoint (atomics).
template <class R>
__global__ void __launch_bounds__(256)
    k_push_rev(const Part<R> *__restrict__ P, int np,
               const R *__restrict__ fi, const R *__restrict__ accb,
               R *__restrict__ pb, R *__restrict__ fib, Coef<R> C, GridDev G) {
  const R one = 1, c3 = R(1.0 / 3.0), c15 = R(2.0 / 15.0);
  const int stride = gridDim.x * blockDim.x;
  const int n_end = (np + 63) & ~63;
  for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < n_end; n += stride) {
    R gI[NI];
#pragma unroll
    for (int k = 0; k < NI; k++) gI[k] = 0;
    int key = -1;
    if (n < np) {
      const Part<R> p = P[n];
      const int ii = p.i;
      key = ii;
      const R *f = fi + (size_t)NI * ii;
      const R d[3] = {p.dx, p.dy, p.dz};
      const R qc = C.qdt_2mc, qw = p.w * C.qsp;
      // forward, in exact-arithmetic form (the branch from the forward's
      // own rounding below)
      const R ha[3] = {qc * ((f[0] + d[1] * f[1]) + d[2] * (f[2] + d[1] * f[3])),
                       qc * ((f[4] + d[2] * f[5]) + d[0] * (f[6] + d[2] * f[7])),
                       qc * ((f[8] + d[0] * f[9]) + d[1] * (f[10] + d[0] * f[11]))};
      const R b[3] = {f[12] + d[0] * f[13], f[14] + d[1] * f[15],
                      f[16] + d[2] * f[17]};
      const R a[3] = {p.ux + ha[0], p.uy + ha[1], p.uz + ha[2]};
      const R g2 = one + (a[0] * a[0] + (a[1] * a[1] + a[2] * a[2]));
      const R v0 = rsqrt_div(qc, g2);
      const R v1 = b[0] * b[0] + (b[1] * b[1] + b[2] * b[2]);
      const R v2 = (v0 * v0) * v1;
      const R v3 = v0 * (one + v2 * (c3 + v2 * c15));
      const R den = one + v1 * (v3 * v3);
      R v4 = v3 / den;
      v4 += v4;
      const R w[3] = {a[0] + v3 * (a[1] * b[2] - a[2] * b[1]),
                      a[1] + v3 * (a[2] * b[0] - a[0] * b[2]),
                      a[2] + v3 * (a[0] * b[1] - a[1] * b[0])};
      const R wxb[3] = {w[1] * b[2] - w[2] * b[1], w[2] * b[0] - w[0] * b[2],
                        w[0] * b[1] - w[1] * b[0]};
      const R un[3] = {a[0] + v4 * wxb[0] + ha[0], a[1] + v4 * wxb[1] + ha[1],
                       a[2] + v4 * wxb[2] + ha[2]};
      const R r2 = rsqrt_div(one, one + (un[0] * un[0] +
                                         (un[1] * un[1] + un[2] * un[2])));
      const R s[3] = {(un[0] * C.cdt[0]) * r2, (un[1] * C.cdt[1]) * r2,
                      (un[2] * C.cdt[2]) * r2};
      R db[3] = {pb[6 * (size_t)n], pb[6 * (size_t)n + 1],
                 pb[6 * (size_t)n + 2]};
      R ub[3] = {pb[6 * (size_t)n + 3], pb[6 * (size_t)n + 4],
                 pb[6 * (size_t)n + 5]};
      R sb[3], d0b[3];
      const R e3[3] = {(d[0] + s[0]) + s[0], (d[1] + s[1]) + s[1],
                       (d[2] + s[2]) + s[2]};
      if (e3[0] <= one && e3[1] <= one && e3[2] <= one && -e3[0] <= one &&
          -e3[1] <= one && -e3[2] <= one) {
        // in-cell: new d = d + 2s, streak (s, d + s) in cell ii
        const R mid[3] = {d[0] + s[0], d[1] + s[1], d[2] + s[2]};
        R sdb[3] = {0, 0, 0}, mb[3] = {0, 0, 0};
        streak_rev(accb + (size_t)12 * ii, qw, s, mid, sdb, mb);
#pragma unroll
        for (int k = 0; k < 3; k++) {
          d0b[k] = db[k] + mb[k];
          sb[k] = 2 * db[k] + mb[k] + sdb[k];
        }
      } else {
        // count the segments (the forward's move_p walk)
        int nseg = 0;
        {
          R dd[3] = {d[0], d[1], d[2]}, ds[3] = {s[0], s[1], s[2]};
          int cell = ii;
          for (; nseg < 64;) {
            const Seg<R> g = seg_geom(dd, ds);
            nseg++;
#pragma unroll
            for (int k = 0; k < 3; k++) {
              const R sd = ds[k] * g.f;
              ds[k] -= sd;
              dd[k] += sd + sd;
            }
            if (g.axis == 3) break;
            const int face = g.axis + (g.dir > R(0) ? 3 : 0);
            cell = (int)(grid_neighbor(G, cell, face) - G.rangel);
            dd[g.axis] = -g.dir;
          }
        }
        // reverse the segments, each replayed from the start
        R Db[3] = {db[0], db[1], db[2]}, Pb[3] = {0, 0, 0};
        for (int j = nseg - 1; j >= 0; j--) {
          R dd[3] = {d[0], d[1], d[2]}, ds[3] = {s[0], s[1], s[2]};
          int cell = ii;
          Seg<R> g = seg_geom(dd, ds);
          for (int t = 0; t < j; t++) {
#pragma unroll
            for (int k = 0; k < 3; k++) {
              const R sd = ds[k] * g.f;
              ds[k] -= sd;
              dd[k] += sd + sd;
            }
            const int face = g.axis + (g.dir > R(0) ? 3 : 0);
            cell = (int)(grid_neighbor(G, cell, face) - G.rangel);
            dd[g.axis] = -g.dir;
            g = seg_geom(dd, ds);
          }
          // segment j: sd = ds f, mid = dd + sd, d' = dd + 2 sd (the crossed
          // coordinate then set to a constant), disp' = ds - sd
          R sd[3], mid[3];
#pragma unroll
          for (int k = 0; k < 3; k++) {
            sd[k] = ds[k] * g.f;
            mid[k] = dd[k] + sd[k];
          }
          if (g.axis < 3) Db[g.axis] = 0;
          R sdb[3], mb[3] = {0, 0, 0};
#pragma unroll
          for (int k = 0; k < 3; k++) sdb[k] = 2 * Db[k] - Pb[k];
          streak_rev(accb + (size_t)12 * cell, qw, sd, mid, sdb, mb);
          R fb = 0;
#pragma unroll
          for (int k = 0; k < 3; k++) {
            sdb[k] += mb[k];
            Db[k] += mb[k];
            Pb[k] += sdb[k] * g.f;
            fb += sdb[k] * ds[k];
          }
          if (g.axis < 3) {
            Db[g.axis] += -R(0.5) * fb / ds[g.axis];
            Pb[g.axis] += -fb * g.f / ds[g.axis];
          }
        }
#pragma unroll
        for (int k = 0; k < 3; k++) {
          d0b[k] = Db[k];
          sb[k] = Pb[k];
        }
      }
      // s = un cdt r2, r2 = (1 + |un|^2)^-1/2
      R rb = 0;
#pragma unroll
      for (int k = 0; k < 3; k++) rb += sb[k] * un[k] * C.cdt[k];
      R unb[3];
#pragma unroll
      for (int k = 0; k < 3; k++)
        unb[k] = ub[k] + sb[k] * C.cdt[k] * r2 - rb * r2 * r2 * r2 * un[k];
      // un = a + v4 (w x b) + ha
      R hab[3] = {unb[0], unb[1], unb[2]};
      R ab[3] = {unb[0], unb[1], unb[2]};
      const R v4b = unb[0] * wxb[0] + unb[1] * wxb[1] + unb[2] * wxb[2];
      R wb[3] = {v4 * (b[1] * unb[2] - b[2] * unb[1]),
                 v4 * (b[2] * unb[0] - b[0] * unb[2]),
                 v4 * (b[0] * unb[1] - b[1] * unb[0])};
      R bb[3] = {v4 * (unb[1] * w[2] - unb[2] * w[1]),
                 v4 * (unb[2] * w[0] - unb[0] * w[2]),
                 v4 * (unb[0] * w[1] - unb[1] * w[0])};
      // w = a + v3 (a x b)
      const R axb[3] = {a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2],
                        a[0] * b[1] - a[1] * b[0]};
      R v3b = wb[0] * axb[0] + wb[1] * axb[1] + wb[2] * axb[2];
      ab[0] += wb[0] + v3 * (b[1] * wb[2] - b[2] * wb[1]);
      ab[1] += wb[1] + v3 * (b[2] * wb[0] - b[0] * wb[2]);
      ab[2] += wb[2] + v3 * (b[0] * wb[1] - b[1] * wb[0]);
      bb[0] += v3 * (wb[1] * a[2] - wb[2] * a[1]);
      bb[1] += v3 * (wb[2] * a[0] - wb[0] * a[2]);
      bb[2] += v3 * (wb[0] * a[1] - wb[1] * a[0]);
      // v4 = 2 v3 / (1 + v1 v3^2)
      v3b += v4b * (2 / den) * (one - 2 * v1 * v3 * v3 / den);
      R v1b = -v4b * 2 * v3 * v3 * v3 / (den * den);
      // v3 = v0 (1 + v2 (1/3 + v2 2/15))
      R v0b = v3b * (one + v2 * (c3 + v2 * c15));
      const R v2b = v3b * v0 * (c3 + 2 * c15 * v2);
      // v2 = v0^2 v1
      v0b += v2b * 2 * v0 * v1;
      v1b += v2b * v0 * v0;
      // v1 = |b|^2
#pragma unroll
      for (int k = 0; k < 3; k++) bb[k] += 2 * v1b * b[k];
      // v0 = qc / sqrt(1 + |a|^2)
      const R t0 = -v0b * v0 / g2;
#pragma unroll
      for (int k = 0; k < 3; k++) {
        ab[k] += t0 * a[k];
        hab[k] += ab[k];  // a = u + ha
      }
      // interpolation
      R g = qc * hab[0];
      gI[0] = g;
      gI[1] = g * d[1];
      gI[2] = g * d[2];
      gI[3] = g * d[1] * d[2];
      d0b[1] += g * (f[1] + d[2] * f[3]);
      d0b[2] += g * (f[2] + d[1] * f[3]);
      g = qc * hab[1];
      gI[4] = g;
      gI[5] = g * d[2];
      gI[6] = g * d[0];
      gI[7] = g * d[2] * d[0];
      d0b[2] += g * (f[5] + d[0] * f[7]);
      d0b[0] += g * (f[6] + d[2] * f[7]);
      g = qc * hab[2];
      gI[8] = g;
      gI[9] = g * d[0];
      gI[10] = g * d[1];
      gI[11] = g * d[0] * d[1];
      d0b[0] += g * (f[9] + d[1] * f[11]);
      d0b[1] += g * (f[10] + d[0] * f[11]);
      gI[12] = bb[0];
      gI[13] = bb[0] * d[0];
      d0b[0] += bb[0] * f[13];
      gI[14] = bb[1];
      gI[15] = bb[1] * d[1];
      d0b[1] += bb[1] * f[15];
      gI[16] = bb[2];
      gI[17] = bb[2] * d[2];
      d0b[2] += bb[2] * f[17];
#pragma unroll
      for (int k = 0; k < 3; k++) {
        pb[6 * (size_t)n + k] = d0b[k];
        pb[6 * (size_t)n + 3 + k] = ab[k];
      }
    }
    wave_run_add<R, NI>(fib, key, gI);
  }
}

// ---------------------------------------------------------------------------
// field operations and their transposes (periodic box)
// ---------------------------------------------------------------------------

__device__ inline bool in_box(const Box &B, int x, int y, int z, int x0,
                              int x1, int y0, int y1, int z0, int z1) {
  return x >= x0 && x <= x1 && y >= y0 && y <= y1 && z >= z0 && z <= z1;
}

template <class R>
__global__ void k_fill(R *a, size_t n, R v) {
  for (size_t k = blockIdx.x * (size_t)blockDim.x + threadIdx.x; k < n;
       k += (size_t)gridDim.x * blockDim.x)
    a[k] = v;
}

// load_interpolator_pipeline_scalar (interpolator_array_pipeline.cc:67-121)
template <class R>
__global__ void k_load(R *fi, const Fld<R> *F, Box B) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= B.nv) return;
  int x, y, z;
  coords(B, v, x, y, z);
  if (!in_box(B, x, y, z, 1, B.nx, 1, B.ny, 1, B.nz)) return;
  const R fourth = R(0.25), half = R(0.5);
  const int sx = 1, sy = B.sy, sz = B.sz;
  R *o = fi + (size_t)NI * v;
  R w0, w1, w2, w3;
  w0 = F[v].ex; w1 = F[v + sy].ex; w2 = F[v + sz].ex; w3 = F[v + sy + sz].ex;
  o[0] = fourth * ((w3 + w0) + (w1 + w2));
  o[1] = fourth * ((w3 - w0) + (w1 - w2));
  o[2] = fourth * ((w3 - w0) - (w1 - w2));
  o[3] = fourth * ((w3 + w0) - (w1 + w2));
  w0 = F[v].ey; w1 = F[v + sz].ey; w2 = F[v + sx].ey; w3 = F[v + sz + sx].ey;
  o[4] = fourth * ((w3 + w0) + (w1 + w2));
  o[5] = fourth * ((w3 - w0) + (w1 - w2));
  o[6] = fourth * ((w3 - w0) - (w1 - w2));
  o[7] = fourth * ((w3 + w0) - (w1 + w2));
  w0 = F[v].ez; w1 = F[v + sx].ez; w2 = F[v + sy].ez; w3 = F[v + sx + sy].ez;
  o[8] = fourth * ((w3 + w0) + (w1 + w2));
  o[9] = fourth * ((w3 - w0) + (w1 - w2));
  o[10] = fourth * ((w3 - w0) - (w1 - w2));
  o[11] = fourth * ((w3 + w0) - (w1 + w2));
  w0 = F[v].cbx; w1 = F[v + sx].cbx;
  o[12] = half * (w1 + w0);
  o[13] = half * (w1 - w0);
  w0 = F[v].cby; w1 = F[v + sy].cby;
  o[14] = half * (w1 + w0);
  o[15] = half * (w1 - w0);
  w0 = F[v].cbz; w1 = F[v + sz].cbz;
  o[16] = half * (w1 + w0);
  o[17] = half * (w1 - w0);
}

// Transpose of k_load: Fb += load^T(fib).  Voxel u receives, for each of
// ex/ey/ez, the four corner weights of the interpolators at u, u - Y, u - Z,
// u - Y - Z ((Y, Z) the component's cyclic pair), and for cb the two face
// weights of u and u - X.
template <class R>
__device__ inline R corner_T(const R *fib, const Box &B, int u, int c0,
                             int oy, int oz) {
  // weights of (w0, w1, w2, w3) on (value, d/dY, d/dZ, d2/dYdZ)
  const R q = R(0.25);
  R s = 0;
  const int cand[4] = {u, u - oy, u - oz, u - oy - oz};
  const R sgn[4][4] = {{1, -1, -1, 1}, {1, 1, -1, -1}, {1, -1, 1, -1},
                       {1, 1, 1, 1}};
#pragma unroll
  for (int k = 0; k < 4; k++) {
    const int v = cand[k];
    if (v < 0) continue;
    int x, y, z;
    coords(B, v, x, y, z);
    if (!in_box(B, x, y, z, 1, B.nx, 1, B.ny, 1, B.nz)) continue;
    const R *g = fib + (size_t)NI * v + c0;
    s += q * (sgn[k][0] * g[0] + sgn[k][1] * g[1] + sgn[k][2] * g[2] +
              sgn[k][3] * g[3]);
  }
  return s;
}

template <class R>
__device__ inline R face_T(const R *fib, const Box &B, int u, int c0, int ox) {
  const R h = R(0.5);
  R s = 0;
  int x, y, z;
  coords(B, u, x, y, z);
  if (in_box(B, x, y, z, 1, B.nx, 1, B.ny, 1, B.nz)) {
    const R *g = fib + (size_t)NI * u + c0;
    s += h * (g[0] - g[1]);
  }
  const int v = u - ox;
  if (v >= 0) {
    coords(B, v, x, y, z);
    if (in_box(B, x, y, z, 1, B.nx, 1, B.ny, 1, B.nz)) {
      const R *g = fib + (size_t)NI * v + c0;
      s += h * (g[0] + g[1]);
    }
  }
  return s;
}

template <class R>
__global__ void k_load_T(Fld<R> *Fb, const R *fib, Box B) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= B.nv) return;
  const int sx = 1, sy = B.sy, sz = B.sz;
  Fld<R> o = Fb[u];
  o.ex += corner_T(fib, B, u, 0, sy, sz);
  o.ey += corner_T(fib, B, u, 4, sz, sx);
  o.ez += corner_T(fib, B, u, 8, sx, sy);
  o.cbx += face_T(fib, B, u, 12, sx);
  o.cby += face_T(fib, B, u, 14, sy);
  o.cbz += face_T(fib, B, u, 16, sz);
  Fb[u] = o;
}

// clear_jf + unload_accumulator over voxels 1..n+1
// (unload_accumulator_pipeline.cc:65-67)
template <class R>
__global__ void k_unload(Fld<R> *F, const R *acc, Coef<R> C, Box B) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= B.nv) return;
  int x, y, z;
  coords(B, v, x, y, z);
  Fld<R> &f = F[v];
  f.jfx = f.jfy = f.jfz = 0;
  if (!in_box(B, x, y, z, 1, B.nx + 1, 1, B.ny + 1, 1, B.nz + 1)) return;
  const int sx = 1, sy = B.sy, sz = B.sz;
  const R *a0 = acc + (size_t)12 * v;
  const R *ax = acc + (size_t)12 * (v - sx), *ay = acc + (size_t)12 * (v - sy),
          *az = acc + (size_t)12 * (v - sz);
  const R *ayz = acc + (size_t)12 * (v - sy - sz),
          *azx = acc + (size_t)12 * (v - sz - sx),
          *axy = acc + (size_t)12 * (v - sx - sy);
  f.jfx += C.uc[0] * (a0[0] + ay[1] + az[2] + ayz[3]);
  f.jfy += C.uc[1] * (a0[4] + az[5] + ax[6] + azx[7]);
  f.jfz += C.uc[2] * (a0[8] + ax[9] + ay[10] + axy[11]);
}

// Transpose of k_unload: accumulator adjoint of voxel u from jf-bar.
template <class R>
__global__ void k_unload_T(R *accb, const Fld<R> *Fb, Coef<R> C, Box B) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= B.nv) return;
  const int sx = 1, sy = B.sy, sz = B.sz;
  auto jb = [&](int v, int comp) -> R {
    if (v >= B.nv) return R(0);
    int x, y, z;
    coords(B, v, x, y, z);
    if (!in_box(B, x, y, z, 1, B.nx + 1, 1, B.ny + 1, 1, B.nz + 1))
      return R(0);
    return comp == 0 ? Fb[v].jfx : comp == 1 ? Fb[v].jfy : Fb[v].jfz;
  };
  R *o = accb + (size_t)12 * u;
  o[0] = C.uc[0] * jb(u, 0);
  o[1] = C.uc[0] * jb(u + sy, 0);
  o[2] = C.uc[0] * jb(u + sz, 0);
  o[3] = C.uc[0] * jb(u + sy + sz, 0);
  o[4] = C.uc[1] * jb(u, 1);
  o[5] = C.uc[1] * jb(u + sz, 1);
  o[6] = C.uc[1] * jb(u + sx, 1);
  o[7] = C.uc[1] * jb(u + sz + sx, 1);
  o[8] = C.uc[2] * jb(u, 2);
  o[9] = C.uc[2] * jb(u + sx, 2);
  o[10] = C.uc[2] * jb(u + sy, 2);
  o[11] = C.uc[2] * jb(u + sx + sy, 2);
}

// synchronize_jf on a periodic box, one axis X (remote.cc:474-502): the
// tangential jf of the two face planes X = 1 and X = n+1 both become their
// sum.  The map is symmetric, so it is its own transpose.
template <class R>
__global__ void k_fold_jf(Fld<R> *F, Box B, int X) {
  const int n[3] = {B.nx, B.ny, B.nz};
  const int Y = (X + 1) % 3, Z = (X + 2) % 3;
  const int nY = n[Y] + 1, nZ = n[Z] + 1;
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nY * nZ) return;
  const int yy = 1 + k % nY, zz = 1 + k / nY;
  int c[3];
  c[X] = 1;
  c[Y] = yy;
  c[Z] = zz;
  const int lo = vox(c[0], c[1], c[2], B.nx, B.ny);
  c[X] = n[X] + 1;
  const int hi = vox(c[0], c[1], c[2], B.nx, B.ny);
  R *flo = &F[lo].jfx, *fhi = &F[hi].jfx;
  // jf_Y on (Y in 1..nY, Z in 1..nZ+1), jf_Z on (Y in 1..nY+1, Z in 1..nZ)
  if (yy <= n[Y]) {
    const R s = flo[Y] + fhi[Y];
    flo[Y] = s;
    fhi[Y] = s;
  }
  if (zz <= n[Z]) {
    const R s = flo[Z] + fhi[Z];
    flo[Z] = s;
    fhi[Z] = s;
  }
}

// advance_b (advance_b_pipeline.h:57-59 over voxels 1..n+1)
template <class R>
__global__ void k_advance_b(Fld<R> *F, Coef<R> C, Box B) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= B.nv) return;
  int x, y, z;
  coords(B, v, x, y, z);
  if (!in_box(B, x, y, z, 1, B.nx + 1, 1, B.ny + 1, 1, B.nz + 1)) return;
  const R px = C.bp[0], py = C.bp[1], pz = C.bp[2];
  Fld<R> *f = F + v;
  // the +1 neighbours are read only where they lie inside the array
  if (y <= B.ny && z <= B.nz) {
    const Fld<R> *fy = F + v + B.sy, *fz = F + v + B.sz;
    f->cbx -= (py * (fy->ez - f->ez) - pz * (fz->ey - f->ey));
  }
  if (x <= B.nx && z <= B.nz) {
    const Fld<R> *fx = F + v + 1, *fz = F + v + B.sz;
    f->cby -= (pz * (fz->ex - f->ex) - px * (fx->ez - f->ez));
  }
  if (x <= B.nx && y <= B.ny) {
    const Fld<R> *fx = F + v + 1, *fy = F + v + B.sy;
    f->cbz -= (px * (fx->ey - f->ey) - py * (fy->ex - f->ex));
  }
}

__device__ inline bool in_bx(const Box &B, int v) {
  if (v < 0 || v >= B.nv) return false;
  int x, y, z;
  coords(B, v, x, y, z);
  return in_box(B, x, y, z, 1, B.nx + 1, 1, B.ny, 1, B.nz);
}
__device__ inline bool in_by(const Box &B, int v) {
  if (v < 0 || v >= B.nv) return false;
  int x, y, z;
  coords(B, v, x, y, z);
  return in_box(B, x, y, z, 1, B.nx, 1, B.ny + 1, 1, B.nz);
}
__device__ inline bool in_bz(const Box &B, int v) {
  if (v < 0 || v >= B.nv) return false;
  int x, y, z;
  coords(B, v, x, y, z);
  return in_box(B, x, y, z, 1, B.nx, 1, B.ny, 1, B.nz + 1);
}
__device__ inline bool in_ex(const Box &B, int v) {
  if (v < 0 || v >= B.nv) return false;
  int x, y, z;
  coords(B, v, x, y, z);
  return in_box(B, x, y, z, 1, B.nx, 1, B.ny + 1, 1, B.nz + 1);
}
__device__ inline bool in_ey(const Box &B, int v) {
  if (v < 0 || v >= B.nv) return false;
  int x, y, z;
  coords(B, v, x, y, z);
  return in_box(B, x, y, z, 1, B.nx + 1, 1, B.ny, 1, B.nz + 1);
}
__device__ inline bool in_ez(const Box &B, int v) {
  if (v < 0 || v >= B.nv) return false;
  int x, y, z;
  coords(B, v, x, y, z);
  return in_box(B, x, y, z, 1, B.nx + 1, 1, B.ny + 1, 1, B.nz);
}

// Transpose of k_advance_b: E-bar += T_b^T cb-bar (cb-bar passes through).
template <class R>
__global__ void k_advance_b_T(Fld<R> *Fb, Coef<R> C, Box B) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= B.nv) return;
  const R px = C.bp[0], py = C.bp[1], pz = C.bp[2];
  const int sx = 1, sy = B.sy, sz = B.sz;
  auto bx = [&](int v) { return in_bx(B, v) ? Fb[v].cbx : R(0); };
  auto by = [&](int v) { return in_by(B, v) ? Fb[v].cby : R(0); };
  auto bz = [&](int v) { return in_bz(B, v) ? Fb[v].cbz : R(0); };
  // cbx[v] += -py ez[v+sy] + py ez[v] + pz ey[v+sz] - pz ey[v]
  // cby[v] += -pz ex[v+sz] + pz ex[v] + px ez[v+sx] - px ez[v]
  // cbz[v] += -px ey[v+sx] + px ey[v] + py ex[v+sy] - py ex[v]
  const R exb = pz * (by(u) - by(u - sz)) + py * (bz(u - sy) - bz(u));
  const R eyb = pz * (bx(u - sz) - bx(u)) + px * (bz(u) - bz(u - sx));
  const R ezb = py * (bx(u) - bx(u - sy)) + px * (by(u - sx) - by(u));
  Fb[u].ex += exb;
  Fb[u].ey += eyb;
  Fb[u].ez += ezb;
}

// Tangential cb ghosts of a periodic box (remote.cc:61-134 with the
// self-loop port): the ghost plane X = n+1 takes face plane 1, X = 0 takes
// face plane n, for the two tangential components.
template <class R>
__global__ void k_ghost_b(Fld<R> *F, Box B) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= B.nv) return;
  int c[3];
  coords(B, v, c[0], c[1], c[2]);
  const int n[3] = {B.nx, B.ny, B.nz};
  for (int X = 0; X < 3; X++) {
    const int src = c[X] == 0 ? n[X] : c[X] == n[X] + 1 ? 1 : -1;
    if (src < 0) continue;
    const int Y = (X + 1) % 3, Z = (X + 2) % 3;
    int s[3] = {c[0], c[1], c[2]};
    s[X] = src;
    const int u = vox(s[0], s[1], s[2], B.nx, B.ny);
    R *dst = &F[v].cbx;
    const R *from = &F[u].cbx;
    // cb_Y on (Y in 1..nY+1, Z in 1..nZ); cb_Z on (Y in 1..nY, Z in 1..nZ+1)
    if (c[Y] >= 1 && c[Y] <= n[Y] + 1 && c[Z] >= 1 && c[Z] <= n[Z])
      dst[Y] = from[Y];
    if (c[Y] >= 1 && c[Y] <= n[Y] && c[Z] >= 1 && c[Z] <= n[Z] + 1)
      dst[Z] = from[Z];
  }
}

// Transpose of k_ghost_b, first half: each face-plane element adds the
// adjoints of the ghosts copied from it.
template <class R>
__global__ void k_ghost_b_T_add(Fld<R> *Fb, Box B) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= B.nv) return;
  int c[3];
  coords(B, v, c[0], c[1], c[2]);
  const int n[3] = {B.nx, B.ny, B.nz};
  R add[3] = {0, 0, 0};
  for (int X = 0; X < 3; X++) {
    const int Y = (X + 1) % 3, Z = (X + 2) % 3;
    for (int side = 0; side < 2; side++) {
      // face plane 1 feeds ghost n+1; face plane n feeds ghost 0
      const int face = side == 0 ? 1 : n[X];
      if (c[X] != face) continue;
      int g[3] = {c[0], c[1], c[2]};
      g[X] = side == 0 ? n[X] + 1 : 0;
      const R *gb = &Fb[vox(g[0], g[1], g[2], B.nx, B.ny)].cbx;
      if (c[Y] >= 1 && c[Y] <= n[Y] + 1 && c[Z] >= 1 && c[Z] <= n[Z])
        add[Y] += gb[Y];
      if (c[Y] >= 1 && c[Y] <= n[Y] && c[Z] >= 1 && c[Z] <= n[Z] + 1)
        add[Z] += gb[Z];
    }
  }
  Fb[v].cbx += add[0];
  Fb[v].cby += add[1];
  Fb[v].cbz += add[2];
}

// Second half: the ghosts' adjoints are consumed (the forward overwrote
// the ghosts).
template <class R>
__global__ void k_ghost_b_T_zero(Fld<R> *Fb, Box B) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= B.nv) return;
  int c[3];
  coords(B, v, c[0], c[1], c[2]);
  const int n[3] = {B.nx, B.ny, B.nz};
  R *dst = &Fb[v].cbx;
  for (int X = 0; X < 3; X++) {
    if (c[X] != 0 && c[X] != n[X] + 1) continue;
    const int Y = (X + 1) % 3, Z = (X + 2) % 3;
    if (c[Y] >= 1 && c[Y] <= n[Y] + 1 && c[Z] >= 1 && c[Z] <= n[Z])
      dst[Y] = 0;
    if (c[Y] >= 1 && c[Y] <= n[Y] && c[Z] >= 1 && c[Z] <= n[Z] + 1)
      dst[Z] = 0;
  }
}

// vacuum_advance_e with no damping (vacuum_advance_e_pipeline.h:56-69;
// decay = drive = rmu = 1), voxels 1..n+1.
template <class R>
__global__ void k_advance_e(Fld<R> *F, Coef<R> C, Box B) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= B.nv) return;
  int x, y, z;
  coords(B, v, x, y, z);
  if (!in_box(B, x, y, z, 1, B.nx + 1, 1, B.ny + 1, 1, B.nz + 1)) return;
  const R px = C.ep[0], py = C.ep[1], pz = C.ep[2], cj = C.cj, one = 1;
  Fld<R> &f = F[v];
  const Fld<R> &fx = F[v - 1], &fy = F[v - B.sy], &fz = F[v - B.sz];
  if (x <= B.nx) {
    const R t = py * (f.cbz - fy.cbz) - pz * (f.cby - fz.cby);
    f.ex = one * f.ex + one * (t - cj * f.jfx);
  }
  if (y <= B.ny) {
    const R t = pz * (f.cbx - fz.cbx) - px * (f.cbz - fx.cbz);
    f.ey = one * f.ey + one * (t - cj * f.jfy);
  }
  if (z <= B.nz) {
    const R t = px * (f.cby - fx.cby) - py * (f.cbx - fy.cbx);
    f.ez = one * f.ez + one * (t - cj * f.jfz);
  }
}

// Transpose of k_advance_e: cb-bar += T_e^T E-bar, jf-bar = -cj E-bar
// (E-bar passes through).
template <class R>
__global__ void k_advance_e_T(Fld<R> *Fb, Coef<R> C, Box B) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= B.nv) return;
  const R px = C.ep[0], py = C.ep[1], pz = C.ep[2], cj = C.cj;
  const int sx = 1, sy = B.sy, sz = B.sz;
  auto ex = [&](int v) { return in_ex(B, v) ? Fb[v].ex : R(0); };
  auto ey = [&](int v) { return in_ey(B, v) ? Fb[v].ey : R(0); };
  auto ez = [&](int v) { return in_ez(B, v) ? Fb[v].ez : R(0); };
  // ex[v] += py (cbz[v] - cbz[v-sy]) - pz (cby[v] - cby[v-sz])
  // ey[v] += pz (cbx[v] - cbx[v-sz]) - px (cbz[v] - cbz[v-sx])
  // ez[v] += px (cby[v] - cby[v-sx]) - py (cbx[v] - cbx[v-sy])
  const R cbxb = pz * (ey(u) - ey(u + sz)) - py * (ez(u) - ez(u + sy));
  const R cbyb = px * (ez(u) - ez(u + sx)) - pz * (ex(u) - ex(u + sz));
  const R cbzb = py * (ex(u) - ex(u + sy)) - px * (ey(u) - ey(u + sx));
  Fld<R> &o = Fb[u];
  o.cbx += cbxb;
  o.cby += cbyb;
  o.cbz += cbzb;
  o.jfx = -cj * ex(u);
  o.jfy = -cj * ey(u);
  o.jfz = -cj * ez(u);
}

// The magnetic part of energy_f (vacuum_energy_f_pipeline.h:44-75): per
// voxel R terms, FP64 block sums.
template <class R>
__global__ void k_benergy(const Fld<R> *F, Box B, double *part) {
  __shared__ double red[256];
  double acc = 0;
  const int ncell = B.nx * B.ny * B.nz;
  for (int l = blockIdx.x * blockDim.x + threadIdx.x; l < ncell;
       l += gridDim.x * blockDim.x) {
    const int x = 1 + l % B.nx, y = 1 + (l / B.nx) % B.ny,
              z = 1 + l / (B.nx * B.ny);
    const int v = vox(x, y, z, B.nx, B.ny);
    const R h = R(0.5);
    const Fld<R> &f = F[v], &fx = F[v + 1], &fy = F[v + B.sy],
                 &fz = F[v + B.sz];
    acc += (double)(h * (f.cbx * f.cbx + fx.cbx * fx.cbx));
    acc += (double)(h * (f.cby * f.cby + fy.cby * fy.cby));
    acc += (double)(h * (f.cbz * f.cbz + fz.cbz * fz.cbz));
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// d(energy)/d(field): the energy counts cbx[u] once as a voxel's own face
// and once as its +x neighbour's (likewise y, z).
template <class R>
__global__ void k_benergy_grad(Fld<R> *Fb, const Fld<R> *F, Box B, R scale) {
  const int u = blockIdx.x * blockDim.x + threadIdx.x;
  if (u >= B.nv) return;
  int x, y, z;
  coords(B, u, x, y, z);
  Fld<R> o;
  o.ex = o.ey = o.ez = o.jfx = o.jfy = o.jfz = 0;
  const bool iy = y >= 1 && y <= B.ny, iz = z >= 1 && z <= B.nz,
             ix = x >= 1 && x <= B.nx;
  const R nxc = R((x >= 1 && x <= B.nx) + (x >= 2 && x <= B.nx + 1));
  const R nyc = R((y >= 1 && y <= B.ny) + (y >= 2 && y <= B.ny + 1));
  const R nzc = R((z >= 1 && z <= B.nz) + (z >= 2 && z <= B.nz + 1));
  o.cbx = (iy && iz) ? scale * F[u].cbx * nxc : R(0);
  o.cby = (ix && iz) ? scale * F[u].cby * nyc : R(0);
  o.cbz = (ix && iy) ? scale * F[u].cbz * nzc : R(0);
  Fb[u] = o;
}

// The caller's field array (vpic_field_t, or its FP64 twin with the same
// members as doubles) to the run's fields, and the particles' momentum
// adjoints out.
template <class R>
__global__ void k_fields_in(Fld<R> *F, const R *src, int stride_r, int nv) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const R *s = src + (size_t)stride_r * v;
  Fld<R> o;
  o.ex = s[0];
  o.ey = s[1];
  o.ez = s[2];
  o.cbx = s[4];
  o.cby = s[5];
  o.cbz = s[6];
  o.jfx = o.jfy = o.jfz = 0;
  F[v] = o;
}

template <class R>
__global__ void k_split_pb(const R *pb, R *ub, R *db, int np) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= np) return;
  for (int k = 0; k < 3; k++) {
    ub[3 * (size_t)n + k] = pb[6 * (size_t)n + 3 + k];
    if (db) db[3 * (size_t)n + k] = pb[6 * (size_t)n + k];
  }
}

template <class R>
__global__ void k_fields_out(R *dst, const Fld<R> *Fb, int nv) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= nv) return;
  const Fld<R> &f = Fb[v];
  R *o = dst + 6 * (size_t)v;
  o[0] = f.ex;
  o[1] = f.ey;
  o[2] = f.ez;
  o[3] = f.cbx;
  o[4] = f.cby;
  o[5] = f.cbz;
}

// ---------------------------------------------------------------------------
// host driver
// ---------------------------------------------------------------------------

struct Buf {
  void *p = nullptr;
  ~Buf() {
    if (p) (void)hipFree(p);
  }
};

int alloc(Buf &b, size_t bytes) {
  hipError_t e = hipMalloc(&b.p, bytes < 16 ? 16 : bytes);
  if (e != hipSuccess) {
    b.p = nullptr;
    return vpic_set_error(VPIC_ENOMEM, "adjoint: hipMalloc(%zu): %s", bytes,
                          hipGetErrorString(e));
  }
  return VPIC_OK;
}

inline dim3 cover(int64_t n, int bs = 256) {
  int64_t b = (n + bs - 1) / bs;
  return dim3((unsigned)(b < 1 ? 1 : b));
}

template <class R>
struct Run {
  Box B;
  GridDev G;
  Coef<R> C;
  int np = 0;
  hipStream_t s = nullptr;
  Part<R> *p = nullptr;
  Fld<R> *F = nullptr;
  R *fi = nullptr, *acc = nullptr;
  int *err = nullptr;
  double *part = nullptr;
  int eblocks = 256;
  double escale = 0;  // 0.5 eps0 dV

  int setup(const vpic_grid_t *g, int np_, float q, float m, hipStream_t st) {
    B = {g->nx, g->ny, g->nz, g->sy, g->sz, g->nv};
    G = make_grid_dev(g);
    G.neighbor = nullptr;
    np = np_;
    s = st;
    // advance_p_pipeline.cc:279-283 (in float for the FP32 run)
    const R rq = q, rm = m, dt = g->dt, c = g->cvac;
    C.qdt_2mc = (rq * dt) / (2 * rm * c);
    C.cdt[0] = c * dt * R(g->rdx);
    C.cdt[1] = c * dt * R(g->rdy);
    C.cdt[2] = c * dt * R(g->rdz);
    C.qsp = rq;
    const R frac = R(0.5);
    C.bp[0] = g->nx > 1 ? frac * c * dt * R(g->rdx) : R(0);
    C.bp[1] = g->ny > 1 ? frac * c * dt * R(g->rdy) : R(0);
    C.bp[2] = g->nz > 1 ? frac * c * dt * R(g->rdz) : R(0);
    const R one = 1;
    C.ep[0] = (g->nx > 1 ? (one + R(0)) * c * dt * R(g->rdx) : R(0)) * one;
    C.ep[1] = (g->ny > 1 ? (one + R(0)) * c * dt * R(g->rdy) : R(0)) * one;
    C.ep[2] = (g->nz > 1 ? (one + R(0)) * c * dt * R(g->rdz) : R(0)) * one;
    C.cj = dt / R(g->eps0);
    // unload_accumulator_pipeline.cc:137-139: in double, stored in R
    C.uc[0] = (R)(0.25 * g->rdy * g->rdz / g->dt);
    C.uc[1] = (R)(0.25 * g->rdz * g->rdx / g->dt);
    C.uc[2] = (R)(0.25 * g->rdx * g->rdy / g->dt);
    escale = 0.5 * (double)g->eps0 * (double)g->dV;
    return VPIC_OK;
  }

#define ADJ_LAUNCH(kern, grid, ...)                                   \
  do {                                                                \
    hipLaunchKernelGGL(kern, grid, dim3(256), 0, s, __VA_ARGS__);     \
    VPIC_LAUNCHED(#kern);                                             \
  } while (0)

  int load() {
    ADJ_LAUNCH(k_load<R>, cover(B.nv), fi, (const Fld<R> *)F, B);
    return VPIC_OK;
  }

  // One step (advance.cc:25-185 on a periodic box, one species); fi holds
  // the interpolators of the step's fields on entry and of the new fields
  // on return.
  int step() {
    ADJ_LAUNCH(k_fill<R>, cover(12 * (int64_t)B.nv), acc, (size_t)12 * B.nv,
               R(0));
    if (np > 0) {
      int64_t want = ((int64_t)np + 255) / 256;
      dim3 grid((unsigned)(want < 8192 ? want : 8192));
      ADJ_LAUNCH(k_push_fwd<R>, grid, p, np, (const R *)fi, acc, C, G, err);
    }
    ADJ_LAUNCH(k_unload<R>, cover(B.nv), F, (const R *)acc, C, B);
    for (int X = 0; X < 3; X++) {
      const int n[3] = {B.nx, B.ny, B.nz};
      ADJ_LAUNCH(k_fold_jf<R>,
                 cover((int64_t)(n[(X + 1) % 3] + 1) * (n[(X + 2) % 3] + 1)), F,
                 B, X);
    }
    ADJ_LAUNCH(k_advance_b<R>, cover(B.nv), F, C, B);
    ADJ_LAUNCH(k_ghost_b<R>, cover(B.nv), F, B);
    ADJ_LAUNCH(k_advance_e<R>, cover(B.nv), F, C, B);
    ADJ_LAUNCH(k_advance_b<R>, cover(B.nv), F, C, B);
    return load();
  }

  int energy(double *out) {
    ADJ_LAUNCH(k_benergy<R>, dim3(eblocks), (const Fld<R> *)F, B, part);
    std::vector<double> h(eblocks);
    VPIC_HIP(hipMemcpyAsync(h.data(), part, sizeof(double) * eblocks,
                            hipMemcpyDeviceToHost, s));
    VPIC_HIP(hipStreamSynchronize(s));
    double e = 0;
    for (double v : h) e += v;
    *out = e * escale;
    return VPIC_OK;
  }

  int check_err() {
    int h = 0;
    VPIC_HIP(hipMemcpyAsync(&h, err, sizeof(int), hipMemcpyDeviceToHost, s));
    VPIC_HIP(hipStreamSynchronize(s));
    if (h)
      return vpic_set_error(VPIC_EINVAL,
                            "adjoint run: %d particles left the periodic box "
                            "(the differentiated run has no particle "
                            "boundaries)", h);
    return VPIC_OK;
  }

  // Reverse of one step from its starting state (p_n, F_n), with pb and Fb
  // the adjoint of the state after the step on entry, before it on return.
  int step_T(const Part<R> *pn, const Fld<R> *Fn, R *pb, Fld<R> *Fb, R *fib,
             R *accb) {
    ADJ_LAUNCH(k_advance_b_T<R>, cover(B.nv), Fb, C, B);
    ADJ_LAUNCH(k_advance_e_T<R>, cover(B.nv), Fb, C, B);
    ADJ_LAUNCH(k_ghost_b_T_add<R>, cover(B.nv), Fb, B);
    ADJ_LAUNCH(k_ghost_b_T_zero<R>, cover(B.nv), Fb, B);
    ADJ_LAUNCH(k_advance_b_T<R>, cover(B.nv), Fb, C, B);
    for (int X = 2; X >= 0; X--) {
      const int n[3] = {B.nx, B.ny, B.nz};
      ADJ_LAUNCH(k_fold_jf<R>,
                 cover((int64_t)(n[(X + 1) % 3] + 1) * (n[(X + 2) % 3] + 1)),
                 Fb, B, X);
    }
    ADJ_LAUNCH(k_unload_T<R>, cover(B.nv), accb, (const Fld<R> *)Fb, C, B);
    ADJ_LAUNCH(k_load<R>, cover(B.nv), fi, Fn, B);
    ADJ_LAUNCH(k_fill<R>, cover((int64_t)NI * B.nv), fib, (size_t)NI * B.nv,
               R(0));
    if (np > 0) {
      int64_t want = ((int64_t)np + 255) / 256;
      dim3 grid((unsigned)(want < 8192 ? want : 8192));
      ADJ_LAUNCH(k_push_rev<R>, grid, pn, np, (const R *)fi, (const R *)accb,
                 pb, fib, C, G);
    }
    ADJ_LAUNCH(k_load_T<R>, cover(B.nv), Fb, (const R *)fib, B);
    return VPIC_OK;
  }
#undef ADJ_LAUNCH
};

int check_periodic(const vpic_grid_t *g) {
  if (!g || g->nx < 1 || g->ny < 1 || g->nz < 1)
    return vpic_set_error(VPIC_EINVAL, "adjoint run: bad grid");
  if (g->sy != g->nx + 2 || g->sz != (g->ny + 2) * g->sy ||
      g->nv != (g->nz + 2) * g->sz)
    return vpic_set_error(VPIC_EINVAL, "adjoint run: inconsistent strides");
  for (int k = 0; k < 6; k++)
    if (g->pbc[k] != VPIC_BC_PERIODIC || g->fbc[k] != VPIC_BC_PERIODIC)
      return vpic_set_error(VPIC_EUNSUPPORTED,
                            "adjoint run: only periodic boxes are "
                            "differentiated (face %d)", k);
  return VPIC_OK;
}

// The forward with checkpoints, then (when grads are wanted) the reverse.
template <class R>
int adjoint_run(const void *p0, int np, float q, float m, const void *f0,
                const vpic_grid_t *g, int n_step, int ckpt, double *host_L,
                void *u_bar, void *d_bar, void *f_bar, hipStream_t s) {
  int rc;
  if ((rc = check_periodic(g))) return rc;
  if (np < 0 || n_step < 0 || (np > 0 && !p0) || !f0 || !host_L)
    return vpic_set_error(VPIC_EINVAL, "adjoint run: bad arguments");
  const bool grads = u_bar != nullptr;
  if (ckpt < 1) ckpt = n_step > 0 ? n_step : 1;
  Run<R> run;
  run.setup(g, np, q, m, s);
  const int nv = g->nv;
  const size_t pbytes = sizeof(Part<R>) * (size_t)np;
  const size_t fbytes = sizeof(Fld<R>) * (size_t)nv;
  Buf bp, bF, bfi, bacc, berr, bpart;
  if ((rc = alloc(bp, pbytes)) || (rc = alloc(bF, fbytes)) ||
      (rc = alloc(bfi, sizeof(R) * NI * (size_t)nv)) ||
      (rc = alloc(bacc, sizeof(R) * 12 * (size_t)nv)) ||
      (rc = alloc(berr, 64)) ||
      (rc = alloc(bpart, sizeof(double) * run.eblocks)))
    return rc;
  run.p = (Part<R> *)bp.p;
  run.F = (Fld<R> *)bF.p;
  run.fi = (R *)bfi.p;
  run.acc = (R *)bacc.p;
  run.err = (int *)berr.p;
  run.part = (double *)bpart.p;
  VPIC_HIP(hipMemsetAsync(run.err, 0, 64, s));
  if (np) VPIC_HIP(hipMemcpyAsync(run.p, p0, pbytes, hipMemcpyDeviceToDevice, s));
  // the caller's fields: 20 (float) or 18 (double) R per record
  const int rstride = sizeof(R) == 4 ? 20 : 18;
  hipLaunchKernelGGL(k_fields_in<R>, cover(nv), dim3(256), 0, s, run.F,
                     (const R *)f0, rstride, nv);
  VPIC_LAUNCHED("k_fields_in");
  if ((rc = run.load())) return rc;
  const int nseg = grads ? (n_step + ckpt - 1) / ckpt : 0;
  std::vector<Buf> cp(nseg), cF(nseg);
  for (int k = 0; k < nseg; k++)
    if ((rc = alloc(cp[k], pbytes)) || (rc = alloc(cF[k], fbytes))) return rc;
  for (int n = 0; n < n_step; n++) {
    if (grads && n % ckpt == 0) {
      const int k = n / ckpt;
      if (np)
        VPIC_HIP(hipMemcpyAsync(cp[k].p, run.p, pbytes,
                                hipMemcpyDeviceToDevice, s));
      VPIC_HIP(hipMemcpyAsync(cF[k].p, run.F, fbytes, hipMemcpyDeviceToDevice,
                              s));
    }
    if ((rc = run.step())) return rc;
  }
  if ((rc = run.check_err())) return rc;
  if ((rc = run.energy(host_L))) return rc;
  if (!grads) return VPIC_OK;

  // reverse
  Buf bpb, bFb, bfib, baccb;
  std::vector<Buf> sp(ckpt), sF(ckpt);
  if ((rc = alloc(bpb, sizeof(R) * 6 * (size_t)np)) ||
      (rc = alloc(bFb, fbytes)) ||
      (rc = alloc(bfib, sizeof(R) * NI * (size_t)nv)) ||
      (rc = alloc(baccb, sizeof(R) * 12 * (size_t)nv)))
    return rc;
  for (int j = 0; j < ckpt; j++)
    if ((rc = alloc(sp[j], pbytes)) || (rc = alloc(sF[j], fbytes))) return rc;
  R *pb = (R *)bpb.p, *fib = (R *)bfib.p, *accb = (R *)baccb.p;
  Fld<R> *Fb = (Fld<R> *)bFb.p;
  hipLaunchKernelGGL(k_fill<R>, cover(6 * (int64_t)np), dim3(256), 0, s, pb,
                     (size_t)6 * np, R(0));
  VPIC_LAUNCHED("k_fill");
  hipLaunchKernelGGL(k_benergy_grad<R>, cover(nv), dim3(256), 0, s, Fb,
                     (const Fld<R> *)run.F, run.B, (R)run.escale);
  VPIC_LAUNCHED("k_benergy_grad");
  for (int k = nseg - 1; k >= 0; k--) {
    const int n0 = k * ckpt, n1 = n0 + ckpt < n_step ? n0 + ckpt : n_step;
    if (np)
      VPIC_HIP(hipMemcpyAsync(run.p, cp[k].p, pbytes, hipMemcpyDeviceToDevice,
                              s));
    VPIC_HIP(hipMemcpyAsync(run.F, cF[k].p, fbytes, hipMemcpyDeviceToDevice,
                            s));
    if ((rc = run.load())) return rc;
    for (int n = n0; n < n1; n++) {
      const int j = n - n0;
      if (np)
        VPIC_HIP(hipMemcpyAsync(sp[j].p, run.p, pbytes,
                                hipMemcpyDeviceToDevice, s));
      VPIC_HIP(hipMemcpyAsync(sF[j].p, run.F, fbytes, hipMemcpyDeviceToDevice,
                              s));
      if (n + 1 < n1 && (rc = run.step())) return rc;
    }
    for (int n = n1 - 1; n >= n0; n--) {
      const int j = n - n0;
      if ((rc = run.step_T((const Part<R> *)sp[j].p, (const Fld<R> *)sF[j].p,
                           pb, Fb, fib, accb)))
        return rc;
    }
  }
  hipLaunchKernelGGL(k_split_pb<R>, cover(np), dim3(256), 0, s,
                     (const R *)pb, (R *)u_bar, (R *)d_bar, np);
  VPIC_LAUNCHED("k_split_pb");
  if (f_bar) {
    hipLaunchKernelGGL(k_fields_out<R>, cover(nv), dim3(256), 0, s,
                       (R *)f_bar, (const Fld<R> *)Fb, nv);
    VPIC_LAUNCHED("k_fields_out");
  }
  VPIC_HIP(hipStreamSynchronize(s));
  return VPIC_OK;
}

}  // namespace

extern "C" int vpic_hip_adjoint_bfield_energy(
    const void *p0, int32_t np, float q, float m, const void *f0,
    const vpic_grid_t *g, int32_t n_step, int32_t ckpt_interval,
    int32_t precision, double *host_energy, void *u_bar, void *d_bar,
    void *f_bar, void *stream) {
  const hipStream_t s = (hipStream_t)stream;
  if (precision == 4)
    return adjoint_run<float>(p0, np, q, m, f0, g, n_step, ckpt_interval,
                              host_energy, u_bar, d_bar, f_bar, s);
  if (precision == 8)
    return adjoint_run<double>(p0, np, q, m, f0, g, n_step, ckpt_interval,
                               host_energy, u_bar, d_bar, f_bar, s);
  return vpic_set_error(VPIC_EINVAL, "adjoint run: precision %d (4 or 8)",
                        precision);
}
