// ensemble.hip -- many independent small periodic boxes advanced together:
// the packed layout, the batched step (one persistent launch per phase over
// a work queue of (box, range) items, captured once in a hipGraph and
// replayed), the tuner of the launch shapes, and the C-ABI.
//
// Layout: box b's voxels (nv_b = (nx+2)(ny+2)(nz+2)) start at voxel offset
// voff_b of one packed field, interpolator and accumulator array; its
// particles start at particle offset poff_b of one packed particle array.
// Every array keeps the reference's per-box layout (particle_t with the
// box-local voxel index, field_t, interpolator_t, accumulator_t), so a box
// is a plain single-grid run at an offset.
//
// The step of every box is vpic_simulation::advance's hot path
// (src/vpic/advance.cc:25-185) for one pushed species (immobile ions are a
// neutralising background the step never sees): clear_accumulator_array,
// advance_p, unload_accumulator with clear_jf and synchronize_jf,
// advance_b(0.5), the tangential-B ghosts and vacuum_advance_e,
// advance_b(0.5), load_interpolator_array.  Each phase is one launch for
// all boxes; the jf fold runs one launch per axis (edges are shared between
// axes, as in the reference's x, y, z order).  Work queues are reset by a
// kernel at the start of each step (no memset nodes in the graph).
#include "common.h"

#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

namespace {

enum {
  Q_PUSH,
  Q_UNLOAD,
  Q_FOLDX,
  Q_FOLDY,
  Q_FOLDZ,
  Q_ADVB1,
  Q_GHOST,
  Q_ADVE,
  Q_ADVB2,
  Q_LOAD,
  Q_TILES,  // the tiled push's (box, tile) queue (items held by EnsTiled)
  Q_COUNT
};

struct Shape {
  int push_bs, push_chunk, field_bs, field_chunk, blocks_per_cu;
};

// Tile of the tiled push (cells per axis, clamped to each box); 0: the
// flat push.
struct TileShape {
  int tx, ty, tz;
  int halo;  // cells of the LDS window beyond the tile (1 to ENS_MAX_HALO)
};

}  // namespace

struct vpic_ensemble {
  int nbox = 0;
  int device = 0, ncu = 256;
  std::vector<vpic_grid_t> grids;
  std::vector<int> np;
  std::vector<float> q, m;
  std::vector<int64_t> voff, poff;
  int64_t nv_total = 0, np_total = 0;
  vpic_particle_t *P = nullptr;
  vpic_field_t *F = nullptr;
  vpic_interpolator_t *FI = nullptr;
  vpic_accumulator_t *A = nullptr;
  EnsPushBox *pbox = nullptr;
  EnsFieldBox *fbox = nullptr;
  EnsItem *items = nullptr;  // all phases' items, concatenated
  int item_off[Q_COUNT + 1] = {0};
  int *queues = nullptr;     // Q_COUNT counters + nbad[2]
  Shape shape = {256, 4096, 256, 2048, 4};
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  // sort_p of every box every sort_interval steps (0: never), before the
  // step whose count of steps done is a multiple of it, as the reference's
  // advance sorts a species (advance.cc:31-37); its own captured graph
  int sort_interval = 20;
  vpic_particle_t *aux = nullptr;
  int *skeys = nullptr, *scount = nullptr, *sstart = nullptr,
      *ssums = nullptr;
  hipGraph_t sgraph = nullptr;
  hipGraphExec_t sexec = nullptr;
  // the tiled push (after a first sort: it reads the sort's partitions)
  TileShape tile = {16, 8, 4, 1};
  EnsTiled *tiled = nullptr;
  bool sorted_once = false;
  bool graph_tiled = false;  // the captured step uses the tiled push
  hipStream_t stream = nullptr;
  bool own_stream = false;
  int64_t steps = 0;

  ~vpic_ensemble() {
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    if (sexec) (void)hipGraphExecDestroy(sexec);
    if (sgraph) (void)hipGraphDestroy(sgraph);
    if (tiled) ens_tiled_destroy(tiled);
    if (own_stream && stream) (void)hipStreamDestroy(stream);
    void *bufs[] = {P,   F,     FI,     A,      pbox,  fbox,  items,
                    queues, aux, skeys, scount, sstart, ssums};
    for (void *b : bufs)
      if (b) (void)hipFree(b);
  }
};

namespace {

int ens_alloc(void **p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes < 16 ? 16 : bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    return vpic_set_error(VPIC_ENOMEM, "ensemble: hipMalloc(%zu): %s", bytes,
                          hipGetErrorString(e));
  }
  return VPIC_OK;
}

void drop_graph(vpic_ensemble *e) {
  if (e->exec) (void)hipGraphExecDestroy(e->exec);
  if (e->graph) (void)hipGraphDestroy(e->graph);
  if (e->sexec) (void)hipGraphExecDestroy(e->sexec);
  if (e->sgraph) (void)hipGraphDestroy(e->sgraph);
  e->exec = nullptr;
  e->graph = nullptr;
  e->sexec = nullptr;
  e->sgraph = nullptr;
}

// The sort's buffers: the aux particle array, keys, counts and offsets.
int ensure_sort_buffers(vpic_ensemble *e) {
  if (e->aux) return VPIC_OK;
  const int64_t nslots = e->nv_total + 1;
  const int64_t nsums = (nslots + 2047) / 2048 + 1;
  int rc;
  if ((rc = ens_alloc((void **)&e->aux,
                      sizeof(vpic_particle_t) * e->np_total)) ||
      (rc = ens_alloc((void **)&e->skeys, sizeof(int) * e->np_total)) ||
      (rc = ens_alloc((void **)&e->scount, sizeof(int) * nslots)) ||
      (rc = ens_alloc((void **)&e->sstart, sizeof(int) * nslots)) ||
      (rc = ens_alloc((void **)&e->ssums, sizeof(int) * nsums))) {
    void *bufs[] = {e->aux, e->skeys, e->scount, e->sstart, e->ssums};
    for (void *b : bufs)
      if (b) (void)hipFree(b);
    e->aux = nullptr;
    e->skeys = e->scount = e->sstart = e->ssums = nullptr;
    return rc;
  }
  return VPIC_OK;
}

// The tiled push's state for the current tile shape (none for tile 0).
int ensure_tiled(vpic_ensemble *e) {
  if (e->tiled || e->tile.tx <= 0) return VPIC_OK;
  return ens_tiled_create(&e->tiled, e->nbox, e->grids.data(), e->np.data(),
                          e->q.data(), e->m.data(), e->voff.data(),
                          e->poff.data(), e->P, e->FI, e->A, e->tile.tx,
                          e->tile.ty, e->tile.tz, e->tile.halo);
}

int enqueue_sort(vpic_ensemble *e, hipStream_t s) {
  return launch_ens_sort(e->P, e->aux, e->pbox,
                         e->items + e->item_off[Q_PUSH],
                         e->item_off[Q_PUSH + 1] - e->item_off[Q_PUSH],
                         e->nbox, e->np_total, e->nv_total, e->skeys,
                         e->scount, e->sstart, e->ssums, e->queues + Q_COUNT,
                         ens_tiled_epart(e->tiled), s);
}

// The step uses the tiled push once a sort has written the partitions.
bool use_tiled(const vpic_ensemble *e) {
  return e->tiled && e->sorted_once;
}

// Work items of every phase for the current shape.
int build_items(vpic_ensemble *e) {
  std::vector<EnsItem> all;
  std::vector<int> off(Q_COUNT + 1, 0);
  auto add_range = [&](int box, int axis, int64_t n, int chunk) {
    for (int64_t f = 0; f < n; f += chunk) {
      EnsItem it;
      it.box = box;
      it.axis = axis;
      it.first = (int)f;
      it.count = (int)std::min<int64_t>(chunk, n - f);
      all.push_back(it);
    }
  };
  for (int q = 0; q < Q_COUNT; q++) {
    off[q] = (int)all.size();
    for (int b = 0; b < e->nbox; b++) {
      const vpic_grid_t &g = e->grids[b];
      const int n[3] = {g.nx, g.ny, g.nz};
      if (q == Q_TILES) {
        // (box, tile) items live with the tiled push's state
      } else if (q == Q_PUSH) {
        add_range(b, -1, e->np[b], e->shape.push_chunk);
      } else if (q == Q_FOLDX || q == Q_FOLDY || q == Q_FOLDZ) {
        const int X = q - Q_FOLDX;
        add_range(b, X, (int64_t)(n[(X + 1) % 3] + 1) * (n[(X + 2) % 3] + 1),
                  e->shape.field_chunk);
      } else if (q == Q_GHOST) {
        for (int X = 0; X < 3; X++)
          add_range(b, X,
                    (int64_t)(n[(X + 1) % 3] + 1) * (n[(X + 2) % 3] + 1),
                    e->shape.field_chunk);
      } else {
        add_range(b, -1, (int64_t)(g.nx + 1) * (g.ny + 1) * (g.nz + 1),
                  e->shape.field_chunk);
      }
    }
  }
  off[Q_COUNT] = (int)all.size();
  if (e->items) (void)hipFree(e->items);
  e->items = nullptr;
  int rc;
  if ((rc = ens_alloc((void **)&e->items, sizeof(EnsItem) * all.size())))
    return rc;
  VPIC_HIP(hipMemcpy(e->items, all.data(), sizeof(EnsItem) * all.size(),
                     hipMemcpyHostToDevice));
  for (int q = 0; q <= Q_COUNT; q++) e->item_off[q] = off[q];
  drop_graph(e);
  return VPIC_OK;
}

int field_phase(vpic_ensemble *e, int q, int phase, hipStream_t s) {
  const int nblk = e->shape.blocks_per_cu * e->ncu;
  return launch_ens_field(phase, e->F, e->FI, e->A, e->fbox,
                          e->items + e->item_off[q],
                          e->item_off[q + 1] - e->item_off[q], e->queues + q,
                          nblk, e->shape.field_bs, s);
}

// One step of every box, enqueued on stream s.
int enqueue_step(vpic_ensemble *e, hipStream_t s) {
  int rc;
  if ((rc = launch_ens_reset(e->queues, Q_COUNT, s))) return rc;
  if ((rc = launch_clear_accumulator_n(e->A, e->nv_total, s))) return rc;
  const int nblk = e->shape.blocks_per_cu * e->ncu;
  if (use_tiled(e)) {
    // one persistent workgroup per CU (the LDS window takes the CU)
    if ((rc = launch_ens_push_tiled(e->tiled, e->queues + Q_TILES, e->ncu, s)))
      return rc;
  } else if ((rc = launch_ens_push(e->P, e->FI, e->A, e->pbox,
                                   e->items + e->item_off[Q_PUSH],
                                   e->item_off[Q_PUSH + 1] -
                                       e->item_off[Q_PUSH],
                                   e->queues + Q_PUSH, e->queues + Q_COUNT,
                                   nblk, e->shape.push_bs, s))) {
    return rc;
  }
  if ((rc = field_phase(e, Q_UNLOAD, ENS_UNLOAD, s)) ||
      (rc = field_phase(e, Q_FOLDX, ENS_FOLD, s)) ||
      (rc = field_phase(e, Q_FOLDY, ENS_FOLD, s)) ||
      (rc = field_phase(e, Q_FOLDZ, ENS_FOLD, s)) ||
      (rc = field_phase(e, Q_ADVB1, ENS_ADVB, s)) ||
      (rc = field_phase(e, Q_GHOST, ENS_GHOST, s)) ||
      (rc = field_phase(e, Q_ADVE, ENS_ADVE, s)) ||
      (rc = field_phase(e, Q_ADVB2, ENS_ADVB, s)) ||
      (rc = field_phase(e, Q_LOAD, ENS_LOAD, s)))
    return rc;
  return VPIC_OK;
}

// Capture what `enqueue` puts on the ensemble's stream into a graph.
int capture_one(vpic_ensemble *e, int (*enqueue)(vpic_ensemble *, hipStream_t),
                hipGraph_t *graph, hipGraphExec_t *exec) {
  hipStream_t s = e->stream;
  VPIC_HIP(hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal));
  int rc = enqueue(e, s);
  hipGraph_t g = nullptr;
  hipError_t ce = hipStreamEndCapture(s, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ce != hipSuccess) return vpic_check_hip(ce, "hipStreamEndCapture");
  *graph = g;
  VPIC_HIP(hipGraphInstantiate(exec, g, nullptr, nullptr, 0));
  return VPIC_OK;
}

int capture(vpic_ensemble *e) {
  drop_graph(e);
  int rc;
  e->graph_tiled = use_tiled(e);
  if ((rc = capture_one(e, enqueue_step, &e->graph, &e->exec))) return rc;
  if (e->sort_interval > 0) {
    if ((rc = ensure_sort_buffers(e)) ||
        (rc = capture_one(e, enqueue_sort, &e->sgraph, &e->sexec)))
      return rc;
  }
  return VPIC_OK;
}

int run_steps(vpic_ensemble *e, int n, int use_graph) {
  int rc;
  if (e->sort_interval > 0 &&
      ((rc = ensure_sort_buffers(e)) || (rc = ensure_tiled(e))))
    return rc;
  for (int k = 0; k < n; k++) {
    if (e->sort_interval > 0 && e->steps % e->sort_interval == 0) {
      if (use_graph && !e->sexec && (rc = capture(e))) return rc;
      if (use_graph)
        VPIC_HIP(hipGraphLaunch(e->sexec, e->stream));
      else if ((rc = enqueue_sort(e, e->stream)))
        return rc;
      e->sorted_once = true;
    }
    // (re)capture when the step's push changed (first sort, shape change)
    if (use_graph && (!e->exec || e->graph_tiled != use_tiled(e)) &&
        (rc = capture(e)))
      return rc;
    if (use_graph)
      VPIC_HIP(hipGraphLaunch(e->exec, e->stream));
    else if ((rc = enqueue_step(e, e->stream)))
      return rc;
    e->steps++;
  }
  return VPIC_OK;
}

int check_box(const vpic_ensemble *e, int b) {
  if (!e) return vpic_set_error(VPIC_EINVAL, "ensemble is NULL");
  if (b < 0 || b >= e->nbox)
    return vpic_set_error(VPIC_EINVAL, "ensemble: box %d of %d", b, e->nbox);
  return VPIC_OK;
}

}  // namespace

extern "C" int vpic_hip_ensemble_create(vpic_ensemble_t **out, int32_t nbox,
                                        const vpic_grid_t *grids,
                                        const int32_t *np, const float *q,
                                        const float *m, void *stream) {
  if (!out || nbox < 1 || !grids || !np || !q || !m)
    return vpic_set_error(VPIC_EINVAL, "ensemble_create: bad arguments");
  *out = nullptr;
  vpic_ensemble *e = new vpic_ensemble;
  e->nbox = nbox;
  e->stream = (hipStream_t)stream;
  if (!e->stream) {
    // a graph cannot be captured on the legacy default stream: the ensemble
    // keeps a stream of its own (default priority)
    hipError_t se = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (se != hipSuccess) {
      delete e;
      return vpic_check_hip(se, "ensemble_create: hipStreamCreateWithFlags");
    }
    e->own_stream = true;
  }
  (void)hipGetDevice(&e->device);
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, e->device) == hipSuccess)
    e->ncu = prop.multiProcessorCount;
  int64_t voff = 0, poff = 0;
  for (int b = 0; b < nbox; b++) {
    const vpic_grid_t &g = grids[b];
    bool periodic = true;
    for (int k = 0; k < 6; k++)
      periodic &= g.pbc[k] == VPIC_BC_PERIODIC && g.fbc[k] == VPIC_BC_PERIODIC;
    if (!periodic || g.nx < 1 || g.ny < 1 || g.nz < 1 || np[b] < 1 ||
        g.sy != g.nx + 2 || g.sz != (g.ny + 2) * g.sy ||
        g.nv != (g.nz + 2) * g.sz || g.neighbor) {
      delete e;
      return vpic_set_error(VPIC_EINVAL,
                            "ensemble_create: box %d must be a periodic grid "
                            "with consistent strides, no neighbor table and "
                            "at least one particle", b);
    }
    e->grids.push_back(g);
    e->np.push_back(np[b]);
    e->q.push_back(q[b]);
    e->m.push_back(m[b]);
    e->voff.push_back(voff);
    e->poff.push_back(poff);
    voff += g.nv;
    poff += np[b];
  }
  e->nv_total = voff;
  e->np_total = poff;
  if (voff >= INT32_MAX || poff >= INT32_MAX) {
    delete e;
    return vpic_set_error(VPIC_EINVAL,
                          "ensemble_create: %lld voxels and %lld particles "
                          "in all (each must stay below 2^31)",
                          (long long)voff, (long long)poff);
  }
  int rc;
  if ((rc = ens_alloc((void **)&e->P, sizeof(vpic_particle_t) * poff)) ||
      (rc = ens_alloc((void **)&e->F, sizeof(vpic_field_t) * voff)) ||
      (rc = ens_alloc((void **)&e->FI, sizeof(vpic_interpolator_t) * voff)) ||
      (rc = ens_alloc((void **)&e->A, sizeof(vpic_accumulator_t) * voff)) ||
      (rc = ens_alloc((void **)&e->pbox, sizeof(EnsPushBox) * nbox)) ||
      (rc = ens_alloc((void **)&e->fbox, sizeof(EnsFieldBox) * nbox)) ||
      (rc = ens_alloc((void **)&e->queues, sizeof(int) * (Q_COUNT + 2)))) {
    delete e;
    return rc;
  }
  std::vector<EnsPushBox> pb(nbox);
  std::vector<EnsFieldBox> fb(nbox);
  vpic_field_params_t mp;
  mp.decayx = mp.decayy = mp.decayz = 1.0f;
  mp.drivex = mp.drivey = mp.drivez = 1.0f;
  mp.rmux = mp.rmuy = mp.rmuz = 1.0f;
  mp.epsx = mp.epsy = mp.epsz = 1.0f;
  mp.damp = 0.0f;
  mp.nonconductive = 1.0f;
  for (int b = 0; b < nbox; b++) {
    const vpic_grid_t &g = e->grids[b];
    // advance_p_pipeline.cc:279-283, in float
    pb[b].qdt_2mc = (e->q[b] * g.dt) / (2 * e->m[b] * g.cvac);
    pb[b].cdt_dx = g.cvac * g.dt * g.rdx;
    pb[b].cdt_dy = g.cvac * g.dt * g.rdy;
    pb[b].cdt_dz = g.cvac * g.dt * g.rdz;
    pb[b].qsp = e->q[b];
    pb[b].G = make_grid_dev(&g);
    pb[b].poff = e->poff[b];
    pb[b].voff = e->voff[b];
    ens_make_field_box(&fb[b], &g, &mp, e->voff[b]);
  }
  hipError_t he;
  if ((he = hipMemcpy(e->pbox, pb.data(), sizeof(EnsPushBox) * nbox,
                      hipMemcpyHostToDevice)) != hipSuccess ||
      (he = hipMemcpy(e->fbox, fb.data(), sizeof(EnsFieldBox) * nbox,
                      hipMemcpyHostToDevice)) != hipSuccess ||
      (he = hipMemset(e->F, 0, sizeof(vpic_field_t) * voff)) != hipSuccess ||
      (he = hipMemset(e->FI, 0, sizeof(vpic_interpolator_t) * voff)) !=
          hipSuccess ||
      (he = hipMemset(e->queues, 0, sizeof(int) * (Q_COUNT + 2))) !=
          hipSuccess) {
    delete e;
    return vpic_check_hip(he, "ensemble_create");
  }
  if ((rc = build_items(e))) {
    delete e;
    return rc;
  }
  *out = e;
  return VPIC_OK;
}

extern "C" int vpic_hip_ensemble_destroy(vpic_ensemble_t *e) {
  delete e;
  return VPIC_OK;
}

extern "C" int vpic_hip_ensemble_box(vpic_ensemble_t *e, int32_t b,
                                     vpic_particle_t **p, int32_t *np,
                                     vpic_field_t **f,
                                     vpic_interpolator_t **fi,
                                     vpic_accumulator_t **a) {
  int rc;
  if ((rc = check_box(e, b))) return rc;
  if (p) *p = e->P + e->poff[b];
  if (np) *np = e->np[b];
  if (f) *f = e->F + e->voff[b];
  if (fi) *fi = e->FI + e->voff[b];
  if (a) *a = e->A + e->voff[b];
  return VPIC_OK;
}

extern "C" int vpic_hip_ensemble_prepare(vpic_ensemble_t *e) {
  if (!e) return vpic_set_error(VPIC_EINVAL, "ensemble is NULL");
  int rc;
  if ((rc = launch_ens_reset(e->queues, Q_COUNT + 2, e->stream)) ||
      (rc = field_phase(e, Q_LOAD, ENS_LOAD, e->stream)))
    return rc;
  VPIC_HIP(hipStreamSynchronize(e->stream));
  return VPIC_OK;
}

extern "C" int vpic_hip_ensemble_step(vpic_ensemble_t *e, int32_t nstep,
                                      int32_t use_graph) {
  if (!e || nstep < 0)
    return vpic_set_error(VPIC_EINVAL, "ensemble_step: bad arguments");
  int rc;
  int *tcnt = ens_tiled_counters(e->tiled);
  if (tcnt)
    VPIC_HIP(hipMemsetAsync(tcnt, 0, 4 * sizeof(int), e->stream));
  if ((rc = run_steps(e, nstep, use_graph))) return rc;
  VPIC_HIP(hipStreamSynchronize(e->stream));
  int bad[2] = {0, 0};
  VPIC_HIP(hipMemcpy(bad, e->queues + Q_COUNT, sizeof bad,
                     hipMemcpyDeviceToHost));
  tcnt = ens_tiled_counters(e->tiled);  // (created by the first sorting step)
  if (tcnt) {
    int tc[4] = {0, 0, 0, 0};
    VPIC_HIP(hipMemcpy(tc, tcnt, sizeof tc, hipMemcpyDeviceToHost));
    bad[0] += tc[3];
    bad[1] += tc[1];
  }
  if (bad[0] || bad[1])
    return vpic_set_error(VPIC_EINVAL,
                          "ensemble_step: %d particles with a voxel off their "
                          "box, %d movers (the boxes are periodic)",
                          bad[0], bad[1]);
  return VPIC_OK;
}

extern "C" int vpic_hip_ensemble_shape(vpic_ensemble_t *e, int32_t push_bs,
                                       int32_t push_chunk, int32_t field_bs,
                                       int32_t field_chunk,
                                       int32_t blocks_per_cu) {
  if (!e) return vpic_set_error(VPIC_EINVAL, "ensemble is NULL");
  if (push_bs < 64 || push_bs > 256 || (push_bs & 63) || field_bs < 64 ||
      field_bs > 512 || (field_bs & 63) || push_chunk < 64 ||
      field_chunk < 64 || blocks_per_cu < 1 || blocks_per_cu > 64)
    return vpic_set_error(VPIC_EINVAL, "ensemble_shape: out of range");
  e->shape = {push_bs, push_chunk, field_bs, field_chunk, blocks_per_cu};
  return build_items(e);
}

extern "C" int vpic_hip_ensemble_tile_window(vpic_ensemble_t *e, int32_t tx,
                                             int32_t ty, int32_t tz,
                                             int32_t halo) {
  if (!e || tx < 0 || ty < 0 || tz < 0 || (tx == 0) != (ty == 0) ||
      (tx == 0) != (tz == 0) || halo < 1 || halo > ENS_MAX_HALO)
    return vpic_set_error(VPIC_EINVAL, "ensemble_tile: bad arguments");
  if (tx == e->tile.tx && ty == e->tile.ty && tz == e->tile.tz &&
      halo == e->tile.halo)
    return VPIC_OK;
  VPIC_HIP(hipStreamSynchronize(e->stream));
  if (e->tiled) ens_tiled_destroy(e->tiled);
  e->tiled = nullptr;
  e->sorted_once = false;
  e->tile = {tx, ty, tz, halo};
  drop_graph(e);
  return VPIC_OK;
}

extern "C" int vpic_hip_ensemble_tile(vpic_ensemble_t *e, int32_t tx,
                                      int32_t ty, int32_t tz) {
  return vpic_hip_ensemble_tile_window(e, tx, ty, tz, 1);
}

extern "C" int vpic_hip_ensemble_get_tile(const vpic_ensemble_t *e,
                                          int32_t out[3]) {
  if (!e || !out) return vpic_set_error(VPIC_EINVAL, "bad arguments");
  out[0] = e->tile.tx;
  out[1] = e->tile.ty;
  out[2] = e->tile.tz;
  return VPIC_OK;
}

extern "C" int vpic_hip_ensemble_get_tile_window(const vpic_ensemble_t *e,
                                                 int32_t out[4]) {
  if (!e || !out) return vpic_set_error(VPIC_EINVAL, "bad arguments");
  out[0] = e->tile.tx;
  out[1] = e->tile.ty;
  out[2] = e->tile.tz;
  out[3] = e->tile.halo;
  return VPIC_OK;
}

extern "C" int vpic_hip_ensemble_sort_interval(vpic_ensemble_t *e,
                                               int32_t interval) {
  if (!e || interval < 0)
    return vpic_set_error(VPIC_EINVAL, "ensemble_sort_interval: bad arguments");
  if (interval != e->sort_interval) {
    e->sort_interval = interval;
    drop_graph(e);
  }
  return VPIC_OK;
}

extern "C" int vpic_hip_ensemble_get_shape(const vpic_ensemble_t *e,
                                           int32_t out[5]) {
  if (!e || !out) return vpic_set_error(VPIC_EINVAL, "bad arguments");
  out[0] = e->shape.push_bs;
  out[1] = e->shape.push_chunk;
  out[2] = e->shape.field_bs;
  out[3] = e->shape.field_chunk;
  out[4] = e->shape.blocks_per_cu;
  return VPIC_OK;
}

// The tuner: every candidate is timed over `trial_steps` graph replays
// (hipEvents on the ensemble's stream) from the same saved state, restored
// after each; first the tile of the tiled push (with the current launch
// shape), then the launch shape (with the best tile).  With sorting on,
// each trial starts from a sort of the saved state (untimed; it writes the
// partitions the tiled push reads) and runs without further sorts.
extern "C" int vpic_hip_ensemble_tune(vpic_ensemble_t *e, int32_t trial_steps,
                                      double *host_best_ms) {
  if (!e || trial_steps < 1)
    return vpic_set_error(VPIC_EINVAL, "ensemble_tune: bad arguments");
  static const Shape cands[] = {
      {256, 4096, 256, 2048, 4},  {256, 8192, 256, 4096, 4},
      {256, 2048, 256, 1024, 8},  {128, 2048, 256, 2048, 8},
      {256, 16384, 256, 8192, 2}, {256, 4096, 512, 4096, 4},
      {64, 1024, 128, 1024, 16},  {256, 32768, 256, 16384, 1},
      {128, 4096, 128, 2048, 8},  {256, 8192, 512, 8192, 2},
  };
  // (a box one cell thick along an axis has no window halo along it, so
  // the flat shapes' windows are the tile and a one-cell ring in-plane)
  // (and wider halos, which on flat boxes keep the deposits of particles
  // that drifted further since the sort in LDS; a shape whose window does
  // not fit some box is skipped)
  static const TileShape tiles[] = {
      {32, 16, 1, 1}, {32, 16, 1, 2}, {32, 16, 1, 3}, {64, 8, 1, 1},
      {64, 8, 1, 2},  {16, 16, 1, 1}, {16, 16, 1, 2}, {16, 16, 1, 3},
      {32, 8, 1, 1},  {32, 8, 1, 2},  {16, 8, 1, 1},  {8, 8, 1, 1},
      {16, 8, 4, 1},  {7, 8, 7, 1}};
  const int ncand = sizeof(cands) / sizeof(cands[0]);
  const int ntile = sizeof(tiles) / sizeof(tiles[0]);
  void *saveP = nullptr, *saveF = nullptr, *saveFI = nullptr;
  int rc;
  const size_t pbytes = sizeof(vpic_particle_t) * e->np_total;
  const size_t fbytes = sizeof(vpic_field_t) * e->nv_total;
  const size_t ibytes = sizeof(vpic_interpolator_t) * e->nv_total;
  if ((rc = ens_alloc(&saveP, pbytes)) || (rc = ens_alloc(&saveF, fbytes)) ||
      (rc = ens_alloc(&saveFI, ibytes))) {
    if (saveP) (void)hipFree(saveP);
    if (saveF) (void)hipFree(saveF);
    return rc;
  }
  hipStream_t s = e->stream;
  hipEvent_t t0, t1;
  (void)hipEventCreate(&t0);
  (void)hipEventCreate(&t1);
  const Shape orig = e->shape;
  const TileShape orig_tile = e->tile;
  const int64_t steps0 = e->steps;
  const int interval0 = e->sort_interval;
  e->sort_interval = 0;  // no sorts inside a trial
  (void)hipMemcpyAsync(saveP, e->P, pbytes, hipMemcpyDeviceToDevice, s);
  (void)hipMemcpyAsync(saveF, e->F, fbytes, hipMemcpyDeviceToDevice, s);
  (void)hipMemcpyAsync(saveFI, e->FI, ibytes, hipMemcpyDeviceToDevice, s);
  auto restore = [&]() {
    (void)hipMemcpyAsync(e->P, saveP, pbytes, hipMemcpyDeviceToDevice, s);
    (void)hipMemcpyAsync(e->F, saveF, fbytes, hipMemcpyDeviceToDevice, s);
    (void)hipMemcpyAsync(e->FI, saveFI, ibytes, hipMemcpyDeviceToDevice, s);
  };
  auto set_tile = [&](const TileShape &t) -> int {
    if (t.tx == e->tile.tx && t.ty == e->tile.ty && t.tz == e->tile.tz &&
        t.halo == e->tile.halo && (e->tiled || t.tx <= 0))
      return VPIC_OK;
    (void)hipStreamSynchronize(s);
    if (e->tiled) ens_tiled_destroy(e->tiled);
    e->tiled = nullptr;
    e->sorted_once = false;
    e->tile = t;
    return ensure_tiled(e);
  };
  // one trial of the current shape and tile: ms per step
  auto trial = [&](float &ms) -> int {
    int r;
    if (interval0 > 0) {
      if ((r = ensure_sort_buffers(e)) || (r = enqueue_sort(e, s))) return r;
      e->sorted_once = true;
    }
    if ((r = build_items(e)) || (r = capture(e))) return r;
    if ((r = run_steps(e, 1, 1))) return r;  // first-launch costs
    (void)hipEventRecord(t0, s);
    if ((r = run_steps(e, trial_steps, 1))) return r;
    (void)hipEventRecord(t1, s);
    (void)hipEventSynchronize(t1);
    (void)hipEventElapsedTime(&ms, t0, t1);
    ms /= trial_steps;
    restore();
    return VPIC_OK;
  };
  Shape best = orig;
  TileShape best_tile = orig_tile;
  float best_ms = 1e30f;
  rc = VPIC_OK;
  // the tiles (the tiled push needs the sort's partitions)
  for (int c = 0; c < (interval0 > 0 ? ntile : 0) && rc == VPIC_OK; c++) {
    float ms = 0;
    if (set_tile(tiles[c]) != VPIC_OK) continue;  // does not fit a box
    if ((rc = trial(ms))) break;
    if (ms < best_ms) {
      best_ms = ms;
      best_tile = tiles[c];
    }
  }
  if (rc == VPIC_OK) rc = set_tile(best_tile);
  for (int c = 0; c < ncand && rc == VPIC_OK; c++) {
    float ms = 0;
    e->shape = cands[c];
    if ((rc = trial(ms))) break;
    if (ms < best_ms) {
      best_ms = ms;
      best = cands[c];
    }
  }
  (void)hipStreamSynchronize(s);
  e->steps = steps0;  // the trials' steps were undone with the state
  e->sort_interval = interval0;
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  (void)hipFree(saveP);
  (void)hipFree(saveF);
  (void)hipFree(saveFI);
  if (rc) {
    e->shape = orig;
    (void)set_tile(orig_tile);
    e->sorted_once = false;
    (void)build_items(e);
    return rc;
  }
  e->shape = best;
  // the partitions now describe the last trial's sort of the saved state,
  // not the restored order: any sort's partition still gives every particle
  // to exactly one tile, and the next sort refreshes it
  e->sorted_once = interval0 > 0 && e->tiled != nullptr;
  if ((rc = build_items(e))) return rc;  // drops the trial graphs
  if (host_best_ms) *host_best_ms = best_ms;
  return VPIC_OK;
}
