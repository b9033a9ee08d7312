// C-ABI: workspace lifetime, eager warm-up, graph capture and replay.
// Declarations and reference citations: include/peel355.h.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/peel355.h"
#include "workspace.hpp"

namespace peel {
int subdivide_blocks_per_cu();
}

using namespace peel;

struct PeelWorkspace {
  Ws w;
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  bool warmed = false;
  int device = 0;
  std::vector<void *> allocations;
  int64_t bytes = 0;  // device memory held by the pools
  int cus = 1;        // compute units of the device
  int max_per_cu = 1; // resident k_subdivide blocks per CU (occupancy)
};

static thread_local std::string g_error;

static int fail(int code, const std::string &message) {
  g_error = message;
  return code;
}

#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      return fail(-(int)_e, std::string(#expr) + ": " + hipGetErrorString(_e));        \
  } while (0)

template <class T>
static int alloc(PeelWorkspace *ws, T *&ptr, long count) {
  if (count < 1) count = 1;
  void *p = nullptr;
  hipError_t e = hipMalloc(&p, sizeof(T) * (size_t)count);
  if (e != hipSuccess)
    return fail(-(int)e, "hipMalloc of " + std::to_string(sizeof(T) * (size_t)count) +
                             " bytes: " + hipGetErrorString(e));
  e = hipMemset(p, 0, sizeof(T) * (size_t)count);
  if (e != hipSuccess) return fail(-(int)e, std::string("hipMemset: ") + hipGetErrorString(e));
  ws->allocations.push_back(p);
  ws->bytes += (int64_t)(sizeof(T) * (size_t)count);
  ptr = (T *)p;
  return 0;
}

#define ALLOC(ptr, n)                          \
  do {                                         \
    int _r = alloc(ws, ptr, (long)(n));        \
    if (_r) {                                  \
      peel_destroy(ws);                        \
      return _r;                               \
    }                                          \
  } while (0)

static int alloc_hierarchy(PeelWorkspace *ws, Hierarchy &h, int leaves) {
  h.size = pow2_at_least(leaves < 1 ? 1 : leaves);
  int r;
  if ((r = alloc(ws, h.keys, h.size))) return r;
  if ((r = alloc(ws, h.alt_keys, h.size))) return r;
  if ((r = alloc(ws, h.ids, h.size))) return r;
  if ((r = alloc(ws, h.alt_ids, h.size))) return r;
  if ((r = alloc(ws, h.nodes, 2L * h.size))) return r;
  if ((r = alloc(ws, h.leaves, 1))) return r;
  return 0;
}

extern "C" {

const char *peel_last_error(void) { return g_error.c_str(); }

const char *peel_version(void) { return "peel355 0.1 gfx950"; }

int peel_destroy(PeelWorkspace *ws) {
  if (!ws) return 0;
  if (ws->exec) hipGraphExecDestroy(ws->exec);
  if (ws->graph) hipGraphDestroy(ws->graph);
  if (ws->stream) hipStreamSynchronize(ws->stream);
  for (void *p : ws->allocations) hipFree(p);
  if (ws->stream) hipStreamDestroy(ws->stream);
  delete ws;
  return 0;
}

int peel_create(const PeelDesc *d, PeelWorkspace **out) {
  if (!d || !out) return fail(1, "null argument");
  *out = nullptr;
  if (d->num_vertices < 0 || d->num_faces < 0 || d->num_vertices > (1L << 30) ||
      d->num_faces > (1L << 27))
    return fail(2, "vertex or face count out of range");
  if ((d->num_vertices && !d->vertices) || (d->num_faces && !d->faces) || !d->camera || !d->settings)
    return fail(3, "null device array");
  if (d->frontier < 1 || d->regions < 1 || d->hits < 1 || d->vertices_cap < 1 || d->triangles < 1 ||
      d->rounds < 1 || d->layer_capacity < 1 || d->edge_vertices < 1)
    return fail(4, "Capacities must be positive");
  if (d->edge_vertices > 128) return fail(5, "edge_vertices must be <= 128");
  if (d->layers < 0 || d->layers > d->layer_capacity)
    return fail(6, "layers must be zero (all), or fit in the layer capacity");
  if (d->route != PEEL_ROUTE_EVENTS && d->route != PEEL_ROUTE_LAYERED) return fail(7, "unknown route");
  if (d->route == PEEL_ROUTE_LAYERED && (d->workers < 1 || d->pieces < 1))
    return fail(4, "Capacities must be positive");
  PeelWorkspace *ws = new PeelWorkspace();
  HIP_TRY(hipGetDevice(&ws->device));
  Ws &w = ws->w;
  std::memset(&w, 0, sizeof(Ws));
  w.vertices = d->vertices;
  w.faces = d->faces;
  w.camera = d->camera;
  w.settings = d->settings;
  w.nv = (int)d->num_vertices;
  w.nf = (int)d->num_faces;
  w.cap = Caps{d->frontier, d->regions, d->hits, d->vertices_cap, d->triangles, d->rounds,
               d->layer_capacity, d->edge_vertices};
  w.limit = d->layers ? d->layers : d->layer_capacity;
  w.all_layers = d->layers == 0;
  w.share_regions = (d->flags & PEEL_SHARE_REGIONS) != 0;
  w.propagate_ranks = (d->flags & PEEL_PROPAGATE_RANKS) != 0;
  w.merge_regions = (d->flags & PEEL_MERGE_REGIONS) != 0;
  w.route = d->route;
  const bool layered = d->route == PEEL_ROUTE_LAYERED;
  w.workers = layered ? d->workers : 1;
  w.pieces = layered ? d->pieces : 1;
  int nf1 = w.nf > 0 ? w.nf : 1, nv1 = w.nv > 0 ? w.nv : 1;
  ALLOC(w.status, 16);
  ALLOC(w.metrics, M_COUNT);
  ALLOC(w.control, C_COUNT);
  ALLOC(w.inverse, 16);
  ALLOC(w.canonical, nv1);
  w.vhash_size = pow2_at_least(2 * nv1);
  ALLOC(w.vhash, w.vhash_size);
  w.ehash_size = pow2_at_least(6 * nf1);
  ALLOC(w.ekeys, w.ehash_size);
  ALLOC(w.ecount, w.ehash_size);
  ALLOC(w.efirst, w.ehash_size);
  ALLOC(w.esecond, w.ehash_size);
  ALLOC(w.masks, 3L * nf1);
  ALLOC(w.src_xy, (long)SRCV * nf1);
  ALLOC(w.src_n, nf1);
  ALLOC(w.scores, nf1);
  ALLOC(w.facing, nf1);
  ALLOC(w.clipped, nf1);
  ALLOC(w.src_boxes, nf1);
  w.nseg = 8 * nf1;
  ALLOC(w.seg_a, w.nseg);
  ALLOC(w.seg_b, w.nseg);
  ALLOC(w.seg_near, 2L * w.nseg);
  ALLOC(w.seg_rep, w.nseg);
  w.shash_size = pow2_at_least(2 * w.nseg);
  ALLOC(w.shash, w.shash_size);
  ALLOC(w.seg_boxes, w.nseg);
  ALLOC(w.seg_flag, w.nseg);
  ALLOC(w.seg_off, w.nseg);
  ALLOC(w.cseg_a, w.nseg);
  ALLOC(w.cseg_b, w.nseg);
  ALLOC(w.cnear, 2L * w.nseg);
  int r;
  if ((r = alloc_hierarchy(ws, w.source_tree, nf1)) || (r = alloc_hierarchy(ws, w.contour_tree, w.nseg))) {
    peel_destroy(ws);
    return r;
  }
  ALLOC(w.leaf_xy, 8L * w.source_tree.size);
  ALLOC(w.leaf_n, w.source_tree.size);
  ALLOC(w.leaf_score, w.source_tree.size);
  w.filter32 = (d->flags & PEEL_FILTER_FP32) != 0;
  ALLOC(w.leaf_xy32, w.filter32 ? 8L * w.source_tree.size : 1);
  ALLOC(w.leaf_tol, w.filter32 ? w.source_tree.size : 1);
  long F = layered ? 1 : w.cap.frontier, R = w.cap.regions;
  // layered route: ping-pong residual pools sized by the frontier capacity,
  // per-worker piece pools, per-layer statistics
  long RES = layered ? w.cap.frontier : 1, NF = layered ? nf1 : 1;
  long POOL = layered ? 3L * (w.pieces + 3) * w.workers : 1;
  ALLOC(w.ranks, NF);
  ALLOC(w.contour_boxes, NF);
  for (int par = 0; par < 2; ++par) {
    ALLOC(w.res_xy[par], RES * MAXV);
    ALLOC(w.res_n[par], RES);
    ALLOC(w.res_src[par], RES);
    ALLOC(w.res_run[par], RES);
    ALLOC(w.res_link[par], RES);
    ALLOC(w.res_box[par], RES);
    ALLOC(w.heads[par], NF);
  }
  ALLOC(w.pool_xy, POOL * MAXV);
  ALLOC(w.pool_n, POOL);
  ALLOC(w.stats, 3L * w.cap.layers);
  ALLOC(w.fr_xy, F * MAXV);
  ALLOC(w.fr_n, F);
  ALLOC(w.fr_src, F);
  ALLOC(w.fr_node, F);
  ALLOC(w.fr_gen, F);
  ALLOC(w.fr_ready, F);
  ALLOC(w.out_xy, R * MAXV);
  ALLOC(w.out_n, R);
  ALLOC(w.out_src, R);
  ALLOC(w.out_layer, R);
  ALLOC(w.samples, R);
  ALLOC(w.reps, R);
  w.qhash_size = pow2_at_least(2 * w.cap.regions);
  ALLOC(w.qhash, w.qhash_size);
  ALLOC(w.groups, R);
  ALLOC(w.group_counts, R);
  ALLOC(w.group_offsets, R);
  ALLOC(w.group_fill, R);
  ALLOC(w.ray_keys, R);
  ALLOC(w.ray_keys_alt, R);
  ALLOC(w.ray_ids, R);
  ALLOC(w.ray_ids_alt, R);
  ALLOC(w.medium, R);
  ALLOC(w.heavy, R);
  ALLOC(w.beam_box, 4L * BEAM_CELLS);
  ALLOC(w.beam_leader, BEAM_CELLS);
  ALLOC(w.beam_valid, BEAM_CELLS);
  ALLOC(w.beam_jobs, BEAM_CELLS);
  long H = w.cap.hits;
  ALLOC(w.hit_depth, H);
  ALLOC(w.hit_src, H);
  ALLOC(w.hit_group, H);
  ALLOC(w.hkeys, H);
  ALLOC(w.hkeys_alt, H);
  ALLOC(w.hids, H);
  ALLOC(w.hids_alt, H);
  long BV = w.cap.vertices, T = w.cap.triangles;
  ALLOC(w.owners, BV);
  ALLOC(w.starts, BV);
  ALLOC(w.ends, BV);
  ALLOC(w.partners, BV);
  ALLOC(w.parents, R);
  ALLOC(w.choices, R);
  long MR = w.merge_regions ? R : 1;
  ALLOC(w.merge_xy, MR * MAXV);
  ALLOC(w.merge_n, MR);
  ALLOC(w.merge_src, MR);
  ALLOC(w.merge_layer, MR);
  w.ehash2_size = pow2_at_least(2 * (int)std::min<long>(BV, 1L << 29));
  ALLOC(w.eslots, w.ehash2_size);
  ALLOC(w.vcounts, R);
  ALLOC(w.voffsets, R);
  ALLOC(w.bverts, BV);
  ALLOC(w.blayers, BV);
  ALLOC(w.tasks, BV);
  ALLOC(w.crowded, BV);
  w.phash_size = pow2_at_least(2 * (int)std::min<long>(BV, 1L << 29));
  ALLOC(w.phash, w.phash_size);
  w.cell_capacity = 4 * BV;
  ALLOC(w.cell_heads, w.cell_capacity);
  ALLOC(w.cell_next, BV);
  ALLOC(w.tri_uv, T * 6);
  ALLOC(w.tri_world, T * 9);
  ALLOC(w.tri_src, T);
  ALLOC(w.tri_layer, T);
  long sort_cap = H;
  sort_cap = std::max<long>(sort_cap, w.source_tree.size);
  sort_cap = std::max<long>(sort_cap, w.contour_tree.size);
  sort_cap = std::max<long>(sort_cap, R);
  w.sort.max_blocks = (int)((sort_cap + RS_TILE - 1) / RS_TILE);
  ALLOC(w.sort.counts, RS_MAX_SWEEPS * 256);
  ALLOC(w.sort.tickets, RS_MAX_SWEEPS);
  ALLOC(w.sort.tiles, (long)RS_MAX_SWEEPS * w.sort.max_blocks * 256);
  ALLOC(w.sort.flags, RS_MAX_SWEEPS);
  w.sort.status = w.status;
  long scan_cap = std::max<long>(std::max<long>(R, BV), w.nseg);
  w.scan.max_blocks = (int)((scan_cap + 4095) / 4096) + 1;
  ALLOC(w.scan.block_sums, w.scan.max_blocks);
  int cus = 0;
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, ws->device));
  int per_cu = subdivide_blocks_per_cu();
  ws->cus = std::max(1, cus);
  ws->max_per_cu = std::max(1, per_cu);
  // One 256-lane block per CU: enough lanes to hide traversal latency, few
  // enough idle pollers that the queue words do not become an L2 hot spot.
  w.persistent_blocks = ws->cus;
  if (const char *env = std::getenv("PEEL355_PERSISTENT_BLOCKS")) {
    int v = std::atoi(env);
    if (v > 0) w.persistent_blocks = std::min(v, ws->cus * ws->max_per_cu);
  }
  w.ray_lanes_target = 524288;
  w.heavy_lanes = 65536;
  w.long_edge_cells = 48;
  hipError_t e = hipStreamCreateWithFlags(&ws->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    peel_destroy(ws);
    return fail(-(int)e, std::string("hipStreamCreate: ") + hipGetErrorString(e));
  }
  e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    peel_destroy(ws);
    return fail(-(int)e, std::string("hipDeviceSynchronize: ") + hipGetErrorString(e));
  }
  *out = ws;
  return 0;
}

static void enqueue_all(const Ws &w, hipStream_t s, Timeline *t) {
  enqueue_prepare(w, s, t);
  if (w.route == PEEL_ROUTE_LAYERED) {
    enqueue_layered(w, s, t);
  } else {
    enqueue_subdivide(w, s, t);
    enqueue_rank(w, s, t);
  }
  enqueue_mesh(w, s, t);
}

int peel_enqueue(PeelWorkspace *ws) {
  if (!ws) return fail(1, "null workspace");
  enqueue_all(ws->w, ws->stream, nullptr);
  HIP_TRY(hipGetLastError());
  return 0;
}

int peel_profile(PeelWorkspace *ws, int32_t *count, float *ms, char *names, int32_t names_len) {
  if (!ws || !count) return fail(1, "null argument");
  Timeline t;
  for (int i = 0; i < Timeline::MAX; ++i) HIP_TRY(hipEventCreate(&t.events[i]));
  const Ws &w = ws->w;
  mark(&t, ws->stream, "begin");
  enqueue_all(w, ws->stream, &t);
  hipError_t e = hipStreamSynchronize(ws->stream);
  std::string joined;
  int n = 0;
  for (int i = 1; i < t.count && e == hipSuccess; ++i) {
    float v = 0.f;
    e = hipEventElapsedTime(&v, t.events[i - 1], t.events[i]);
    if (ms && n < *count) ms[n] = v;
    ++n;
    if (!joined.empty()) joined += ",";
    joined += t.names[i];
  }
  for (int i = 0; i < Timeline::MAX; ++i) (void)hipEventDestroy(t.events[i]);
  if (e != hipSuccess) return fail(-(int)e, std::string("profile: ") + hipGetErrorString(e));
  *count = n;
  if (names && names_len > 0) {
    std::strncpy(names, joined.c_str(), names_len - 1);
    names[names_len - 1] = 0;
  }
  return 0;
}

int peel_warmup(PeelWorkspace *ws) {
  if (!ws) return fail(1, "null workspace");
  hipStreamCaptureStatus st;
  HIP_TRY(hipStreamIsCapturing(ws->stream, &st));
  if (st != hipStreamCaptureStatusNone) return fail(8, "Warm-up must run before graph capture");
  int r = peel_enqueue(ws);
  if (r) return r;
  HIP_TRY(hipStreamSynchronize(ws->stream));
  ws->warmed = true;
  return 0;
}

int peel_capture(PeelWorkspace *ws) {
  if (!ws) return fail(1, "null workspace");
  if (!ws->warmed) {
    int r = peel_warmup(ws);
    if (r) return r;
  }
  if (ws->exec) {
    hipGraphExecDestroy(ws->exec);
    ws->exec = nullptr;
  }
  if (ws->graph) {
    hipGraphDestroy(ws->graph);
    ws->graph = nullptr;
  }
  HIP_TRY(hipStreamBeginCapture(ws->stream, hipStreamCaptureModeThreadLocal));
  int r = peel_enqueue(ws);
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(ws->stream, &g);
  if (r) {
    if (g) hipGraphDestroy(g);
    return r;
  }
  if (e != hipSuccess) return fail(-(int)e, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
  ws->graph = g;
  HIP_TRY(hipGraphInstantiate(&ws->exec, ws->graph, nullptr, nullptr, 0));
  return 0;
}

int peel_get_tuning(PeelWorkspace *ws, PeelTuning *t) {
  if (!ws || !t) return fail(1, "null argument");
  const Ws &w = ws->w;
  t->persistent_blocks_per_cu = std::max(1, w.persistent_blocks / ws->cus);
  t->ray_lanes_target = w.ray_lanes_target;
  t->heavy_lanes = w.heavy_lanes;
  t->long_edge_cells = w.long_edge_cells;
  t->max_blocks_per_cu = ws->max_per_cu;
  return 0;
}

int peel_set_tuning(PeelWorkspace *ws, const PeelTuning *t) {
  if (!ws || !t) return fail(1, "null argument");
  if (t->persistent_blocks_per_cu < 1 || t->persistent_blocks_per_cu > ws->max_per_cu)
    return fail(2, "persistent_blocks_per_cu must be in [1, " + std::to_string(ws->max_per_cu) + "]");
  if (t->ray_lanes_target < 64) return fail(2, "ray_lanes_target must be at least 64");
  if (t->heavy_lanes < 1024) return fail(2, "heavy_lanes must be at least 1024");
  if (t->long_edge_cells < 1) return fail(2, "long_edge_cells must be positive");
  Ws &w = ws->w;
  w.persistent_blocks = ws->cus * t->persistent_blocks_per_cu;
  w.ray_lanes_target = t->ray_lanes_target;
  w.heavy_lanes = t->heavy_lanes;
  w.long_edge_cells = t->long_edge_cells;
  // the captured graph holds the old launch shapes and kernel arguments
  if (ws->exec) {
    hipGraphExecDestroy(ws->exec);
    ws->exec = nullptr;
  }
  if (ws->graph) {
    hipGraphDestroy(ws->graph);
    ws->graph = nullptr;
  }
  return 0;
}

int peel_launch(PeelWorkspace *ws) {
  if (!ws) return fail(1, "null workspace");
  if (!ws->exec) {
    int r = peel_capture(ws);
    if (r) return r;
  }
  HIP_TRY(hipGraphLaunch(ws->exec, ws->stream));
  return 0;
}

int peel_synchronize(PeelWorkspace *ws) {
  if (!ws) return fail(1, "null workspace");
  HIP_TRY(hipStreamSynchronize(ws->stream));
  return 0;
}

int peel_read_status(PeelWorkspace *ws, int32_t *status, int64_t *metrics) {
  if (!ws) return fail(1, "null workspace");
  HIP_TRY(hipStreamSynchronize(ws->stream));
  if (status) HIP_TRY(hipMemcpy(status, ws->w.status, 16 * sizeof(int32_t), hipMemcpyDeviceToHost));
  if (metrics)
    HIP_TRY(hipMemcpy(metrics, ws->w.metrics, M_COUNT * sizeof(int64_t), hipMemcpyDeviceToHost));
  return 0;
}

int peel_get_buffers(PeelWorkspace *ws, PeelBuffers *b) {
  if (!ws || !b) return fail(1, "null argument");
  const Ws &w = ws->w;
  b->output_xy = (double *)w.out_xy;
  b->output_n = w.out_n;
  b->output_sources = w.out_src;
  b->output_layers = w.out_layer;
  b->regions = w.cap.regions;
  b->max_vertices = MAXV;
  b->triangles_uv = w.tri_uv;
  b->triangles_world = w.tri_world;
  b->triangle_sources = w.tri_src;
  b->triangle_layers = w.tri_layer;
  b->triangles = w.cap.triangles;
  b->status = w.status;
  b->metrics = (int64_t *)w.metrics;
  b->metrics_count = M_COUNT;
  b->canonical = w.canonical;
  b->stream = (void *)ws->stream;
  b->stats = (int64_t *)w.stats;
  b->stats_rows = w.cap.layers;
  b->device_bytes = ws->bytes;
  return 0;
}

// ---------------------------------------------------------------------------
// Batches of workspaces in one graph.
struct PeelBatch {
  std::vector<PeelWorkspace *> members;
  std::vector<hipEvent_t> events;  // fork (one) and joins (one per member)
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
};

int peel_batch_destroy(PeelBatch *b) {
  if (!b) return 0;
  if (b->exec) hipGraphExecDestroy(b->exec);
  if (b->graph) hipGraphDestroy(b->graph);
  for (hipEvent_t e : b->events) hipEventDestroy(e);
  delete b;
  return 0;
}

int peel_batch_create(PeelWorkspace *const *list, int32_t count, PeelBatch **out) {
  if (!list || !out || count < 1) return fail(1, "null argument or empty batch");
  *out = nullptr;
  for (int i = 0; i < count; ++i) {
    if (!list[i]) return fail(1, "null workspace in batch");
    for (int j = 0; j < i; ++j)
      if (list[j] == list[i]) return fail(2, "a workspace appears twice in the batch");
    if (list[i]->device != list[0]->device) return fail(2, "batch workspaces must share one device");
  }
  for (int i = 0; i < count; ++i) {
    if (!list[i]->warmed) {
      int r = peel_warmup(list[i]);
      if (r) return r;
    }
  }
  PeelBatch *b = new PeelBatch();
  b->members.assign(list, list + count);
  b->events.resize(count + 1, nullptr);
  for (auto &e : b->events) {
    hipError_t err = hipEventCreateWithFlags(&e, hipEventDisableTiming);
    if (err != hipSuccess) {
      peel_batch_destroy(b);
      return fail(-(int)err, std::string("hipEventCreate: ") + hipGetErrorString(err));
    }
  }
  hipStream_t root = list[0]->stream;
  hipError_t err = hipStreamBeginCapture(root, hipStreamCaptureModeThreadLocal);
  if (err != hipSuccess) {
    peel_batch_destroy(b);
    return fail(-(int)err, std::string("hipStreamBeginCapture: ") + hipGetErrorString(err));
  }
  // fork: every other member stream waits on the root, joining the capture
  err = hipEventRecord(b->events[0], root);
  for (int i = 1; i < count && err == hipSuccess; ++i) err = hipStreamWaitEvent(list[i]->stream, b->events[0], 0);
  for (int i = 0; i < count && err == hipSuccess; ++i) {
    enqueue_all(list[i]->w, list[i]->stream, nullptr);
    err = hipGetLastError();
  }
  // join: the root waits on every member stream
  for (int i = 1; i < count && err == hipSuccess; ++i) {
    err = hipEventRecord(b->events[i], list[i]->stream);
    if (err == hipSuccess) err = hipStreamWaitEvent(root, b->events[i], 0);
  }
  hipGraph_t g = nullptr;
  hipError_t end = hipStreamEndCapture(root, &g);
  if (err == hipSuccess) err = end;
  if (err != hipSuccess) {
    if (g) hipGraphDestroy(g);
    peel_batch_destroy(b);
    return fail(-(int)err, std::string("batch capture: ") + hipGetErrorString(err));
  }
  b->graph = g;
  err = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
  if (err != hipSuccess) {
    peel_batch_destroy(b);
    return fail(-(int)err, std::string("hipGraphInstantiate: ") + hipGetErrorString(err));
  }
  *out = b;
  return 0;
}

int peel_batch_launch(PeelBatch *b) {
  if (!b) return fail(1, "null batch");
  HIP_TRY(hipGraphLaunch(b->exec, b->members[0]->stream));
  return 0;
}

int peel_batch_synchronize(PeelBatch *b) {
  if (!b) return fail(1, "null batch");
  HIP_TRY(hipStreamSynchronize(b->members[0]->stream));
  return 0;
}

}  // extern "C"
