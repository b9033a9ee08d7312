// C-ABI: workspace lifetime, eager warm-up, graph capture and replay.
// Declarations and reference citations: include/peel355.h.
#include <cstdio>
#include <cstdlib>
#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "layout.hpp"

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
  void *block = nullptr;  // device memory of every pool (one allocation)
  int64_t bytes = 0;      // its size
  int cus = 1;        // compute units of the device
  int max_per_cu = 1; // resident k_subdivide blocks per CU (occupancy)
};

static thread_local std::string g_error;

int peel::fail(int code, const std::string &message) {
  g_error = message;
  return code;
}


// Power of two at or above n, or 0 when it would exceed 2^30 (the hash tables
// and hierarchies index with int).
static int pow2_checked(long n) {
  long s = 1;
  while (s < n) s <<= 1;
  return s > (1L << 30) ? 0 : (int)s;
}

// Validate the descriptor and set every scalar of the workspace layout.
// Returns 0 or a positive argument error (message in g_error).
int peel::plan(Ws &w, const PeelDesc *d) {
  if (d->num_vertices < 0 || d->num_faces < 0 || d->num_vertices > (1L << 30) ||
      d->num_faces > (1L << 26))
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
  w.filter32 = (d->flags & PEEL_FILTER_FP32) != 0;
  w.route = d->route;
  const bool layered = d->route == PEEL_ROUTE_LAYERED;
  w.workers = layered ? d->workers : 1;
  w.pieces = layered ? d->pieces : 1;
  const long nf1 = w.nf > 0 ? w.nf : 1, nv1 = w.nv > 0 ? w.nv : 1;
  const long BV = w.cap.vertices;
  w.nseg = (int)(8 * nf1);
  w.vhash_size = pow2_checked(2 * nv1);
  w.ehash_size = pow2_checked(6 * nf1);
  w.shash_size = pow2_checked(2L * w.nseg);
  w.qhash_size = pow2_checked(2L * w.cap.regions);
  w.ehash2_size = pow2_checked(2 * std::min<long>(BV, 1L << 29));
  w.phash_size = pow2_checked(2 * std::min<long>(BV, 1L << 29));
  w.source_tree.size = pow2_checked(nf1);
  w.contour_tree.size = pow2_checked(w.nseg);
  for (int s : {w.vhash_size, w.ehash_size, w.shash_size, w.qhash_size, w.ehash2_size, w.phash_size,
                w.source_tree.size, w.contour_tree.size})
    if (s == 0) return fail(2, "mesh or capacities too large: a hash table or hierarchy would exceed 2^30 slots");
  w.cell_capacity = std::min<long>(4 * BV, 1L << 30);
  auto bits_for = [](long n) {
    int b = 1;
    while ((1L << b) < n) ++b;
    return b;
  };
  w.res_parent_bits = bits_for(layered ? w.cap.frontier : 1);
  w.res_key_bits = w.res_parent_bits + bits_for(nf1);
  long sort_cap = std::max<long>({(long)w.cap.hits, (long)w.source_tree.size, (long)w.contour_tree.size,
                                  (long)w.cap.regions, layered ? (long)w.cap.frontier : 1L});
  w.sort.max_blocks = (int)((sort_cap + RS_TILE - 1) / RS_TILE);
  long scan_cap = std::max<long>({(long)w.cap.regions, BV, (long)w.nseg, w.cell_capacity});
  w.scan.max_blocks = (int)((scan_cap + 4095) / 4096) + 1;
  w.ray_lanes_target = 4194304;
  w.heavy_lanes = 65536;
  w.long_edge_cells = 48;
  w.bucket_log = 2;
  return 0;
}

static void carve_hierarchy(Carver &c, Hierarchy &h) {
  c.take(h.keys, h.size);
  c.take(h.alt_keys, h.size);
  c.take(h.ids, h.size);
  c.take(h.alt_ids, h.size);
  c.take(h.nodes, 2L * h.size + 4);  // four boxes of slack for group reads
  c.take(h.leaves, 1);
}

// Every pool of a planned workspace. `outputs` false leaves the caller-owned
// outputs (polygons, triangles, status, metrics) to the caller.
void peel::carve(Ws &w, Carver &c, bool outputs) {
  const bool layered = w.route == PEEL_ROUTE_LAYERED;
  const long nf1 = w.nf > 0 ? w.nf : 1, nv1 = w.nv > 0 ? w.nv : 1;
  const long R = w.cap.regions, H = w.cap.hits, BV = w.cap.vertices, T = w.cap.triangles;
  if (outputs) {
    c.take(w.status, 16);
    c.take(w.metrics, M_COUNT);
    c.take(w.out_xy, R * MAXV);
    c.take(w.out_n, R);
    c.take(w.out_src, R);
    c.take(w.out_layer, R);
    c.take(w.tri_uv, T * 6);
    c.take(w.tri_world, T * 9);
    c.take(w.tri_src, T);
    c.take(w.tri_layer, T);
  }
  c.take(w.control, C_COUNT);
  c.take(w.metric_rows, (long)METRIC_ROWS * METRIC_STRIDE);
  c.take(w.inverse, 16);
  c.take(w.canonical, nv1);
  c.take(w.vhash, w.vhash_size);
  c.take(w.etable, w.ehash_size);
  c.take(w.masks, 3 * nf1);
  c.take(w.src_xy, SRCV * nf1);
  c.take(w.src_n, nf1);
  c.take(w.scores, nf1);
  c.take(w.facing, nf1);
  c.take(w.clipped, nf1);
  c.take(w.src_boxes, nf1);
  c.take(w.seg_a, w.nseg);
  c.take(w.seg_b, w.nseg);
  c.take(w.seg_near, 2L * w.nseg);
  w.cl_a = w.seg_a;
  w.cl_b = w.seg_b;
  w.cl_near = reinterpret_cast<double2 *>(w.seg_near);
  c.take(w.seg_rep, w.nseg);
  c.take(w.shash, w.shash_size);
  c.take(w.seg_boxes, w.nseg);
  c.take(w.seg_flag, w.nseg);
  c.take(w.seg_off, w.nseg);
  c.take(w.cseg_a, w.nseg);
  c.take(w.cseg_b, w.nseg);
  c.take(w.cnear, 2L * w.nseg);
  carve_hierarchy(c, w.source_tree);
  carve_hierarchy(c, w.contour_tree);
  c.take(w.leaf_xy, 8L * w.source_tree.size);
  c.take(w.leaf_n, w.source_tree.size);
  c.take(w.leaf_score, w.source_tree.size);
  c.take(w.leaf_xy32, w.filter32 ? 8L * w.source_tree.size : 1);
  c.take(w.leaf_tol, w.filter32 ? w.source_tree.size : 1);
  c.take(w.scores32, w.filter32 ? nf1 : 1);
  // layered route: residual pools sized by the frontier capacity, per-worker
  // piece pools, per-layer statistics
  const long RES = layered ? w.cap.frontier : 1, NF = layered ? nf1 : 1;
  const long POOL = layered ? 3L * (w.pieces + 3) * w.workers : 1;
  c.take(w.ranks, NF);
  c.take(w.contour_boxes, NF);
  for (int par = 0; par < 2; ++par) {
    c.take(w.res_xy[par], RES * MAXV);
    c.take(w.res_n[par], RES);
    c.take(w.res_src[par], RES);
    c.take(w.res_box[par], RES);
  }
  c.take(w.res_keys, RES);
  c.take(w.res_keys_alt, RES);
  c.take(w.res_ids, RES);
  c.take(w.res_ids_alt, RES);
  c.take(w.run_first, NF);
  c.take(w.run_last, NF);
  c.take(w.pool_xy, POOL * MAXV);
  c.take(w.pool_n, POOL);
  c.take(w.stats, 3L * w.cap.layers);
  const long F = layered ? 1 : w.cap.frontier;
  c.take(w.fr_xy, F * MAXV);
  c.take(w.fr_n, F);
  c.take(w.fr_src, F);
  c.take(w.fr_node, F);
  c.take(w.fr_gen, F);
  c.take(w.sub_partials, layered ? 1L : (long)SUB_PARTIAL_WAVES * SUB_COUNTERS);
  c.take(w.samples, R);
  c.take(w.reps, R);
  c.take(w.qhash, w.qhash_size);
  c.take(w.groups, R);
  c.take(w.group_counts, R);
  c.take(w.group_offsets, R);
  c.take(w.group_fill, R);
  c.take(w.ray_ids, R);
  c.take(w.ray_ids_alt, R);
  c.take(w.medium, R);
  c.take(w.heavy, R);
  c.take(w.beam_box, 4L * BEAM_CELLS);
  c.take(w.beam_leader, BEAM_CELLS);
  c.take(w.beam_valid, BEAM_CELLS);
  c.take(w.beam_jobs, BEAM_CELLS);
  c.take(w.hit_depth, H);
  c.take(w.hit_src, H);
  c.take(w.hit_group, H);
  c.take(w.hkeys, H);
  c.take(w.hkeys_alt, H);
  c.take(w.hids, H);
  c.take(w.hids_alt, H);
  c.take(w.owners, BV);
  c.take(w.starts, BV);
  c.take(w.ends, BV);
  c.take(w.partners, BV);
  c.take(w.parents, R);
  c.take(w.choices, R);
  const long MR = w.merge_regions ? R : 1;
  c.take(w.merge_xy, MR * MAXV);
  c.take(w.merge_n, MR);
  c.take(w.merge_src, MR);
  c.take(w.merge_layer, MR);
  c.take(w.eslots, w.ehash2_size);
  c.take(w.vcounts, R);
  c.take(w.voffsets, R);
  c.take(w.centers, R);
  c.take(w.bverts, BV);
  c.take(w.blayers, BV);
  c.take(w.tasks, BV);
  c.take(w.crowded, BV);
  c.take(w.phash, w.phash_size);
  c.take(w.cell_start, w.cell_capacity);
  c.take(w.cell_of, BV);
  c.take(w.cell_xy, BV);
  c.take(w.sort.counts, RS_MAX_SWEEPS * 256);
  c.take(w.sort.tickets, RS_MAX_SWEEPS);
  c.take(w.sort.tiles, (long)RS_MAX_SWEEPS * w.sort.max_blocks * 256);
  c.take(w.sort.flags, RS_MAX_SWEEPS);
  c.take(w.scan.words, w.scan.max_blocks);
  c.take(w.scan.counters, 2);
  w.sort.status = w.status;
  w.scan.status = w.status;
}

extern "C" {

const char *peel_last_error(void) { return g_error.c_str(); }

const char *peel_version(void) { return "peel355 0.2 gfx950"; }

int peel_destroy(PeelWorkspace *ws) {
  if (!ws) return 0;
  if (ws->exec) hipGraphExecDestroy(ws->exec);
  if (ws->graph) hipGraphDestroy(ws->graph);
  if (ws->stream) hipStreamSynchronize(ws->stream);
  if (ws->block) hipFree(ws->block);
  if (ws->stream) hipStreamDestroy(ws->stream);
  delete ws;
  return 0;
}

}  // extern "C"

// The body of peel_create on a fresh workspace; the caller destroys it on
// failure, which releases whatever was already created.
static int create_into(PeelWorkspace *ws, const PeelDesc *d) {
  Ws &w = ws->w;
  int r = plan(w, d);
  if (r) return r;
  HIP_TRY(hipGetDevice(&ws->device));
  Carver measure;
  carve(w, measure, true);
  void *block = nullptr;
  hipError_t e = hipMalloc(&block, measure.used);
  if (e != hipSuccess)
    return fail(-(int)e, "hipMalloc of " + std::to_string(measure.used) + " bytes: " + hipGetErrorString(e));
  ws->block = block;
  ws->bytes = (int64_t)measure.used;
  HIP_TRY(hipMemset(block, 0, measure.used));
  Carver pools;
  pools.base = static_cast<char *>(block);
  carve(w, pools, true);
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
  HIP_TRY(hipStreamCreateWithFlags(&ws->stream, hipStreamNonBlocking));
  HIP_TRY(hipDeviceSynchronize());
  return 0;
}

extern "C" int peel_create(const PeelDesc *d, PeelWorkspace **out) {
  if (!d || !out) return fail(1, "null argument");
  *out = nullptr;
  PeelWorkspace *ws = new PeelWorkspace();
  int r = create_into(ws, d);
  if (r) {
    std::string message = g_error;
    peel_destroy(ws);
    g_error = message;
    return r;
  }
  *out = ws;
  return 0;
}

void peel::enqueue_all(const Ws &w, hipStream_t s, Timeline *t) {
  enqueue_prepare(w, s, t);
  if (w.route == PEEL_ROUTE_LAYERED) {
    enqueue_layered(w, s, t);
  } else {
    enqueue_subdivide(w, s, t);
    enqueue_rank(w, s, t);
  }
  enqueue_mesh(w, s, t);
}

extern "C" {

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

// Node counts per hipGraphNodeType of a captured graph.
static int census(hipGraph_t g, int32_t *counts, int32_t n) {
  if (!g) return fail(9, "no captured graph: capture first");
  for (int i = 0; i < n; ++i) counts[i] = 0;
  size_t total = 0;
  HIP_TRY(hipGraphGetNodes(g, nullptr, &total));
  std::vector<hipGraphNode_t> nodes(total);
  if (total) HIP_TRY(hipGraphGetNodes(g, nodes.data(), &total));
  for (size_t i = 0; i < total; ++i) {
    hipGraphNodeType type;
    HIP_TRY(hipGraphNodeGetType(nodes[i], &type));
    if ((int)type >= 0 && (int)type < n) counts[(int)type] += 1;
  }
  return 0;
}

int peel_graph_census(PeelWorkspace *ws, int32_t *counts, int32_t n) {
  if (!ws || !counts || n < 1) return fail(1, "null argument");
  return census(ws->graph, counts, n);
}

int peel_get_tuning(PeelWorkspace *ws, PeelTuning *t) {
  if (!ws || !t) return fail(1, "null argument");
  const Ws &w = ws->w;
  t->persistent_blocks_per_cu = std::max(1, w.persistent_blocks / ws->cus);
  t->ray_lanes_target = w.ray_lanes_target;
  t->heavy_lanes = w.heavy_lanes;
  t->long_edge_cells = w.long_edge_cells;
  t->leaf_bucket_log2 = w.bucket_log;
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
  if (t->leaf_bucket_log2 < 0 || t->leaf_bucket_log2 > 4) return fail(2, "leaf_bucket_log2 must be in [0, 4]");
  Ws &w = ws->w;
  w.persistent_blocks = ws->cus * t->persistent_blocks_per_cu;
  w.ray_lanes_target = t->ray_lanes_target;
  w.heavy_lanes = t->heavy_lanes;
  w.long_edge_cells = t->long_edge_cells;
  w.bucket_log = t->leaf_bucket_log2;
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
    // members may run concurrently: no launch of theirs may rely on its
    // blocks being co-resident
    Ws member = list[i]->w;
    member.split_huge = 1;
    enqueue_all(member, list[i]->stream, nullptr);
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

int peel_batch_graph_census(PeelBatch *b, int32_t *counts, int32_t n) {
  if (!b || !counts || n < 1) return fail(1, "null argument");
  return census(b->graph, counts, n);
}

int peel_batch_synchronize(PeelBatch *b) {
  if (!b) return fail(1, "null batch");
  HIP_TRY(hipStreamSynchronize(b->members[0]->stream));
  return 0;
}

}  // extern "C"
