/* peel355 — C-ABI of the MI355X (gfx950) analytic depth peeler.
 *
 * Plain pointers and sizes only; device pointers come from any allocator
 * (the Python host layer passes torch tensor data pointers). Every function
 * returns 0 on success, a negative value for a HIP runtime error and a
 * positive value for an argument error; peel_last_error() describes the last
 * failure of the calling thread.
 *
 * The entry points replace the host interface of the reference's device
 * route (alecjacobson/exact-depth-peeling, warp_peel/), as cited per
 * function. The reference binds its Python classes directly to Warp; the
 * Python package hip_peel binds these symbols with ctypes.
 */
#ifndef PEEL355_H
#define PEEL355_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct PeelWorkspace PeelWorkspace;

enum PeelRoute {
  PEEL_ROUTE_EVENTS = 0,   /* EventPeeler: warp_peel/events.py:36-845 */
  PEEL_ROUTE_LAYERED = 1   /* DevicePeeler API: warp_peel/solver.py:92-519 */
};

enum PeelFlags {
  PEEL_SHARE_REGIONS = 1,    /* events.py:52 share_regions */
  PEEL_PROPAGATE_RANKS = 2,  /* events.py:53 propagate_ranks */
  PEEL_MERGE_REGIONS = 4,    /* events.py:54 merge_regions */
  PEEL_FILTER_FP32 = 8       /* binary32-filtered predicates, binary64 fallback */
};

/* Constructor arguments: warp_peel/solver.py:103-147 (DevicePeeler.__init__)
 * and warp_peel/events.py:43-138 (EventPeeler.__init__, EventCapacities
 * events.py:19-33). Inputs are device arrays: vertices f64 [V][3], faces i32
 * [F][3], camera f64 [4][4] row-major affine world-to-camera, settings f64 [6]
 * = xmin, ymin, xmax, ymax, perspective (0/1), near. The caller owns them;
 * their values may change between launches, their addresses may not. */
typedef struct PeelDesc {
  const double *vertices;
  int64_t num_vertices;
  const int32_t *faces;
  int64_t num_faces;
  const double *camera;
  const double *settings;
  int32_t frontier;       /* frontier slots: seeds + two per contour cut (event
                             route); residual pool (layered route) */
  int32_t regions;        /* output polygons */
  int32_t hits;           /* shared ray hits */
  int32_t vertices_cap;   /* boundary vertices */
  int32_t triangles;      /* output triangles */
  int32_t rounds;         /* subdivision generations */
  int32_t layer_capacity; /* maximum layers */
  int32_t edge_vertices;  /* edge-noding candidates, <= 128 */
  int32_t layers;         /* 0 = all layers, else prefix length */
  int32_t route;          /* PeelRoute */
  int32_t flags;          /* PeelFlags */
  int32_t workers;        /* layered route: worker threads (solver.py:10-31 Capacities.workers) */
  int32_t pieces;         /* layered route: pieces per worker pool (Capacities.pieces_per_worker) */
} PeelDesc;

/* Device outputs (warp_peel/README.md:186-189 and solver.py:156-200): valid
 * prefixes are delimited by status[2] (polygons) and status[3] (triangles).
 * Polygon rings are vertex-major: vertex j of polygon i is output_xy[2 *
 * (j * regions + i) + {0, 1}]. Triangles are [T][3][2] uv and [T][3][3] world. */
typedef struct PeelBuffers {
  double *output_xy;
  int32_t *output_n;
  int32_t *output_sources;
  int32_t *output_layers;
  int32_t regions;
  int32_t max_vertices;
  double *triangles_uv;
  double *triangles_world;
  int32_t *triangle_sources;
  int32_t *triangle_layers;
  int32_t triangles;
  int32_t *status;     /* [16] */
  int64_t *metrics;    /* [metrics_count] */
  int32_t metrics_count;
  int32_t *canonical;  /* [V] welded vertex ids */
  void *stream;        /* hipStream_t owned by the workspace */
  int64_t *stats;      /* layered route: [stats_rows][3] targets, candidate tests,
                          fully hidden targets per layer (solver.py:153) */
  int32_t stats_rows;
  int64_t device_bytes; /* device memory held by the workspace's pools */
} PeelBuffers;

/* Allocate every pool and create the workspace stream
 * (replaces DevicePeeler.__init__ / EventPeeler.__init__). */
int peel_create(const PeelDesc *desc, PeelWorkspace **out);

/* Run the pipeline once, eagerly, before any capture
 * (replaces warmup(), solver.py:447-457 and events.py:782-796). */
int peel_warmup(PeelWorkspace *ws);

/* Record the pipeline on the workspace stream: launches only, no allocation,
 * synchronisation or readback; valid inside an external hipStreamBeginCapture
 * on that stream (replaces enqueue(), solver.py:459-469, events.py:798-806). */
int peel_enqueue(PeelWorkspace *ws);

/* Capture the pipeline into the workspace graph (replaces capture(),
 * solver.py:471-480). */
int peel_capture(PeelWorkspace *ws);

/* Asynchronous replay on the workspace stream, capturing first if needed
 * (replaces run(), solver.py:482-486). */
int peel_launch(PeelWorkspace *ws);

/* Device pointers and capacities of every output. */
int peel_get_buffers(PeelWorkspace *ws, PeelBuffers *out);

/* Synchronise the workspace stream and copy status[16] and the metrics to
 * host memory (replaces read_status(), solver.py:488-519, events.py:808-845;
 * error decoding stays in the host layer). */
int peel_read_status(PeelWorkspace *ws, int32_t *status, int64_t *metrics);

/* Run the pipeline once, eagerly, with hipEvents between stages and around
 * the hot kernels; writes up to *count stage times (ms) and their
 * comma-separated names, then sets *count. Never valid inside a capture
 * (replaces the per-stage diagnostics of scripts/profile_events.py:36-55). */
int peel_profile(PeelWorkspace *ws, int32_t *count, float *ms, char *names, int32_t names_len);

/* Launch tuning of one workspace. No counterpart in the reference, whose
 * launch shapes are fixed; it serves callers that tune a workload once and
 * replay it many times (hip_peel.tune, an autotuner over these fields).
 * Outputs do not depend on these values, only the time. */
typedef struct PeelTuning {
  int32_t persistent_blocks_per_cu; /* k_subdivide work-queue blocks per CU, 1..max_blocks_per_cu */
  int32_t ray_lanes_target;         /* lanes in flight allowed for rays of 64 lanes each (at most 2,048 rays), else 16 lanes per ray; >= 64 */
  int32_t heavy_lanes;              /* lane budget of the heavy-ray and beam subtree jobs, >= 1024 */
  int32_t long_edge_cells;          /* grid cells beyond which a polygon edge is noded by a wave, >= 1 */
  int32_t leaf_bucket_log2;         /* ray walks test 2^n (or 2^(n+1)) hierarchy leaves together, 0..4 */
  int32_t max_blocks_per_cu;        /* read only: resident k_subdivide blocks per CU */
} PeelTuning;
int peel_get_tuning(PeelWorkspace *ws, PeelTuning *tuning);
/* Validates every field (2 on a bad value) and drops the captured graph: the
 * next peel_launch captures again. A batch holding the workspace must be
 * created again. */
int peel_set_tuning(PeelWorkspace *ws, const PeelTuning *tuning);

/* Nodes of the captured graph by type: counts[t] = nodes of hipGraphNodeType
 * t for t < n (kernel 0, memcpy 1, memset 2, host 3, empty 5, event wait 6,
 * event record 7, alloc 10, free 11). Returns 9 before the first capture.
 * It restates the reference's graph-content check (tests/gpu/test_peeler.py:
 * 197-202: no device-to-host, host-to-device or allocation nodes) and
 * counts the launches of one extraction. */
int peel_graph_census(PeelWorkspace *ws, int32_t *counts, int32_t n);

/* Unit-test entry into the device's ring functions, on one lane (no
 * counterpart in the reference's API; it restates the reference's tests that
 * launch its device functions directly, tests/gpu/test_peeler.py:331-375).
 * op 0: clean (warp_peel/kernels.py:146-176); op 1: split_edge by the edge
 * param = (ax, ay, bx, by) (kernels.py:206-212, outputs made strictly convex
 * as the layered route makes them); op 2: the strict-convexity rule. Device
 * pointers: xy [n <= 32][2], out_xy [2][32][2] (inside, outside), out_n [2];
 * enqueued on `stream` without synchronising. Returns 0, 1 on bad arguments,
 * -1 on a launch failure. */
int peel_ring_test(int32_t op, const double *xy, int32_t n, const double *param, double *out_xy, int32_t *out_n,
                   void *stream);

/* Synchronise the workspace stream. */
int peel_synchronize(PeelWorkspace *ws);

/* Release every pool, the graph and the stream. */
int peel_destroy(PeelWorkspace *ws);

/* Batches: one graph that replays several workspaces' pipelines, each on its
 * own stream, forked from and joined to the first workspace's stream, so
 * independent meshes (each with its own camera) peel in one launch and may
 * overlap on the GPU. Every workspace must stay alive, and must not be
 * launched on its own, while the batch is in use; each keeps its own outputs
 * and status (read them with peel_read_status after peel_batch_synchronize).
 * No counterpart in the reference, which peels one mesh per graph
 * (warp_peel/events.py:798-806); it serves callers with many small meshes. */
typedef struct PeelBatch PeelBatch;
int peel_batch_create(PeelWorkspace *const *workspaces, int32_t count, PeelBatch **out);
int peel_batch_launch(PeelBatch *batch);
int peel_batch_synchronize(PeelBatch *batch);
int peel_batch_destroy(PeelBatch *batch);
/* Node census of the batch graph, as peel_graph_census. */
int peel_batch_graph_census(PeelBatch *batch, int32_t *counts, int32_t n);

/* Learned pool sizes. No counterpart in the reference, whose capacities are
 * fixed multiples of the face count (warp_peel/events.py:19-33); it serves
 * callers that admit many extractions into bounded memory. A multilayer
 * perceptron (inputs <= 16 -> 64 tanh -> 64 tanh -> outputs <= 8, binary32,
 * parameters flat [W1 b1 W2 b2 W3 b3], row-major) predicts per-pool usage
 * from mesh statistics; an overflowing prediction is caught by the status
 * bits and rerun with the conservative capacities. All pointers are device
 * memory; `stream` is a hipStream_t (NULL = default stream); nothing here
 * allocates or synchronises. Returns 0, or 1 for bad shapes/null pointers,
 * 2 for bad scalars, -1 for a failed launch. */
int32_t pool_mlp_param_count(int32_t inputs, int32_t outputs);
/* Floats of the gradient workspace `partial` for a batch. */
int32_t pool_mlp_partial_count(int32_t inputs, int32_t outputs, int32_t batch);
/* y[batch][outputs] = network(x[batch][inputs]). */
int pool_mlp_forward(const float *params, int32_t inputs, int32_t outputs, const float *x, int32_t batch,
                     float *y, void *stream);
/* grad[params] and *loss of the asymmetric squared error
 * 1/(N O) sum w(r) r^2, r = y - target, w = under_weight where r < 0, else 1
 * (loss may be NULL); deterministic (fixed summation order). */
int pool_mlp_gradient(const float *params, int32_t inputs, int32_t outputs, const float *x, const float *target,
                      int32_t batch, float under_weight, float *partial, float *grad, float *loss, void *stream);
/* One Adam step; bias1 = 1 - beta1^t and bias2 = 1 - beta2^t. */
int pool_mlp_adam(float *params, const float *grad, float *m, float *v, int32_t count, float lr, float beta1,
                  float beta2, float eps, float bias1, float bias2, void *stream);

/* One extraction that allocates nothing. No counterpart in the reference; it
 * serves callers that cannot allocate device memory (code under graph
 * capture, tight serving loops, harnesses that forbid hipMalloc). Inputs and
 * capacities as for peel_create; the caller owns the outputs, sized by
 * desc->regions and desc->triangles: output_xy [32][regions][2] vertex-major,
 * output_n/sources/layers [regions], triangles_uv [triangles][3][2],
 * triangles_world [triangles][3][3], triangle_sources/layers [triangles],
 * status [16] and metrics [PEEL_METRICS_COUNT], with the meaning of
 * PeelBuffers. Every other pool is carved from a device array of
 * PEEL_STATIC_SCRATCH_BYTES that each GPU holds for the life of the process;
 * a descriptor needing more returns 2 with the needed size in
 * peel_last_error(). The pipeline is enqueued on `stream` (hipStream_t, NULL =
 * the null stream) without synchronising; read the status after the stream.
 * Calls on one GPU share its array: calls that overlap on one GPU (other
 * streams, other threads) race and must be ordered by the caller. */
#define PEEL_STATIC_SCRATCH_BYTES (256ull << 20)
#define PEEL_METRICS_COUNT 30
typedef struct PeelOutputs {
  double *output_xy;
  int32_t *output_n;
  int32_t *output_sources;
  int32_t *output_layers;
  double *triangles_uv;
  double *triangles_world;
  int32_t *triangle_sources;
  int32_t *triangle_layers;
  int32_t *status;
  int64_t *metrics;
} PeelOutputs;
int peel_extract_static(const PeelDesc *desc, const PeelOutputs *outputs, void *stream);

/* Message of the last failure on this thread ("" if none). */
const char *peel_last_error(void);

/* Library version string. */
const char *peel_version(void);

#ifdef __cplusplus
}
#endif
#endif
