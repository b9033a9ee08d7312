// Device workspace of one extraction: every pool, counter and hierarchy,
// allocated once at creation. Kernels receive this struct by value (it lives in
// the kernel-argument segment and is read through the scalar cache).
#pragma once
#include "primitives.hpp"

namespace peel {

struct Caps {
  int frontier;       // frontier slots (seeds + two per accepted split)
  int regions;        // output polygons
  int hits;           // shared ray hits
  int vertices;       // boundary vertices
  int triangles;      // output triangles
  int rounds;         // subdivision generations
  int layers;         // layer capacity
  int edge_vertices;  // edge-noding candidates (<= 128)
};

// One slot of the undirected-edge table: the key (the two canonical vertex
// ids) and, in one word that every incidence adds to, the incidence count
// (bits 0-31), the incidences that start at the smaller vertex (bits 32-47),
// and those of front-facing (48-55) and back-facing (56-63) faces. The fields
// above the count are read only when the count is 2, where none can carry.
struct EdgeSlot {
  unsigned long long key, info;
};
constexpr int EDGE_DIR_SHIFT = 32, EDGE_FRONT_SHIFT = 48, EDGE_BACK_SHIFT = 56;

struct Ws {
  // inputs (caller-owned device arrays)
  const double *vertices;  // [nv][3]
  const int *faces;        // [nf][3]
  const double *camera;    // [4][4] row-major
  const double *settings;  // [6]
  int nv, nf;
  Caps cap;
  int limit, all_layers;
  int share_regions, propagate_ranks, merge_regions;

  // status / metrics / control
  int *status;             // [16]
  long long *metrics;      // [M_COUNT]
  int *control;            // [C_COUNT]
  double *inverse;         // [16] inverse camera

  // welding and edge topology
  int *canonical;          // [nv]
  int *vhash;              // [vhash_size] vertex ids
  int vhash_size;
  EdgeSlot *etable;        // [ehash_size] undirected edge -> incidence summary
  int ehash_size;
  int *masks;              // [3 nf] 1 = contour edge

  // projected sources: rings [SRCV][nf], count, depth plane, facing, clipped
  double2 *src_xy;
  int *src_n;
  double4 *scores;
  int *facing, *clipped;
  Box *src_boxes;          // [nf]

  // contour segments (8 slots per source)
  int nseg;                // 8 nf
  double2 *seg_a, *seg_b;  // [nseg]
  unsigned long long *seg_near;  // [2][nseg] order-preserving keys of min depths
  int *seg_rep;            // [nseg] representative slot (min slot of duplicates)
  int *shash;              // [shash_size]
  int shash_size;
  Box *seg_boxes;          // [nseg] boxes of the compacted unique segments
  int *seg_flag, *seg_off; // [nseg] unique-segment flags and their scan
  double2 *cseg_a, *cseg_b;  // [nseg] compacted unique segments
  double *cnear;           // [2][nseg] their minimum endpoint depths
  // the unique segments again in contour-hierarchy leaf order (endpoints and
  // the two minimum depths), in the memory of seg_a, seg_b and seg_near,
  // which are dead once the segments are compacted
  double2 *cl_a, *cl_b, *cl_near;

  // hierarchies
  Hierarchy source_tree, contour_tree;
  double2 *leaf_xy;        // [source_tree.size][8] rings in leaf order
  int *leaf_n;             // [source_tree.size]
  double4 *leaf_score;     // [source_tree.size]
  int filter32;            // binary32-filtered orientation tests (PEEL_FILTER_FP32)
  float2 *leaf_xy32;       // [source_tree.size][8] binary32 copies of the rings
  float *leaf_tol;         // [source_tree.size] their static error bounds
  float4 *scores32;        // [nf] binary32 copies of the depth planes (filtered depth order)

  // frontier pool [MAXV][frontier]
  double2 *fr_xy;
  int *fr_n, *fr_src, *fr_node, *fr_gen;
  long long *sub_partials;  // [SUB_PARTIAL_WAVES][SUB_COUNTERS]

  // output polygons [MAXV][regions]
  double2 *out_xy;
  int *out_n, *out_src, *out_layer;

  // ranks
  double2 *samples;        // [regions]
  int *reps;               // [regions]
  int *qhash;              // [qhash_size]
  int qhash_size;
  int *groups;             // [regions]
  int *group_counts;       // [regions]
  int *group_offsets;      // [regions]
  int *group_fill;         // [regions] segment fill cursors
  int *ray_ids, *ray_ids_alt;         // [regions] ray leaders; rays left to the depth-first walk
  int *medium;             // [regions] groups sorted by one workgroup
  int4 *heavy;             // [regions] (ray, first unvisited leaf, lane subtree root)
  // rectangle certification of shared rays over 32 x 32 window cells
  unsigned long long *beam_box;  // [4][BEAM_CELLS] order-preserving keys: min x, min y, max x, max y
  int *beam_leader;        // [BEAM_CELLS] smallest patch index in the cell
  int *beam_valid;         // [BEAM_CELLS] 1 = every covering candidate certified
  int2 *beam_jobs;         // [BEAM_CELLS] (cell, first unvisited leaf) handed to the subtree workers
  double *hit_depth;       // [hits]
  int *hit_src, *hit_group;  // [hits]
  uint64_t *hkeys, *hkeys_alt;  // [hits]
  int *hids, *hids_alt;         // [hits]

  // connectivity (patch edges at the scanned ring offsets, [vertices])
  unsigned long long *owners;
  double2 *starts, *ends;
  int *partners;
  int *parents;            // [regions] union-find
  int *choices;            // [regions] merge proposal (smallest patch-edge index)
  double2 *merge_xy;       // [MAXV][regions] compaction pool after merging
  int *merge_n, *merge_src, *merge_layer;  // [regions]
  int *eslots;             // [ehash2_size] reversed-edge hash
  int ehash2_size;

  // meshing
  int *vcounts, *voffsets;  // [regions]
  double2 *centers;         // [regions] polygon centroids
  double2 *bverts;          // [vertices]
  int *blayers;             // [vertices]
  int2 *tasks;              // [vertices] (polygon, edge | vertex count << 16)
  int *crowded;             // [vertices] edges for the per-lane-array noding path
  int *phash;               // [phash_size] boundary vertex dedupe (layer, x, y)
  // per-layer cell grid of the distinct boundary vertices, compressed rows:
  // cell c holds cell_xy[cell_start[c - 1], cell_start[c]) once it is filled
  int *cell_start;          // [cell_capacity] counts, then (scanned and filled) row ends
  long cell_capacity;       // 4 * vertices (at most 2^30): layers * G^2 never exceeds it
  int *cell_of;             // [vertices] cell of a first copy, -1 for a repeat
  double2 *cell_xy;         // [vertices] the distinct vertices, cell by cell
  int phash_size;
  double *tri_uv;           // [triangles][3][2]
  double *tri_world;        // [triangles][3][3]
  int *tri_src, *tri_layer;

  // layered route (DevicePeeler): whole-face ranks, then one clipping pass
  // per layer over the residual pieces
  int route;                // PEEL_ROUTE_EVENTS or PEEL_ROUTE_LAYERED
  int workers, pieces;      // worker threads and pieces per worker pool
  int *ranks;               // [nf] whole-face rank, -1 = residual
  Box *contour_boxes;       // [nf] boxes of each source's contour edges
  // residual pools [MAXV][cap.frontier]: pool 0 holds the pieces of the
  // current pass in (source, parent slot) order, pool 1 receives appends
  double2 *res_xy[2];
  int *res_n[2], *res_src[2];
  Box *res_box[2];
  uint64_t *res_keys, *res_keys_alt;  // [cap.frontier] (source, parent slot) of appended pieces
  int *res_ids, *res_ids_alt;         // [cap.frontier] their append slots
  int res_parent_bits;      // bits of the parent slot in a key
  int res_key_bits;         // bits of a whole key
  int *run_first, *run_last;  // [nf] each source's pieces in pool 0 (-1: none)
  double2 *pool_xy;         // [3][pieces + 3][MAXV][workers] visible, temporary, hidden
  int *pool_n;              // [3][pieces + 3][workers]
  long long *stats;         // [cap.layers][3] targets, candidate tests, fully hidden

  // metric partials: wave v adds to row v % METRIC_ROWS (rows on lines of
  // their own), folded into `metrics` by the pipeline's last kernel
  long long *metric_rows;   // [METRIC_ROWS][METRIC_STRIDE]

  SortScratch sort;
  ScanScratch scan;
  int persistent_blocks;    // grid of the persistent kernels
  // launch tuning (PeelTuning): lanes in flight before rays drop from one
  // subtree per lane to one lane per ray, subtree-job lanes of heavy rays and
  // beams, grid cells beyond which an edge is noded by a whole wave
  int ray_lanes_target;
  int heavy_lanes;
  int long_edge_cells;
  int bucket_log;           // leaf buckets of the ray walks: 2^bucket_log (or twice) leaves
  // hit groups above 2,048 are sorted one phase per launch instead of by one
  // launch with grid barriers (set for graphs whose members run concurrently)
  int split_huge;
};

// Wave-level metric updates. Device atomics on one address serialise (about
// 12 ns each on MI355X, 6 ns on one line), so thousands of waves adding to
// the same metric word would queue behind each other; spread over
// METRIC_ROWS lines they cost next to nothing. Called by every lane of the
// wave (the value is summed, or maximised, across the wave first).
__device__ __forceinline__ long long *metric_row(const Ws &w) {
  int wave = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  return w.metric_rows + (long)(wave & (METRIC_ROWS - 1)) * METRIC_STRIDE;
}
__device__ __forceinline__ void metric_add(const Ws &w, int m, long long v) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd((unsigned long long *)&metric_row(w)[m], (unsigned long long)v);
}
// `v` already uniform across the wave; lane 0 records it
__device__ __forceinline__ void metric_max(const Ws &w, int m, long long v) {
  if ((threadIdx.x & 63) == 0 && v > 0) atomicMax((unsigned long long *)&metric_row(w)[m], (unsigned long long)v);
}

// Optional host-side timeline: hipEvents recorded between stages when the
// pipeline runs eagerly for a stage profile (never inside a capture).
struct Timeline {
  static constexpr int MAX = 64;
  hipEvent_t events[MAX];
  const char *names[MAX];
  int count = 0;
};
inline void mark(Timeline *t, hipStream_t s, const char *name) {
  if (!t || t->count >= Timeline::MAX) return;
  (void)hipEventRecord(t->events[t->count], s);
  t->names[t->count++] = name;
}

// Stage launchers (one translation unit each).
void enqueue_prepare(const Ws &w, hipStream_t s, Timeline *t = nullptr);
void enqueue_subdivide(const Ws &w, hipStream_t s, Timeline *t = nullptr);
void enqueue_rank(const Ws &w, hipStream_t s, Timeline *t = nullptr);
void enqueue_mesh(const Ws &w, hipStream_t s, Timeline *t = nullptr);
void enqueue_layered(const Ws &w, hipStream_t s, Timeline *t = nullptr);
void enqueue_connect(const Ws &w, hipStream_t s);
void enqueue_merge(const Ws &w, hipStream_t s);

}  // namespace peel
