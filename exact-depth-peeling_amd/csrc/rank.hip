// Stage 3: shared ray ranks for the finished patches.
//
// Reference semantics (file:line in alecjacobson/exact-depth-peeling):
//  samples          warp_peel/event_kernels.py:28-48
//  exact groups     warp_peel/event_kernels.py:51-92 (group_queries)
//  shared rays      warp_peel/beams.py:18-207 (cell, group, covers, certify,
//                   certify_heavy, apply)
//  contains_flat    warp_peel/event_kernels.py:131-148 (half-open ownership)
//  store_hit        warp_peel/event_kernels.py:151-169
//  ray_hits         warp_peel/event_kernels.py:196-245
//  heavy_ray_hits   warp_peel/event_kernels.py:248-298
//  hit order        warp_peel/events.py:703-718 (stable source, depth, group)
//  assign_ranks     warp_peel/event_kernels.py:342-385
//
// Samples with identical coordinates share one ray group; the group leader is
// the patch that claimed the sample's hash slot first (a hash table instead of
// the reference's two float64 radix sorts and a binary search). Which member
// leads does not matter: all members cast the same ray.
#include <algorithm>

#include "workspace.hpp"

namespace peel {

__device__ __forceinline__ int npolys(const Ws &w) {
  int p = load_i(w.status + ST_POLYS);
  return p < w.cap.regions ? p : w.cap.regions;
}

__device__ __forceinline__ uint64_t point_hash(double2 x) {
  return mix64(dbits(x.x) ^ mix64(dbits(x.y) + 0x9e3779b97f4a7c15ull));
}

__device__ __forceinline__ int qhash_live(const Ws &w) { return live_table(npolys(w), w.qhash_size); }

// Connectivity reset (components.py:32-52): every patch its own parent and
// representative, and whether patches connect at all: only after a split or
// with an internal mesh edge.
__device__ __forceinline__ void connect_reset(const Ws &w, int i) {
  int P = npolys(w);
  for (int j = i; j < P; j += gridDim.x * blockDim.x) {
    w.parents[j] = j;
    w.reps[j] = j;
    w.choices[j] = 0x7fffffff;
  }
  if (i == 0) {
    int any = (w.metrics[M_SPLITS] > 0) || (w.control[C_INTERNAL] != 0);
    w.control[C_CONNECT] = any;
    if (!any && w.propagate_ranks) w.metrics[M_COMPONENTS] = P;
  }
}

// The subdivision kernel's per-wave partial counters, summed by block 0 (256
// threads) into the metrics.
__device__ void sum_subdivide_partials(const Ws &w) {
  __shared__ long long sums[SUB_COUNTERS][4];
  const int waves = min(w.persistent_blocks * 4, SUB_PARTIAL_WAVES);
  long long part[SUB_COUNTERS];
  for (int k = 0; k < SUB_COUNTERS; ++k) part[k] = 0;
  for (int v = threadIdx.x; v < waves; v += blockDim.x)
    for (int k = 0; k < SUB_COUNTERS; ++k) {
      long long x = w.sub_partials[(long)v * SUB_COUNTERS + k];
      part[k] = k >= SUB_SUMS ? max(part[k], x) : part[k] + x;
    }
  for (int k = 0; k < SUB_COUNTERS; ++k) {
    long long x = part[k];
    for (int off = 32; off > 0; off >>= 1) {
      long long y = __shfl_xor(x, off, 64);
      x = k >= SUB_SUMS ? max(x, y) : x + y;
    }
    if ((threadIdx.x & 63) == 0) sums[k][threadIdx.x >> 6] = x;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int slot[SUB_SUMS] = {M_VISITS, M_CONTOUR_TESTS, M_SPLITS, M_FILTER_STRADDLE,
                                M_FALLBACK_STRADDLE, M_FILTER_FRONT, M_FALLBACK_FRONT};
    for (int k = 0; k < SUB_COUNTERS; ++k) {
      long long x = sums[k][0];
      for (int q = 1; q < (int)(blockDim.x >> 6); ++q)
        x = k >= SUB_SUMS ? max(x, sums[k][q]) : x + sums[k][q];
      if (k < SUB_SUMS)
        w.metrics[slot[k]] += x;  // waves past the partials added theirs directly
      else if (k == SUB_SUMS)
        w.metrics[M_ROUNDS] = max(w.metrics[M_ROUNDS], x);
      else
        w.status[ST_PEAK] = max(w.status[ST_PEAK], (int)x);
    }
  }
}

// Resets of the rank stage and of the connectivity it starts with, one launch.
// Block 0 first sums the subdivision counters (the connectivity reset reads
// the split count).
__global__ __launch_bounds__(256) void k_rank_reset(Ws w) {
  if (blockIdx.x == 0) sum_subdivide_partials(w);
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int size = qhash_live(w);
  for (int j = i; j < size; j += gridDim.x * blockDim.x) w.qhash[j] = -1;
  for (int j = i; j < BEAM_CELLS; j += gridDim.x * blockDim.x) {
    w.beam_box[j] = ~0ull;
    w.beam_box[BEAM_CELLS + j] = ~0ull;
    w.beam_box[2 * BEAM_CELLS + j] = 0ull;
    w.beam_box[3 * BEAM_CELLS + j] = 0ull;
    w.beam_leader[j] = 0x7fffffff;
    w.beam_valid[j] = 0;
  }
  if (i == 0) {
    w.control[C_BEAM_JOBS] = 0;
    w.control[C_HEAVY] = 0;
    w.control[C_MEDIUM] = 0;
    w.control[C_HUGE] = 0;
    w.control[C_HUGE_SYNC] = 0;
    w.control[C_LEADERS] = 0;
    // subdivision totals, for read_status
    w.metrics[M_SLOTS] = w.control[C_ALLOC];
    if (w.nf == 0) w.metrics[M_ROUNDS] = 0;
  }
  if (w.propagate_ranks || w.merge_regions) connect_reset(w, i);
}

// Canonical sample of the representative patch: centroid summed from the
// lexicographically smallest vertex (rotation invariant).
__device__ void sample_one(const Ws &w, int i);

__global__ void k_samples(Ws w) {
  if (w.status[ST_ERR] != 0) return;
  ITEM_LOOP(i, npolys(w)) sample_one(w, i);
}

__device__ void sample_one(const Ws &w, int i) {
  if (!w.propagate_ranks) w.reps[i] = i;
  int r = w.reps[i];
  int n = w.out_n[r];
  long R = w.cap.regions;
  int anchor = 0;
  double2 best = w.out_xy[r];
  for (int j = 1; j < n; ++j) {
    double2 a = w.out_xy[j * R + r];
    if (lexless(a, best)) {
      anchor = j;
      best = a;
    }
  }
  double2 x = make_double2(0.0, 0.0);
  for (int j = 0; j < n; ++j) {
    double2 a = w.out_xy[((anchor + j) % n) * R + r];
    x.x += a.x / (double)n;
    x.y += a.y / (double)n;
  }
  w.samples[i] = x;
  w.group_counts[i] = 0;
  w.group_fill[i] = 0;
}

// Exact sample groups: the first patch to claim a sample's slot leads it. The
// samples compared here were written by k_samples (a previous kernel), so the
// probe needs no fences. Only component representatives claim: every other
// patch carries its representative's sample bit for bit, and k_groups finds
// the claimed slot by probing. A representative whose left neighbour in the
// wave is a representative with the same sample leaves the claim to it, so a
// run of equal samples (the identical-footprint stack) issues one compare-and-
// swap per wave instead of one per lane on a single address.
__device__ void sample_insert_one(const Ws &w, int i, double2 x);

__global__ void k_sample_insert(Ws w) {
  if (w.status[ST_ERR] != 0) return;
  const int n = npolys(w);
  const int lane = threadIdx.x & 63;
  WAVE_LOOP(base, n) {
    const int i = base + threadIdx.x;
    const bool live = i < n;
    const double2 x = live ? w.samples[i] : make_double2(0.0, 0.0);
    const int claim = live && w.reps[i] == i;
    const double left_x = __shfl_up(x.x, 1, 64), left_y = __shfl_up(x.y, 1, 64);
    const int left_claim = __shfl_up(claim, 1, 64);
    const bool repeat = lane > 0 && left_claim && left_x == x.x && left_y == x.y;
    if (claim && !repeat) sample_insert_one(w, i, x);
  }
}

__device__ void sample_insert_one(const Ws &w, int i, double2 x) {
  int size = qhash_live(w), mask = size - 1;
  int slot = (int)(point_hash(x) & (uint64_t)mask);
  for (int probe = 0; probe < size; ++probe) {
    int cur = load_i(&w.qhash[slot]);
    if (cur < 0) {
      cur = atomicCAS(&w.qhash[slot], -1, i);
      if (cur < 0) return;
    }
    if (same(w.samples[cur], x)) return;
    slot = (slot + 1) & mask;
  }
}

// cell (beams.py:18-22): 32 x 32 buckets of the unit window, clamped
__device__ __forceinline__ int beam_cell(double2 x) {
  double cx = fmin(fmax(x.x * (double)BEAM_SIDE, 0.0), (double)(BEAM_SIDE - 1));
  double cy = fmin(fmax(x.y * (double)BEAM_SIDE, 0.0), (double)(BEAM_SIDE - 1));
  return (int)cx * BEAM_SIDE + (int)cy;
}

__device__ void group_one(const Ws &w, int i);

__global__ void k_groups(Ws w) {
  if (w.status[ST_ERR] != 0) return;
  ITEM_LOOP(i, npolys(w)) group_one(w, i);
}

__device__ void group_one(const Ws &w, int i) {
  double2 x = w.samples[i];
  int size = qhash_live(w), mask = size - 1;
  int slot = (int)(point_hash(x) & (uint64_t)mask);
  int leader = i;
  for (int probe = 0; probe < size; ++probe) {
    int cur = w.qhash[slot];
    if (cur < 0) break;
    if (same(w.samples[cur], x)) {
      leader = cur;
      break;
    }
    slot = (slot + 1) & mask;
  }
  w.groups[i] = leader;
  if (w.share_regions && leader == i) {
    // beams.py:53-85: the cell's bounding rectangle of samples and its leader
    // (one contribution per distinct sample)
    int c = beam_cell(x);
    atomicMin(&w.beam_box[c], (unsigned long long)dkey(x.x));
    atomicMin(&w.beam_box[BEAM_CELLS + c], (unsigned long long)dkey(x.y));
    atomicMax(&w.beam_box[2 * BEAM_CELLS + c], (unsigned long long)dkey(x.x));
    atomicMax(&w.beam_box[3 * BEAM_CELLS + c], (unsigned long long)dkey(x.y));
    atomicMin(&w.beam_leader[c], i);
  }
}

// ---------------------------------------------------------------------------
// Shared rays over certified rectangles (beams.py). Samples fall into 32 x 32
// cells of the unit window. A cell whose samples are not all identical
// proposes one ray for all its patches; the proposal is accepted only if every
// source whose padded box meets the samples' bounding rectangle strictly
// contains all four corners (margin EPS * edge length). On resolved input the
// set of covering sources and their depth order are then constant over the
// rectangle, so one ray orders them for every patch in the cell.
__device__ __forceinline__ double2 beam_corner(const Ws &w, int c, int k) {
  return make_double2(unkey(w.beam_box[(k & 1 ? 2 : 0) * BEAM_CELLS + c]),
                      unkey(w.beam_box[(k & 2 ? 3 : 1) * BEAM_CELLS + c]));
}

__device__ __forceinline__ void beam_box(const Ws &w, int c, float4 &lo, float4 &hi) {
  double2 a = beam_corner(w, c, 0), b = beam_corner(w, c, 3);
  lo = make_float4((float)a.x - PAD, (float)a.y - PAD, -1.0e30f, 0.f);
  hi = make_float4((float)b.x + PAD, (float)b.y + PAD, 1.0e30f, 0.f);
}

// covers (beams.py:88-97) on a leaf-ordered source record
__device__ bool covers_leaf(const Ws &w, int leaf, const double2 *corner) {
  int n = w.leaf_n[leaf];
  const double2 *r = w.leaf_xy + 8L * leaf;
  for (int j = 0; j < n; ++j) {
    double2 a = r[j], b = r[(j + 1) % n];
    double margin = EPS * dlength(make_double2(b.x - a.x, b.y - a.y));
    for (int k = 0; k < 4; ++k)
      if (cross3(a, b, corner[k]) <= margin) return false;
  }
  return n > 0;
}

// certify (beams.py:109-147): one lane per cell, the first HANDOFF candidates
__global__ __launch_bounds__(256) void k_certify(Ws w) {
  int c = blockIdx.x * blockDim.x + threadIdx.x;
  bool ok = c < BEAM_CELLS && w.status[ST_ERR] == 0 && w.beam_leader[c] != 0x7fffffff;
  double2 corner[4];
  if (ok) {
    for (int k = 0; k < 4; ++k) corner[k] = beam_corner(w, c, k);
    // identical samples already share their exact group
    ok = !(corner[0].x == corner[3].x && corner[0].y == corner[3].y);
  }
  int node = ok ? 1 : 0, visits = 0;  // binary continuation, for the hand-off
  bool accepted = ok;
  Tree tree = w.source_tree.tree();
  float4 lo, hi;
  TrailWalk walk;
  walk.trail = 0;
  walk.bucket = 0;
  if (ok) {
    beam_box(w, c, lo, hi);
    walk.start(tree, 1, w.bucket_log, lo, hi);
  }
  for (int step = 0; step < 32 && node > 0; ++step) {
    int leaf_node = walk.next_node(tree, lo, hi);
    if (leaf_node < 0) break;
    node = successor(leaf_node);
    int leaf = leaf_node - tree.leaves;
    visits += 1;
    if (!covers_leaf(w, leaf, corner)) {
      accepted = false;
      break;
    }
  }
  if (ok) w.beam_valid[c] = accepted;
  bool handoff = accepted && node > 0;
  if (handoff)
    while (node < tree.leaves) node *= 2;
  int slot = wave_claim32(w.control + C_BEAM_JOBS, handoff);
  if (handoff) w.beam_jobs[slot] = make_int2(c, node);
  metric_add(w, M_BEAM_ATTEMPTS, ok);
  metric_add(w, M_BEAM_TESTS, visits);
  metric_add(w, M_HEAVY_BEAMS, handoff);
}

// certify_heavy (beams.py:150-189): the unvisited suffix of each handed-off
// cell in disjoint leaf subtrees; a failure anywhere rejects the cell.
__global__ __launch_bounds__(256) void k_certify_heavy(Ws w, int workers) {
  bool ok = w.status[ST_ERR] == 0;
  Tree tree = w.source_tree.tree();
  int jobs = ok ? load_i(w.control + C_BEAM_JOBS) : 0;
  int parts = min(8192, tree.leaves);
  while (parts > 16 && (long long)parts * jobs > w.heavy_lanes) parts /= 2;
  int width = tree.leaves / parts;
  long total = (long)jobs * parts;
  long long visits = 0;
  for (long task = (long)blockIdx.x * blockDim.x + threadIdx.x; task < total; task += workers) {
    int2 job = w.beam_jobs[task / parts];
    int c = job.x, part = (int)(task % parts);
    int first = tree.leaves + part * width;
    if (job.y >= first + width || load_i(w.beam_valid + c) == 0) continue;
    int root = parts + part;
    double2 corner[4];
    for (int k = 0; k < 4; ++k) corner[k] = beam_corner(w, c, k);
    float4 lo, hi;
    beam_box(w, c, lo, hi);
    TrailWalk walk;
    if (job.y > first)
      walk.start_at(tree, root, job.y, w.bucket_log, lo, hi);
    else
      walk.start(tree, root, w.bucket_log, lo, hi);
    while (true) {
      int leaf_node = walk.next_node(tree, lo, hi);
      if (leaf_node < 0) break;
      int leaf = leaf_node - tree.leaves;
      visits += 1;
      if (!covers_leaf(w, leaf, corner)) {
        atomicExch(w.beam_valid + c, 0);
        break;
      }
    }
  }
  metric_add(w, M_BEAM_TESTS, visits);
}

// apply (beams.py:192-207): every patch of a certified cell joins its leader
__global__ __launch_bounds__(256) void k_beam_apply(Ws w) {
  long long certified = 0;
  if (w.status[ST_ERR] == 0) {
    ITEM_LOOP(i, npolys(w)) {
      int c = beam_cell(w.samples[i]);
      if (w.beam_valid[c]) {
        int leader = w.beam_leader[c];
        w.groups[i] = leader;
        certified += leader == i;
      }
    }
  }
  metric_add(w, M_BEAMS, certified);
}

// Source rings in hierarchy leaf order: leaf k holds its ring as one 128-byte
// record (8 double2), so the leaves a ray visits in depth-first order are
// adjacent in memory (the flat source cache of event_kernels.py:110-128,
// reordered by Morton code).
__global__ void k_leaf_sources(Ws w) {
  int k = blockIdx.x * blockDim.x + threadIdx.x;
  const Hierarchy &h = w.source_tree;
  if (k >= w.nf) return;
  int id = h.ids[k];
  int n = 0;
  // the ring is loaded whole before it is stored (a store between two loads
  // would make each load wait for the one before)
  double2 ring[SRCV];
  if (id < w.nf && w.status[ST_ERR] == 0) {
    n = w.src_n[id];
#pragma unroll
    for (int j = 0; j < SRCV; ++j)
      if (j < n) ring[j] = w.src_xy[(long)j * w.nf + id];
#pragma unroll
    for (int j = 0; j < SRCV; ++j)
      if (j < n) w.leaf_xy[8L * k + j] = ring[j];
    w.leaf_score[k] = w.scores[id];
  }
  w.leaf_n[k] = n;
  if (w.filter32) {
    if (id < w.nf) {
      double4 sc = w.scores[id];
      w.scores32[id] = make_float4((float)sc.x, (float)sc.y, (float)sc.z, 0.f);
    }
    // binary32 copy of the ring and the static forward error bound of its
    // orientation tests against any point a ray query can bring to this leaf
    double lx = 1e300, ly = 1e300, hx = -1e300, hy = -1e300;
#pragma unroll
    for (int j = 0; j < SRCV; ++j) {
      if (j >= n) continue;
      const double2 v = ring[j];
      w.leaf_xy32[8L * k + j] = make_float2((float)v.x, (float)v.y);
      lx = fmin(lx, v.x), ly = fmin(ly, v.y), hx = fmax(hx, v.x), hy = fmax(hy, v.y);
    }
    float tol = 0.f;
    if (n > 0) {
      // A query point reaches this leaf only if its ray box (x +- 2e-7)
      // meets the leaf box (ring bounds +- 2e-7): it lies within 4e-7 of the
      // ring's bounds plus the binary32 roundings of x, of both paddings and
      // of the box corners (each at most 2^-24 of a coordinate).
      double Mr = fmax(fmax(fabs(lx), fabs(hx)), fmax(fabs(ly), fabs(hy)));
      double D = fmax(hx - lx, hy - ly) + 4.0e-7 + 8.0 * 5.9604644775390625e-8 * (Mr + 1.0e-6);
      double M = Mr + D;
      tol = orientation_tol32(D, M);
    }
    w.leaf_tol[k] = tol;
  }
}

// Half-open point-in-convex test (event_kernels.py:131-148): a point on an
// edge belongs to the ring only when the edge is "owned" (b.y > a.y, or
// b.y == a.y and b.x < a.x), so a tessellated sheet counts once.
// The whole 128-byte record is loaded with the count, so the test waits on
// one round trip instead of the count and then the vertices.
__device__ __forceinline__ bool contains_leaf(const Ws &w, int k, double2 x) {
  const double2 *r = w.leaf_xy + 8L * k;
  double2 v[SRCV];
#pragma unroll
  for (int j = 0; j < SRCV; ++j) v[j] = r[j];
  int n = w.leaf_n[k];
  bool inside = n > 0;
#pragma unroll
  for (int j = 0; j < SRCV; ++j) {
    if (j < n && inside) {
      double2 a = v[j], b = j + 1 < n ? v[(j + 1) % SRCV] : v[0];
      double s = cross3(a, b, x);
      bool owned = b.y > a.y || (b.y == a.y && b.x < a.x);
      if (s < 0.0 || (s == 0.0 && !owned)) inside = false;
    }
  }
  return inside;
}

// The same test with a binary32 filter: an edge whose binary32 orientation
// exceeds the record's static error bound in magnitude has the sign, and a
// nonzero value, that binary64 gives; only the remaining edges are evaluated
// in binary64, with the half-open rule. The decision therefore equals
// contains_leaf's for every edge.
__device__ __forceinline__ bool contains_leaf32(const Ws &w, int k, double2 x, float2 x32, int &filtered,
                                                int &fallbacks) {
  const float2 *r = w.leaf_xy32 + 8L * k;
  float2 v[SRCV];
#pragma unroll
  for (int j = 0; j < SRCV; ++j) v[j] = r[j];
  int n = w.leaf_n[k];
  const float tol = w.leaf_tol[k];
  bool inside = n > 0;
#pragma unroll
  for (int j = 0; j < SRCV; ++j) {
    if (!(j < n && inside)) continue;
    float2 a = v[j], b = j + 1 < n ? v[(j + 1) % SRCV] : v[0];
    float s32 = (b.x - a.x) * (x32.y - a.y) - (b.y - a.y) * (x32.x - a.x);
    if (s32 < -tol) {
      inside = false;
      filtered += 1;
    } else if (s32 > tol) {
      filtered += 1;
    } else {
      fallbacks += 1;
      double2 A = w.leaf_xy[8L * k + j], B = w.leaf_xy[8L * k + (j + 1) % n];
      double s = cross3(A, B, x);
      bool owned = B.y > A.y || (B.y == A.y && B.x < A.x);
      if (s < 0.0 || (s == 0.0 && !owned)) inside = false;
    }
  }
  return inside;
}

__device__ __forceinline__ bool contains_any(const Ws &w, int k, double2 x, float2 x32, int &filtered,
                                             int &fallbacks) {
  return w.filter32 ? contains_leaf32(w, k, x, x32, filtered, fallbacks) : contains_leaf(w, k, x);
}

// store_hit (event_kernels.py:151-169), staged: a wave collects its hits in
// LDS and claims pool slots for them with one atomic per 64 or more. (A claim
// per iteration made every ray walk wait on the one pool counter, whose
// atomics serialise: 5-20 us per candidate on the torus knot and the shells.)
constexpr int HIT_STAGE = 128;  // a flush starts at 64 staged hits, a step adds at most 64
struct HitStage {
  volatile double *depth;
  volatile int *src, *group;
  int staged;  // wave-uniform: every lane calls with the same history
};
__shared__ double hit_stage_depth[4][HIT_STAGE];
__shared__ int hit_stage_src[4][HIT_STAGE], hit_stage_group[4][HIT_STAGE];

__device__ __forceinline__ HitStage hit_stage() {
  int wave = threadIdx.x >> 6;
  return HitStage{hit_stage_depth[wave], hit_stage_src[wave], hit_stage_group[wave], 0};
}
__device__ void flush_hits(const Ws &w, HitStage &st) {
  if (st.staged == 0) return;
  __builtin_amdgcn_wave_barrier();
  const int lane = threadIdx.x & 63;
  unsigned long long base = 0;
  if (lane == 0) base = atomicAdd((unsigned long long *)&w.metrics[M_RAY_HITS], (unsigned long long)st.staged);
  base = __shfl(base, 0, 64);
  for (int j = lane; j < st.staged; j += 64) {
    unsigned long long slot = base + j;
    if (slot < (unsigned long long)w.cap.hits) {
      w.hit_depth[slot] = st.depth[j];
      w.hit_src[slot] = st.src[j];
      w.hit_group[slot] = st.group[j];
    } else {
      set_error(w.status, ERR_HITS);
    }
  }
  __builtin_amdgcn_wave_barrier();
  st.staged = 0;
}
// Called by every lane of the wave together (wave-uniform control flow).
__device__ __forceinline__ void stage_hits(const Ws &w, HitStage &st, bool hit, int leaf, int group, double2 x) {
  uint64_t m = __ballot(hit);
  if (!m) return;
  if (hit) {
    int pos = st.staged + __popcll(m & lanemask_lt());
    st.depth[pos] = value(w.leaf_score[leaf], x);
    st.src[pos] = w.source_tree.ids[leaf];
    st.group[pos] = group;
  }
  st.staged += __popcll(m);
  if (st.staged >= 64) flush_hits(w, st);
}

__device__ __forceinline__ void ray_box(double2 x, float4 &lo, float4 &hi) {
  lo = make_float4((float)x.x - 2.0e-7f, (float)x.y - 2.0e-7f, -1.0e30f, 0.f);
  hi = make_float4((float)x.x + 2.0e-7f, (float)x.y + 2.0e-7f, 1.0e30f, 0.f);
}

// The depth-first walk of a ray, for the rays whose breadth-first level lists
// outgrow LDS: K lanes (a whole wave) split the source hierarchy into its K
// subtrees at depth log2 K and each walks its own, so a ray's candidates are
// visited in parallel and lanes whose subtree box misses the ray stop after
// one node. A lane that reaches HANDOFF candidates hands the unvisited rest of
// its subtree to the subtree-parallel kernel (or keeps walking if that queue
// is full). The reference counts a group as heavy once it passes 32
// candidates (event_kernels.py:222-245): that counter keeps its meaning.
constexpr int HANDOFF = 32;

// Lanes per ray of the breadth-first walk (k_ray_bfs): 16, or a whole wave
// when there are at most RAY_FEW rays and 64 lanes per ray stay within
// `target` lanes in flight. Groups of 16 measured best: the 717,740-face
// shells (79,100 rays) 0.38 ms against 0.46 in groups of 64, 0.39 in groups of
// 8 and 0.80 with one lane per ray walking depth first (rays in Morton order);
// the 65,536-face torus knot (4,124 rays) 0.048 ms against 0.075 in groups of
// 64 and 0.071 in groups of 8; the 256 merged icospheres (1.2 M rays) 2.32 ms
// against 2.40 depth first. A Morton order of the rays gained no more than
// its sort cost.
constexpr int RAY_FEW = 2048;
constexpr int RAY_GROUP = 16;
__device__ __forceinline__ int ray_group(int rays, int target) {
  return rays <= RAY_FEW && (long)64 * rays <= (long)target ? 64 : RAY_GROUP;
}
// Lanes of the depth-first walk of a ray that k_ray_bfs hands over: a whole
// wave, one leaf subtree each.
__device__ __forceinline__ int walk_lanes(const Tree &tree) { return tree.leaves < 64 ? tree.leaves : 64; }

// Sum over the aligned group of K lanes (every lane of the wave active).
__device__ __forceinline__ long long group_sum(long long v, int k) {
  for (int off = k >> 1; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Ray leaders (the patches whose group they lead), compacted in patch order.
// A block claims slots for the leaders among 1,024 patches with one atomic
// (returning atomics on one word serialise at about 15 ns each: one per wave
// cost the shells 0.19 ms).
constexpr int KEY_ITEMS = 4;
__global__ __launch_bounds__(256) void k_ray_leaders(Ws w) {
  __shared__ int block_base, wave_total[4];
  const bool ok = w.status[ST_ERR] == 0;
  const int P = npolys(w), wave = threadIdx.x >> 6;
  for (long base = (long)blockIdx.x * 256 * KEY_ITEMS; base < P; base += (long)gridDim.x * 256 * KEY_ITEMS) {
    bool lead[KEY_ITEMS];
    int mine = 0;
#pragma unroll
    for (int k = 0; k < KEY_ITEMS; ++k) {
      long i = base + threadIdx.x + 256L * k;
      lead[k] = ok && i < P && w.groups[i] == (int)i;
      mine += lead[k];
    }
    int in_wave;
    int before = wave_prefix(mine, in_wave);
    if ((threadIdx.x & 63) == 0) wave_total[wave] = in_wave;
    __syncthreads();
    int all = 0;
    for (int q = 0; q < 4; ++q) {
      if (q < wave) before += wave_total[q];
      all += wave_total[q];
    }
    if (threadIdx.x == 0) block_base = all ? atomicAdd(w.control + C_LEADERS, all) : 0;
    __syncthreads();
    int slot = block_base + before;
#pragma unroll
    for (int k = 0; k < KEY_ITEMS; ++k) {
      if (!lead[k]) continue;
      w.ray_ids[slot++] = (int)(base + threadIdx.x + 256L * k);
    }
    __syncthreads();
  }
}

// Counters of a ray stage, summed per wave at the end of the kernel.
struct RayCounts {
  long long groups = 0, visits = 0, heavy = 0;
  int filtered = 0, fallbacks = 0;
};

// The rays order[base + lane / K] (per wave 64 / K of them), K lanes each:
// every lane walks its subtree at depth log2 K and hands the rest of it to
// k_heavy_hits after HANDOFF candidates. Called by every lane of the wave
// together.
__device__ void walk_ray_lanes(const Ws &w, HitStage &stage, const Tree &tree, int K, long base, int n,
                               const int *order, RayCounts &rc) {
  const int lane = threadIdx.x & 63, sub = lane & (K - 1);
  long r = base + lane / K;
  bool active = r < n;
  int i = active ? order[r] : 0;
  double2 x = active ? w.samples[i] : make_double2(0, 0);
  float2 x32 = make_float2((float)x.x, (float)x.y);
  float4 lo, hi;
  ray_box(x, lo, hi);
  const int root = K + sub;
  int node = active ? root : 0;  // binary continuation, for hand-offs
  const int end = successor(root);
  TrailWalk walk;
  walk.trail = 0;
  walk.bucket = 0;
  if (active) walk.start(tree, root, w.bucket_log, lo, hi);
  int count = 0, visits = 0;
  bool handed = false, queue_full = false;
  bool live = node > 0 && node != end;
  while (__ballot(live)) {  // wave-uniform: the hit stage needs every lane
    bool hit = false;
    int leaf = 0;
    if (live && visits >= HANDOFF && !queue_full) {
      // the unvisited suffix of this subtree starts at the first leaf
      // below the cursor
      int first = node;
      while (first < tree.leaves) first *= 2;
      int slot = wave_claim32(w.control + C_HEAVY, true);
      if (slot < w.cap.regions) {
        w.heavy[slot] = make_int4(i, first, root, 0);
        handed = true;
        live = false;
      } else {
        queue_full = true;
      }
    }
    if (live) {
      int leaf_node = walk.next_node(tree, lo, hi);
      if (leaf_node < 0) {
        live = false;
      } else {
        node = successor(leaf_node);
        leaf = leaf_node - tree.leaves;
        visits += 1;
        hit = contains_any(w, leaf, x, x32, rc.filtered, rc.fallbacks);
        count += hit;
        live = node > 0 && node != end;
      }
    }
    stage_hits(w, stage, hit, leaf, i, x);
  }
  long long total = group_sum(count, K);
  long long tested = group_sum(visits, K);
  long long any_handed = group_sum(handed ? 1 : 0, K);
  if (active && sub == 0) {
    w.group_counts[i] = (int)total;
    rc.groups += 1;
    rc.visits += tested;
    rc.heavy += (any_handed > 0 || tested > 32) ? 1 : 0;
  }
}

__device__ void add_ray_counts(const Ws &w, const RayCounts &rc) {
  metric_add(w, M_RAY_GROUPS, rc.groups);
  metric_add(w, M_RAY_TESTS, rc.visits);
  metric_add(w, M_HEAVY, rc.heavy);
  if (w.filter32) {
    metric_add(w, M_FILTER_TESTS, rc.filtered);
    metric_add(w, M_FALLBACKS, rc.fallbacks);
  }
}

// ray_hits (event_kernels.py:196-245) for the rays k_ray_bfs, launched
// before, listed because their level lists outgrew its LDS: a wave each, one
// leaf subtree per lane.
__global__ __launch_bounds__(256) void k_ray_hits(Ws w) {
  if (w.status[ST_ERR] != 0) return;
  const Tree tree = w.source_tree.tree();
  const int K = walk_lanes(tree);
  const int n = load_i(w.control + C_RAYFALL);
  const int *order = w.ray_ids_alt;
  HitStage stage = hit_stage();
  const int per_wave = 64 / K;
  const long waves = (gridDim.x * blockDim.x) >> 6;
  RayCounts rc;
  for (long base = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * per_wave; base < n; base += waves * per_wave)
    walk_ray_lanes(w, stage, tree, K, base, n, order, rc);
  flush_hits(w, stage);
  add_ray_counts(w, rc);
}

// ray_hits (event_kernels.py:196-245), leaders only: the first C_LEADERS
// entries of ray_ids, breadth first, a group of G lanes (16, or 64) each: each level
// of the source hierarchy's four-wide walk is expanded by the group at once (a
// lane per grandchild box, the overlapping ones appended to the next level's
// list in LDS), then the group's lanes test the leaves of the overlapping
// buckets. A ray costs one dependent read per level plus one per G leaf boxes,
// where a lane walking its own subtree reads once per step of its longest
// walk. The candidates are the same leaves (those whose boxes meet the ray
// box), so the hits and counts are too; the hit order is restored by the sort.
// A ray whose list outgrows its share of RAY_FRONTIER is listed (its patch
// index, in ray_ids_alt) for k_ray_hits. Loops run while any group of the wave has work (the ballots and
// the hit stage need every lane).
constexpr int RAY_FRONTIER = 512;
constexpr int RAY_BFS_BLOCKS = 2048;  // 8,192 waves

__global__ __launch_bounds__(256) void k_ray_bfs(Ws w) {
  __shared__ int frontier_lds[4][2][RAY_FRONTIER];
  if (w.status[ST_ERR] != 0) return;
  const int n = load_i(w.control + C_LEADERS);
  const Tree tree = w.source_tree.tree();
  const int G = ray_group(n, w.ray_lanes_target);
  HitStage stage = hit_stage();
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int per_wave = 64 / G, g = lane / G, gl = lane % G, cap = RAY_FRONTIER / per_wave;
  const uint64_t gmask = G == 64 ? ~0ull : ((1ull << G) - 1) << (G * g);
  const uint64_t below = lanemask_lt() & gmask;
  TrailWalk plan;
  plan.plan(tree, 1, w.bucket_log);
  const int lvls = plan.lvls, blog = plan.blog;
  const long waves = (gridDim.x * blockDim.x) >> 6;
  RayCounts rc;
  for (long r0 = ((blockIdx.x * blockDim.x + threadIdx.x) >> 6) * per_wave; r0 < n; r0 += waves * per_wave) {
    const long r = r0 + g;
    const bool active = r < n;  // group-uniform, as every value below
    const int i = active ? w.ray_ids[r] : 0;
    const double2 x = active ? w.samples[i] : make_double2(0, 0);
    const float2 x32 = make_float2((float)x.x, (float)x.y);
    float4 lo, hi;
    ray_box(x, lo, hi);
    int sel = 0;
    if (gl == 0) frontier_lds[wave][0][g * cap] = 1;
    int count = active ? 1 : 0;
    bool overflow = false;
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    for (int k = 0; k < lvls; ++k) {
      const int *cur = &frontier_lds[wave][sel][g * cap];
      int *next = &frontier_lds[wave][sel ^ 1][g * cap];
      int made = 0;
      for (int c0 = 0; __ballot(c0 < count && !overflow); c0 += G / 4) {
        const int idx = c0 + (gl >> 2);
        bool keep = false;
        int child = 0;
        if (idx < count && !overflow) {
          child = 4 * cur[idx] + (lane & 3);
          keep = overlaps(tree.nodes[child], lo, hi);
        }
        const uint64_t m = __ballot(keep) & gmask;
        const int pos = made + __popcll(m & below);
        if (keep && pos < cap) next[pos] = child;
        made += __popcll(m);
      }
      overflow = overflow || made > cap;
      count = made;
      sel ^= 1;
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
    }
    // the buckets' leaves, G (bucket, leaf) pairs at a time
    const int *cur = &frontier_lds[wave][sel][g * cap];
    const int pairs = overflow ? 0 : count << blog;
    int hits = 0, visits = 0;
    for (int q0 = 0; __ballot(q0 < pairs); q0 += G) {
      const int q = q0 + gl;
      bool hit = false;
      int leaf = 0;
      if (q < pairs) {
        const int leaf_node = (cur[q >> blog] << blog) + (q & ((1 << blog) - 1));
        if (overlaps(tree.nodes[leaf_node], lo, hi)) {
          leaf = leaf_node - tree.leaves;
          visits += 1;
          hit = contains_any(w, leaf, x, x32, rc.filtered, rc.fallbacks);
          hits += hit;
        }
      }
      stage_hits(w, stage, hit, leaf, i, x);
    }
    const long long total = group_sum(hits, G), tested = group_sum(visits, G);
    if (active && gl == 0) {
      if (overflow) {
        w.ray_ids_alt[atomicAdd(w.control + C_RAYFALL, 1)] = i;
      } else {
        w.group_counts[i] = (int)total;
        rc.groups += 1;
        rc.visits += tested;
        rc.heavy += tested > 32 ? 1 : 0;
      }
    }
  }
  flush_hits(w, stage);
  add_ray_counts(w, rc);
}

// heavy_ray_hits (event_kernels.py:248-298): each handed-off job (ray, first
// unvisited leaf, lane subtree) is split into aligned parts of its subtree,
// sized so jobs * parts stays near 65,536; parts before the cursor are skipped
// and the part holding it starts there.
__global__ __launch_bounds__(256) void k_heavy_hits(Ws w, int workers) {
  bool ok = w.status[ST_ERR] == 0;
  HitStage stage = hit_stage();
  Tree tree = w.source_tree.tree();
  int jobs = ok ? load_i(w.control + C_HEAVY) : 0;
  if (jobs > w.cap.regions) jobs = w.cap.regions;
  // leaves per lane subtree, with the lanes per ray that k_ray_hits chose
  const int W = tree.leaves / walk_lanes(tree);
  // a job split into more parts re-enters that many subtrees from their
  // roots: split only while the split jobs still fit the heavy lane budget
  int parts = W;
  while (parts > 1 && (long long)parts * jobs > w.heavy_lanes) parts /= 2;
  const int width = W / parts;
  long total = (long)jobs * parts;
  long long visits = 0;
  int filtered = 0, fallbacks = 0;
  for (long base = (long)blockIdx.x * blockDim.x; base < total; base += workers) {
    long task = base + threadIdx.x;
    int i = 0, part_root = 0, from = 0;
    if (task < total) {
      int4 job = w.heavy[task / parts];
      int part = (int)(task % parts);
      int root = job.z * parts + part;
      int first = root * width;  // first leaf node of the part
      if (job.y < first + width) {
        part_root = root;
        from = job.y > first ? job.y : 0;  // the leaf the job continues from, inside this part
        i = job.x;
      }
    }
    double2 x = part_root > 0 ? w.samples[i] : make_double2(0, 0);
    float4 lo, hi;
    ray_box(x, lo, hi);
    TrailWalk walk;
    walk.trail = 0;
    walk.bucket = 0;
    if (part_root > 0) {
      if (from)
        walk.start_at(tree, part_root, from, w.bucket_log, lo, hi);
      else
        walk.start(tree, part_root, w.bucket_log, lo, hi);
    }
    int count = 0;
    float2 x32 = make_float2((float)x.x, (float)x.y);
    bool live = part_root > 0;
    while (__ballot(live)) {  // wave-uniform: the hit stage needs every lane
      bool hit = false;
      int leaf = 0;
      if (live) {
        int leaf_node = walk.next_node(tree, lo, hi);
        if (leaf_node < 0) {
          live = false;
        } else {
          leaf = leaf_node - tree.leaves;
          visits += 1;
          hit = contains_any(w, leaf, x, x32, filtered, fallbacks);
          count += hit;
        }
      }
      stage_hits(w, stage, hit, leaf, i, x);
    }
    // Consecutive tasks are parts of one job, so the lanes of a wave usually
    // add to one ray's count: those that share the first contributing lane's
    // ray add through it with one atomic (the 65,536-layer stack's single ray
    // took 65,536 atomics on one address), the others add their own.
    const int lane = threadIdx.x & 63;
    uint64_t adding = __ballot(count != 0);
    if (adding) {
      int leader = __ffsll((unsigned long long)adding) - 1;
      int ray = __shfl(i, leader);
      bool same = count != 0 && i == ray;
      long long shared = wave_sum(same ? count : 0);
      if (lane == leader) atomicAdd(&w.group_counts[ray], (int)shared);
      if (count != 0 && !same) atomicAdd(&w.group_counts[i], count);
    }
  }
  flush_hits(w, stage);
  metric_add(w, M_RAY_TESTS, visits);
  if (w.filter32) {
    metric_add(w, M_FILTER_TESTS, filtered);
    metric_add(w, M_FALLBACKS, fallbacks);
  }
}

// ---------------------------------------------------------------------------
// Hit order (group, depth, source). Each group's hits form one contiguous
// segment at its scanned offset; segments are sorted independently: up to 64
// hits by one wave (bitonic network across lanes, in registers), up to 2048 by
// one workgroup (bitonic in LDS), larger ones by the workgroups of
// k_sort_huge together (LDS chunks, then merge passes).
constexpr int SMALL_SEGMENT = 64, MEDIUM_SEGMENT = 2048;

__device__ __forceinline__ bool hit_less(uint64_t ka, int sa, uint64_t kb, int sb) {
  return ka < kb || (ka == kb && sa < sb);
}

// Hits into their group's segment. The order inside a segment does not
// matter (the segment sorts are total). A ray's hits sit next to each other,
// so the lanes of a wave that share the first lane's group take their slots
// with one atomic (one group of 65,536 hits would otherwise serialise 65,536
// atomics on one counter); the other lanes, usually of small groups, take
// theirs one by one.
__global__ void k_hit_scatter(Ws w) {
  // the ray stages' hit total against the pool (check_hits,
  // event_kernels.py:300-309); only stored hits are scattered
  const long long total = w.metrics[M_RAY_HITS];
  if (blockIdx.x == 0 && threadIdx.x == 0 && total > w.cap.hits) set_error(w.status, ERR_HITS);
  if (w.status[ST_ERR] != 0) return;
  const int n = (int)(total < w.cap.hits ? total : w.cap.hits);
  const int lane = threadIdx.x & 63;
  const uint64_t below = (1ull << lane) - 1;
  WAVE_LOOP(base, n) {
    int h = base + threadIdx.x;
    bool valid = h < n;
    int g = valid ? w.hit_group[h] : -1;
    int leader = __ffsll((unsigned long long)__ballot(valid)) - 1;  // lane 0 unless the wave is empty
    int lg = __shfl(g, leader < 0 ? 0 : leader, 64);
    bool shared = valid && g == lg;
    uint64_t peers = __ballot(shared);
    int first = 0;
    if (lane == leader && peers) first = atomicAdd(&w.group_fill[lg], __popcll(peers));
    first = __shfl(first, leader < 0 ? 0 : leader, 64);
    int pos = 0;
    if (shared)
      pos = w.group_offsets[lg] + first + __popcll(peers & below);
    else if (valid)
      pos = w.group_offsets[g] + atomicAdd(&w.group_fill[g], 1);
    if (valid) w.hids_alt[pos] = h;
  }
}

__device__ void sort_small_one(const Ws &w, int g);

// One wave per ray group (leaders are the first C_LEADERS sorted rays).
__global__ __launch_bounds__(256) void k_sort_small(Ws w) {
  if (w.status[ST_ERR] != 0) return;
  int n = load_i(w.control + C_LEADERS);
  for (int t = blockIdx.x * 4 + (threadIdx.x >> 6); t < n; t += gridDim.x * 4) sort_small_one(w, w.ray_ids[t]);
}

__device__ void sort_small_one(const Ws &w, int g) {
  int lane = threadIdx.x & 63;
  int c = w.group_counts[g];
  if (c == 0) return;
  if (c > SMALL_SEGMENT) {
    if (lane == 0) {
      if (c <= MEDIUM_SEGMENT) {
        int k = atomicAdd(w.control + C_MEDIUM, 1);
        w.medium[k] = g;
      } else {
        // huge groups fill the same list from its end (a group is one or
        // the other, so the two ends never meet)
        int k = atomicAdd(w.control + C_HUGE, 1);
        w.medium[w.cap.regions - 1 - k] = g;
      }
    }
    return;
  }
  int off = w.group_offsets[g];
  uint64_t key = ~0ull;
  int src = 0x7fffffff, h = -1;
  if (lane < c) {
    h = w.hids_alt[off + lane];
    key = dkey(w.hit_depth[h]);
    src = w.hit_src[h];
  }
  for (int k = 2; k <= 64; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      uint64_t ok = __shfl_xor(key, j, 64);
      int os = __shfl_xor(src, j, 64), oh = __shfl_xor(h, j, 64);
      bool up = (lane & k) == 0;
      bool lower = (lane & j) == 0;
      bool mine_less = hit_less(key, src, ok, os);
      bool keep = (lower == up) ? mine_less : !mine_less;
      if (!keep) {
        key = ok;
        src = os;
        h = oh;
      }
    }
  }
  if (lane < c) w.hids[off + lane] = h;
}

__global__ __launch_bounds__(256) void k_sort_medium(Ws w) {
  __shared__ unsigned long long kk[MEDIUM_SEGMENT];
  __shared__ int ss[MEDIUM_SEGMENT], hh[MEDIUM_SEGMENT];
  if (w.status[ST_ERR] != 0) return;
  int count = load_i(w.control + C_MEDIUM);
  for (int item = blockIdx.x; item < count; item += gridDim.x) {
    int g = w.medium[item];
    int c = w.group_counts[g], off = w.group_offsets[g];
    int n2 = 1;
    while (n2 < c) n2 <<= 1;
    for (int i = threadIdx.x; i < n2; i += blockDim.x) {
      if (i < c) {
        int h = w.hids_alt[off + i];
        kk[i] = dkey(w.hit_depth[h]);
        ss[i] = w.hit_src[h];
        hh[i] = h;
      } else {
        kk[i] = ~0ull;
        ss[i] = 0x7fffffff;
        hh[i] = -1;
      }
    }
    __syncthreads();
    for (int k = 2; k <= n2; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = threadIdx.x; i < n2; i += blockDim.x) {
          int p = i ^ j;
          if (p > i) {
            bool up = (i & k) == 0;
            bool greater = hit_less(kk[p], ss[p], kk[i], ss[i]);
            if (greater == up) {
              unsigned long long tk = kk[i];
              kk[i] = kk[p];
              kk[p] = tk;
              int t = ss[i];
              ss[i] = ss[p];
              ss[p] = t;
              t = hh[i];
              hh[i] = hh[p];
              hh[p] = t;
            }
          }
        }
        __syncthreads();
      }
    }
    for (int i = threadIdx.x; i < c; i += blockDim.x) w.hids[off + i] = hh[i];
    __syncthreads();
  }
}

// Groups above MEDIUM_SEGMENT hits (e.g. 65,536 identical footprints) are
// sorted by the HUGE_BLOCKS workgroups of one launch together: chunks of
// MEDIUM_SEGMENT bitonic-sorted in LDS (spread over the blocks), then merge
// passes of doubling run width, each grid thread producing a contiguous span
// of outputs from its merge-path split, with a grid barrier after every
// phase. A group's passes alternate between (hkeys, hids) and (hkeys_alt,
// hids_alt); an odd count leaves its order in hids_alt, copied back at the
// end. Sources are unique within a group, so the order (depth key, source) is
// total and equals the small and medium sorts'. With no huge group every
// block returns at once.
constexpr int HUGE_THREADS = 1024, HUGE_BLOCKS = 64;

__device__ __forceinline__ bool hit_before(const Ws &w, uint64_t ka, int ha, uint64_t kb, int hb) {
  return ka < kb || (ka == kb && w.hit_src[ha] < w.hit_src[hb]);
}

// Grid barrier of k_sort_huge: its HUGE_BLOCKS blocks are all resident (the
// launch is small and follows the previous kernel of its stream), every block
// passes the same barriers, and a bounded wait sets ERR_WAIT instead of
// hanging. Agent-scope fences publish the blocks' writes across the chip.
__device__ void huge_barrier(const Ws &w, int &generation) {
  __syncthreads();
  if (threadIdx.x == 0) {
    generation += 1;
    __threadfence();
    atomicAdd(w.control + C_HUGE_SYNC, 1);
    const int target = generation * (int)gridDim.x;
    int spins = 0;
    while (__hip_atomic_load(w.control + C_HUGE_SYNC, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (++spins > (1 << 20)) {
        set_error(w.status, ERR_WAIT);
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    __threadfence();
  }
  __syncthreads();
}

// Phases of the huge-group sort: the LDS chunk sorts, merge pass p of run
// width MEDIUM_SEGMENT << p, and the copy-back of groups with an odd number
// of passes. huge_plan reads the groups' sizes.
__device__ __forceinline__ void huge_plan(const Ws &w, int count, long &tasks, long &largest) {
  tasks = 0, largest = 0;
  for (int item = 0; item < count; ++item) {
    int c = w.group_counts[w.medium[w.cap.regions - 1 - item]];
    tasks += (c + MEDIUM_SEGMENT - 1) / MEDIUM_SEGMENT;
    largest = max(largest, (long)c);
  }
}

__device__ void huge_chunks(const Ws &w, int count, long tasks, unsigned long long *kk, int *ss, int *hh) {
  for (long task = blockIdx.x; task < tasks; task += gridDim.x) {
    int item = 0;
    long first = 0;
    int g = 0, c = 0;
    for (;; ++item) {
      g = w.medium[w.cap.regions - 1 - item];
      c = w.group_counts[g];
      long chunks = (c + MEDIUM_SEGMENT - 1) / MEDIUM_SEGMENT;
      if (task < first + chunks) break;
      first += chunks;
    }
    const int off = w.group_offsets[g];
    const int base = (int)(task - first) * MEDIUM_SEGMENT;
    const int m = min(MEDIUM_SEGMENT, c - base);
    int n2 = 1;
    while (n2 < m) n2 <<= 1;
    for (int i = threadIdx.x; i < n2; i += blockDim.x) {
      if (i < m) {
        int h = w.hids_alt[off + base + i];
        kk[i] = dkey(w.hit_depth[h]);
        ss[i] = w.hit_src[h];
        hh[i] = h;
      } else {
        kk[i] = ~0ull;
        ss[i] = 0x7fffffff;
        hh[i] = -1;
      }
    }
    __syncthreads();
    for (int k = 2; k <= n2; k <<= 1) {
      for (int j = k >> 1; j > 0; j >>= 1) {
        for (int i = threadIdx.x; i < n2; i += blockDim.x) {
          int p = i ^ j;
          if (p > i) {
            bool up = (i & k) == 0;
            bool greater = hit_less(kk[p], ss[p], kk[i], ss[i]);
            if (greater == up) {
              unsigned long long tk = kk[i];
              kk[i] = kk[p];
              kk[p] = tk;
              int t = ss[i];
              ss[i] = ss[p];
              ss[p] = t;
              t = hh[i];
              hh[i] = hh[p];
              hh[p] = t;
            }
          }
        }
        __syncthreads();
      }
    }
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      w.hkeys[off + base + i] = kk[i];
      w.hids[off + base + i] = hh[i];
    }
    __syncthreads();
  }
}

__device__ void huge_merge(const Ws &w, int count, int pass, long width) {
  const long threads = (long)gridDim.x * blockDim.x;
  const long me = (long)blockIdx.x * blockDim.x + threadIdx.x;
  for (int item = 0; item < count; ++item) {
    const int g = w.medium[w.cap.regions - 1 - item];
    const long c = w.group_counts[g];
    if (c <= width) continue;
    const int off = w.group_offsets[g];
    const uint64_t *kin = (pass & 1 ? w.hkeys_alt : w.hkeys) + off;
    const int *iin = (pass & 1 ? w.hids_alt : w.hids) + off;
    uint64_t *kout = (pass & 1 ? w.hkeys : w.hkeys_alt) + off;
    int *iout = (pass & 1 ? w.hids : w.hids_alt) + off;
    const long per = (c + threads - 1) / threads;
    long k = me * per;
    const long k_end = min(c, k + per);
    while (k < k_end) {
      const long a = k / (2 * width) * (2 * width);
      const long mid_run = min(c, a + width), e = min(c, a + 2 * width);
      const long nl = mid_run - a, nr = e - mid_run;
      const uint64_t *lk = kin + a, *rk = kin + mid_run;
      const int *li = iin + a, *ri = iin + mid_run;
      // merge-path split of the first d outputs of this pair of runs
      const long d = k - a;
      long lo = max(0L, d - nr), hi = min(d, nl);
      while (lo < hi) {
        long m = (lo + hi) / 2;
        if (hit_before(w, lk[m], li[m], rk[d - m - 1], ri[d - m - 1]))
          lo = m + 1;
        else
          hi = m;
      }
      long i = lo, j = d - lo;
      const long stop = min(k_end, e);
      for (; k < stop; ++k) {
        bool left = j >= nr || (i < nl && hit_before(w, lk[i], li[i], rk[j], ri[j]));
        if (left) {
          kout[k] = lk[i];
          iout[k] = li[i];
          ++i;
        } else {
          kout[k] = rk[j];
          iout[k] = ri[j];
          ++j;
        }
      }
    }
  }
}

__device__ void huge_finish(const Ws &w, int count) {
  const long threads = (long)gridDim.x * blockDim.x;
  const long me = (long)blockIdx.x * blockDim.x + threadIdx.x;
  for (int item = 0; item < count; ++item) {
    const int g = w.medium[w.cap.regions - 1 - item];
    const long c = w.group_counts[g];
    int passes = 0;
    for (long width = MEDIUM_SEGMENT; width < c; width *= 2) ++passes;
    if (passes & 1) {
      const int off = w.group_offsets[g];
      for (long i = me; i < c; i += threads) w.hids[off + i] = w.hids_alt[off + i];
    }
  }
}

// One launch for every phase, with a grid barrier between phases: its
// HUGE_BLOCKS blocks are all resident when it runs alone after the previous
// kernel of its stream (the single-workspace graph).
__global__ __launch_bounds__(HUGE_THREADS) void k_sort_huge(Ws w) {
  __shared__ unsigned long long kk[MEDIUM_SEGMENT];
  __shared__ int ss[MEDIUM_SEGMENT], hh[MEDIUM_SEGMENT];
  if (w.status[ST_ERR] != 0) return;
  const int count = load_i(w.control + C_HUGE);
  if (count == 0) return;
  int generation = 0;
  long tasks, largest;
  huge_plan(w, count, tasks, largest);
  huge_chunks(w, count, tasks, kk, ss, hh);
  huge_barrier(w, generation);
  int pass = 0;
  for (long width = MEDIUM_SEGMENT; width < largest; width *= 2, ++pass) {
    huge_merge(w, count, pass, width);
    huge_barrier(w, generation);
  }
  huge_finish(w, count);
}

// One phase per launch, no grid barrier (graphs whose members run
// concurrently, where the blocks of one launch need not be co-resident):
// phase 0 the chunk sorts, phase 1 + p merge pass p (nothing to do once the
// runs cover the largest group), phase HUGE_FINISH the copy-back.
constexpr int HUGE_FINISH = 1 << 20;
__global__ __launch_bounds__(HUGE_THREADS) void k_sort_huge_phase(Ws w, int phase) {
  __shared__ unsigned long long kk[MEDIUM_SEGMENT];
  __shared__ int ss[MEDIUM_SEGMENT], hh[MEDIUM_SEGMENT];
  if (w.status[ST_ERR] != 0) return;
  const int count = load_i(w.control + C_HUGE);
  if (count == 0) return;
  long tasks, largest;
  huge_plan(w, count, tasks, largest);
  if (phase == 0) {
    huge_chunks(w, count, tasks, kk, ss, hh);
  } else if (phase == HUGE_FINISH) {
    huge_finish(w, count);
  } else {
    const int pass = phase - 1;
    const long width = (long)MEDIUM_SEGMENT << pass;
    if (width < largest) huge_merge(w, count, pass, width);
  }
}

__device__ int assign_one(const Ws &w, int i, int &filtered, int &fallbacks);

__global__ __launch_bounds__(256) void k_assign(Ws w) {
  int top = 0, filtered = 0, fallbacks = 0;
  if (w.status[ST_ERR] == 0) {
    ITEM_LOOP(i, npolys(w)) top = max(top, assign_one(w, i, filtered, fallbacks));
  }
  top = wave_max(top);
  // most waves see the layer count already at least their top: no atomic
  if ((threadIdx.x & 63) == 0 && top > load_i(w.status + ST_LAYERS)) atomicMax(w.status + ST_LAYERS, top);
  if (w.filter32) {
    metric_add(w, M_FILTER_DEPTH, filtered);
    metric_add(w, M_FALLBACK_DEPTH, fallbacks);
  }
}

__device__ int assign_one(const Ws &w, int i, int &filtered, int &fallbacks) {
  int source = w.out_src[w.reps[i]];
  double2 x = w.samples[i];
  double z = value(w.scores[source], x);
  int g = w.groups[i];
  int low = 0, high = w.group_counts[g];
  int base = w.group_offsets[g];
  float2 x32 = to32(x);
  float Az = 0.f, z32 = 0.f;
  if (w.filter32) z32 = value32(w.scores32[source], x32, Az);
  while (low < high) {
    int mid = (low + high) / 2;
    int h = w.hids[base + mid];
    int hs = w.hit_src[h];
    // binary32 depth order beyond its static bound; binary64 for the rest
    int decided = 0;  // -1 nearer, +1 farther, 0 undecided
    if (w.filter32 && hs == source) {
      decided = 1;  // the patch's own hit: equal depth, and the tie rule says not nearer
    } else if (w.filter32) {
      float Ah;
      float diff = value32(w.scores32[hs], x32, Ah) - z32;
      float tol = DEPTH_TOL32 * (Ah + Az) + 1.0e-37f;
      decided = diff < -tol ? -1 : (diff > tol ? 1 : 0);
      if (decided)
        filtered += 1;
      else
        fallbacks += 1;
    }
    bool nearer = decided < 0;
    if (decided == 0) {
      double hz = value(w.scores[hs], x);
      nearer = hz < z || (hz == z && hs < source);
    }
    if (nearer)
      low = mid + 1;
    else
      high = mid;
  }
  w.out_layer[i] = low;
  if (low >= w.limit) {
    if (w.all_layers)
      set_error(w.status, ERR_LAYERS);
    else
      w.out_n[i] = 0;
    return 0;
  }
  return low + 1;
}

static inline dim3 grid_for(long n, int block = 256) {
  long g = (n + block - 1) / block;
  return dim3((unsigned)(g < 1 ? 1 : g));
}

void enqueue_rank(const Ws &w, hipStream_t s, Timeline *t) {
  long qgrid = ((long)std::max(w.qhash_size, w.cap.regions) + 255) / 256;
  if (qgrid > 1024) qgrid = 1024;
  hipLaunchKernelGGL(k_rank_reset, dim3((unsigned)qgrid), dim3(256), 0, s, w);
  if (w.propagate_ranks || w.merge_regions) enqueue_connect(w, s);
  mark(t, s, "connectivity");
  hipLaunchKernelGGL(k_leaf_sources, grid_for(w.source_tree.size), dim3(256), 0, s, w);
  hipLaunchKernelGGL(k_samples, stride_grid(w.cap.regions), dim3(256), 0, s, w);
  hipLaunchKernelGGL(k_sample_insert, stride_grid(w.cap.regions), dim3(256), 0, s, w);
  hipLaunchKernelGGL(k_groups, stride_grid(w.cap.regions), dim3(256), 0, s, w);
  if (w.share_regions) {
    hipLaunchKernelGGL(k_certify, dim3(BEAM_CELLS / 256), dim3(256), 0, s, w);
    hipLaunchKernelGGL(k_certify_heavy, dim3(32), dim3(256), 0, s, w, 32 * 256);
    hipLaunchKernelGGL(k_beam_apply, stride_grid(w.cap.regions), dim3(256), 0, s, w);
  }
  {
    long kg = ((long)w.cap.regions + 256 * KEY_ITEMS - 1) / (256 * KEY_ITEMS);
    hipLaunchKernelGGL(k_ray_leaders, dim3((unsigned)std::max(1L, std::min(kg, 1024L))), dim3(256), 0, s, w);
  }
  mark(t, s, "samples_groups");
  hipLaunchKernelGGL(k_ray_bfs, dim3(std::min(stride_grid(w.cap.regions).x, (unsigned)RAY_BFS_BLOCKS)), dim3(256), 0,
                     s, w);
  hipLaunchKernelGGL(k_ray_hits, stride_grid(w.cap.regions), dim3(256), 0, s, w);
  mark(t, s, "k_ray_hits");
  {
    // enough workers for many handed-off rays (each continues in one lane
    // when there are more than 65,536 jobs); idle blocks return at once
    dim3 hg = stride_grid(w.cap.regions);
    hipLaunchKernelGGL(k_heavy_hits, hg, dim3(256), 0, s, w, (int)hg.x * 256);
  }
  mark(t, s, "k_heavy_hits");
  scan_exclusive(s, w.group_counts, w.group_offsets, w.status + ST_POLYS, 0, w.cap.regions,
                 nullptr, w.scan);
  hipLaunchKernelGGL(k_hit_scatter, stride_grid(w.cap.hits), dim3(256), 0, s, w);
  hipLaunchKernelGGL(k_sort_small, stride_grid(w.cap.regions, 4), dim3(256), 0, s, w);
  hipLaunchKernelGGL(k_sort_medium, dim3(1024), dim3(256), 0, s, w);
  if (!w.split_huge) {
    hipLaunchKernelGGL(k_sort_huge, dim3(HUGE_BLOCKS), dim3(HUGE_THREADS), 0, s, w);
  } else {
    // enough merge passes for a group as large as the hit pool
    int passes = 0;
    for (long width = MEDIUM_SEGMENT; width < w.cap.hits; width *= 2) ++passes;
    for (int phase = 0; phase <= passes; ++phase)
      hipLaunchKernelGGL(k_sort_huge_phase, dim3(HUGE_BLOCKS), dim3(HUGE_THREADS), 0, s, w, phase);
    hipLaunchKernelGGL(k_sort_huge_phase, dim3(HUGE_BLOCKS), dim3(HUGE_THREADS), 0, s, w, HUGE_FINISH);
  }
  mark(t, s, "hit_sort");
  hipLaunchKernelGGL(k_assign, stride_grid(w.cap.regions), dim3(256), 0, s, w);
  mark(t, s, "k_assign");
  if (w.merge_regions) {
    enqueue_merge(w, s);
    mark(t, s, "merge_regions");
  }
}

}  // namespace peel
