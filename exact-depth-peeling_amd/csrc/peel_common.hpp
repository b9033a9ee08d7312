// Shared device definitions for the gfx950 depth peeler.
//
// Layout conventions (HBM):
//  * Polygon rings are stored vertex-major, ring[j * capacity + slot], so the
//    64 lanes of a wave that work on 64 consecutive slots read vertex j of all
//    of them in one coalesced 1 KiB transaction (double2 per lane).
//  * Hierarchy nodes are 32-byte {float4 lo, float4 hi} records of an implicit
//    complete binary tree (node 1 is the root, leaves at [size, 2 size)),
//    matching the stackless traversal of warp_peel/kernels.py:24-64.
//  * Every pool is preallocated; counts live on the device and every kernel
//    reads its own trip count from device memory, so a captured graph needs no
//    host reads or reallocation between replays.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace peel {

constexpr int MAXV = 32;      // ring capacity (warp_peel/kernels.py:10)
constexpr int SRCV = 8;       // triangle + near plane + 4 window planes
constexpr double EPS = 1.0e-10;  // warp_peel/kernels.py:11
constexpr float PAD = 2.0e-7f;   // box padding (warp_peel/kernels.py:235-250)
constexpr int ROUTE_EVENTS = 0;  // event route (EventPeeler)
constexpr int ROUTE_LAYERED = 1;  // per-layer clipping route (DevicePeeler)
constexpr int BEAM_SIDE = 32;    // window cells per axis for shared rays (warp_peel/beams.py:18-22)
constexpr int BEAM_CELLS = BEAM_SIDE * BEAM_SIDE;

// status[16] slots and error bits: warp_peel/kernels.py:17-21, events.py:811-814.
enum Status : int {
  ST_CUR = 0,
  ST_NEXT = 1,
  ST_POLYS = 2,
  ST_TRIS = 3,
  ST_LAYERS = 4,
  ST_PASS = 5,
  ST_BVERTS = 6,
  ST_TIES = 8,
  ST_ERR = 9,
  ST_MAXH = 10,
  ST_PEAK = 12,
};
enum Err : int {
  ERR_POLY = 1,
  ERR_PIECES = 2,
  ERR_RESIDUAL = 4,
  ERR_OUTPUT = 8,
  ERR_VERTS = 16,
  ERR_TRIS = 32,
  ERR_LAYERS = 64,
  ERR_EDGE = 128,
  ERR_INPUT = 256,
  ERR_HITS = 1024,
  ERR_ROUNDS = 2048,
  ERR_WAIT = 4096,  // a bounded device wait ran out (extension; never expected)
};
// metrics[]: the 19 counters of warp_peel/events.py:819-843, then this build's own.
enum Metric : int {
  M_VISITS = 0,
  M_CONTOUR_TESTS = 1,
  M_SPLITS = 2,
  M_RAY_GROUPS = 3,
  M_RAY_TESTS = 4,
  M_RAY_HITS = 5,
  M_ROUNDS = 6,
  M_UNIQUE = 7,
  M_HEAVY = 8,
  M_BEAM_ATTEMPTS = 9,
  M_BEAMS = 10,
  M_BEAM_TESTS = 11,
  M_HEAVY_BEAMS = 12,
  M_COMPONENTS = 13,
  M_CONNECTIONS = 14,
  M_CONNECTION_TESTS = 15,
  M_MERGED = 16,
  M_MATCHED = 17,
  M_HASH_MISSES = 18,
  M_FILTER_TESTS = 19,   // ray containment edges decided by the binary32 filter
  M_FALLBACKS = 20,      // ray containment edges that needed binary64
  M_SLOTS = 21,          // frontier slots used (seeds + 2 per split)
  M_EDGE_MAX = 22,       // most noding candidates on one polygon edge
  M_UNIQUE_BVERTS = 23,  // distinct (layer, point) boundary vertices
  M_FILTER_STRADDLE = 24,    // subdivision straddle signs decided in binary32
  M_FALLBACK_STRADDLE = 25,  // ... evaluated in binary64
  M_FILTER_FRONT = 26,       // front_segment edge tests decided in binary32
  M_FALLBACK_FRONT = 27,     // ... evaluated in binary64
  M_FILTER_DEPTH = 28,       // rank depth comparisons decided in binary32
  M_FALLBACK_DEPTH = 29,     // ... evaluated in binary64
  M_COUNT = 30,
};

// Control words (device) used by persistent kernels and on-device counts.
// Contended words sit on their own 128-byte lines.
enum Control : int {
  C_ALLOC = 0,      // next free frontier slot
  C_HEAD = 32,      // next seed of the subdivision
  C_HEAVY = 96,     // heavy ray jobs
  C_NBV = 129,      // boundary vertices (scan total)
  C_MEDIUM = 131,   // groups sorted by a workgroup
  C_HUGE = 132,     // groups too large for one LDS sort (listed from the end of `medium`)
  C_NSEG = 133,     // unique contour segments
  C_INTERNAL = 135, // any internal (non-contour) mesh edge
  C_CONNECT = 136,  // connectivity pass enabled for this run
  C_PEDGES = 137,   // patch edges (scan total)
  C_CROWDED = 138,  // edges with more noding candidates than the register path
  C_BEAM_JOBS = 139, // cells whose certification continues in subtree workers
  C_LOOP = 140,      // layered route: another peeling pass is due
  C_LEADERS = 141,   // ray-group leaders (they sort first in the ray order)
  C_GRID = 142,      // cells per axis of the boundary-vertex grid
  C_CELLS = 146,     // live cells of that grid (layers * G^2)
  C_LIVE = 143,      // patches left after convex coalescing
  C_SORTN = 144,     // layered route: residual pieces to order (0 once the loop ends)
  C_HUGE_SYNC = 160, // arrivals at the grid barriers of the huge-group sort
  C_RAYFALL = 224,   // rays k_ray_bfs leaves to k_ray_hits (their level lists outgrew its LDS)
  C_COUNT = 256,
};

// Per-wave partial counters of the subdivision kernel, reduced by the next
// kernel: sums (visits, contour tests, splits, the four binary32 filter
// counts), then maxima (rounds, the peak of a block's live tasks).
constexpr int SUB_SUMS = 7;
// rows of metric partials (a power of two) and their stride in words (two
// 128-byte lines, at least M_COUNT)
constexpr int METRIC_ROWS = 64;
constexpr int METRIC_STRIDE = 32;
constexpr int SUB_COUNTERS = SUB_SUMS + 2;
constexpr int SUB_PARTIAL_WAVES = 8192;

struct __align__(16) Box {
  float4 lo, hi;
};

struct Tree {
  const Box *nodes;  // [2 * leaves]
  const int *ids;    // [leaves] sorted leaf ids
  int leaves;        // power of two
};

// ---------------------------------------------------------------------------
// small helpers
__device__ __forceinline__ double cross3(double2 a, double2 b, double2 c) {
  double ux = b.x - a.x, uy = b.y - a.y, vx = c.x - a.x, vy = c.y - a.y;
  return ux * vy - uy * vx;
}
__device__ __forceinline__ double value(double4 s, double2 x) { return s.x * x.x + s.y * x.y + s.z; }
__device__ __forceinline__ bool lexless(double2 a, double2 b) {
  return a.x < b.x || (a.x == b.x && a.y < b.y);
}
__device__ __forceinline__ bool same(double2 a, double2 b) { return a.x == b.x && a.y == b.y; }
__device__ __forceinline__ double dlength(double2 d) { return sqrt(d.x * d.x + d.y * d.y); }

__device__ __forceinline__ void set_error(int *status, int bit) { atomicOr(status + ST_ERR, bit); }
__device__ __forceinline__ int load_i(const int *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void acquire_fence() { __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent"); }
// Release at agent scope followed by an explicit wait for every outstanding
// memory operation, the L2 writeback of the release included. The compiler
// drops its own wait after the writeback when no store is pending at that
// point (an earlier wait has drained them), and a flag stored next could then
// reach a consumer on another XCD before the data it publishes.
__device__ __forceinline__ void release_fence() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
  __builtin_amdgcn_s_waitcnt(0);
}
__device__ __forceinline__ int load_acquire(const int *p) {
  return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void store_release(int *p, int v) {
  __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ long long load_ll(const long long *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Order-preserving map of binary64 to unsigned 64-bit keys; -0.0 and +0.0 map
// to the same key (sorting must agree with ==, warp_peel tests/gpu
// test_peeler.py:305-319).
__device__ __forceinline__ uint64_t dkey(double x) {
  if (x == 0.0) x = 0.0;
  uint64_t b = (uint64_t)__double_as_longlong(x);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double unkey(uint64_t k) {
  uint64_t b = (k >> 63) ? (k & 0x7fffffffffffffffull) : ~k;
  return __longlong_as_double((long long)b);
}

// Atomic min on binary64 through the order-preserving key (values are finite).
// The word only decreases, so an update that cannot lower the value it reads
// is skipped: many copies of one key (65,536 identical footprints share three
// segments) then cost a load each instead of an atomic on one address.
__device__ __forceinline__ void atomic_min_key(unsigned long long *p, double x) {
  unsigned long long k = (unsigned long long)dkey(x);
  if (__hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > k) atomicMin(p, k);
}

// Static forward error bound of a binary32 orientation test
//   s32 = (b.x - a.x) * (x.y - a.y) - (b.y - a.y) * (x.x - a.x)
// evaluated on binary64 inputs rounded to nearest binary32, when every
// coordinate is at most M in magnitude and every difference taken is at most
// D. With u = 2^-24, each rounded difference is off by at most 2uM + uD, each
// product by (2uM + uD) 2D + uD^2, and the final subtraction adds u|s32|, so
// |s32 - s| <= 8uMD + 8uD^2 to first order (s: the exact value). The bound
// doubles that and adds 1e-37 for flushed subnormal products; the binary64
// evaluation of the same expression is within ~4 2^-53 D^2, far inside it. A
// filter value beyond the bound therefore has the sign, and a nonzero value,
// that the binary64 predicate computes. Rounded up to binary32.
__device__ __forceinline__ float orientation_tol32(double D, double M) {
  const double u = 5.9604644775390625e-8;  // 2^-24
  double t = 16.0 * u * D * (M + D) + 1.0e-37;
  return __double2float_ru(t);
}

// The same orientation in binary32 on rounded inputs (no contraction: every
// HIP unit is built with -ffp-contract=off).
__device__ __forceinline__ float cross32(float2 a, float2 b, float2 c) {
  return (b.x - a.x) * (c.y - a.y) - (b.y - a.y) * (c.x - a.x);
}
__device__ __forceinline__ float2 to32(double2 x) { return make_float2((float)x.x, (float)x.y); }

// Bounds of a subdivision candidate test: every point it touches (the target
// ring inside its box [lx, hx] x [ly, hy], the segment a-b) lies in the union
// box, so D is that box's larger side and M its largest coordinate magnitude.
__device__ __forceinline__ float candidate_tol32(double lx, double ly, double hx, double hy, double2 a,
                                                 double2 b) {
  double x0 = fmin(lx, fmin(a.x, b.x)), x1 = fmax(hx, fmax(a.x, b.x));
  double y0 = fmin(ly, fmin(a.y, b.y)), y1 = fmax(hy, fmax(a.y, b.y));
  double D = fmax(x1 - x0, y1 - y0);
  double M = fmax(fmax(fabs(x0), fabs(x1)), fmax(fabs(y0), fabs(y1)));
  return orientation_tol32(D, M);
}

// Depth of plane s at x, s.x x.x + s.y x.y + s.z, in binary32 on rounded
// inputs, with A = |s.x||x.x| + |s.y||x.y| + |s.z|: each of the five
// roundings (two inputs and the product of each term, the two sums) is at
// most 2^-24 relative, so the value is within 5.01 2^-24 A of the exact depth,
// and the binary64 value within 3.4e-16 A. Two depths whose binary32
// difference exceeds DEPTH_TOL32 (A1 + A2) (three times the sum of both
// errors, which also covers the difference's own rounding and the rounding
// of A) therefore compare in binary64 as they do in binary32, and unequal.
constexpr float DEPTH_TOL32 = 16.0f * 5.9604644775390625e-8f;
__device__ __forceinline__ float value32(float4 s, float2 x, float &A) {
  A = fabsf(s.x) * fabsf(x.x) + fabsf(s.y) * fabsf(x.y) + fabsf(s.z);
  return s.x * x.x + s.y * x.y + s.z;
}

// Kernels over pools whose live count is known only on the device (polygons,
// patch edges, boundary vertices, hits) launch a capped grid and stride over
// the live count, instead of one mostly idle lane per capacity slot.
// WAVE_LOOP keeps every lane of a block in the loop for the same trips, so
// wave collectives in the body see all 64 lanes (lanes past the count run the
// body as inactive); ITEM_LOOP is for bodies without wave collectives.
constexpr int STRIDE_BLOCKS = 8192;  // 32 blocks of 256 lanes per CU
inline dim3 stride_grid(long items, int block = 256) {
  long g = (items + block - 1) / block;
  return dim3((unsigned)(g < 1 ? 1 : (g > STRIDE_BLOCKS ? STRIDE_BLOCKS : g)));
}
#define ITEM_LOOP(i, n) for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += gridDim.x * blockDim.x)
#define WAVE_LOOP(base, n) for (int base = blockIdx.x * blockDim.x; base < (n); base += gridDim.x * blockDim.x)

// Live size of an open-addressing table: the power of two at least twice the
// live key count (load <= 1/2), at least 256, at most the allocated power of
// two. Tables are cleared and probed only over their live size, so a replay
// with few keys touches few lines.
__device__ __forceinline__ int live_table(long count, int capacity) {
  long need = 2 * count;
  int size = 256;
  while (size < capacity && size < need) size <<= 1;
  return size < capacity ? size : capacity;
}

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
  // splitmix64 finaliser
  x = (x ^ (x >> 30)) * 0xbf58476d1ce4e5b9ull;
  x = (x ^ (x >> 27)) * 0x94d049bb133111ebull;
  return x ^ (x >> 31);
}
__device__ __forceinline__ uint64_t dbits(double x) {
  if (x == 0.0) x = 0.0;
  return (uint64_t)__double_as_longlong(x);
}

// ---------------------------------------------------------------------------
// Stackless traversal of the implicit complete tree (warp_peel/kernels.py:32-64,
// event_kernels.py:172-193): returns the next leaf id overlapping [lo, hi] in
// depth-first order from `node`, and the continuation node (0 when exhausted).
__device__ __forceinline__ int successor(int node) {
  while (node > 0 && (node & 1) != 0) node >>= 1;
  if (node > 0) node += 1;
  return node;
}
__device__ __forceinline__ bool overlaps(const Box &b, float4 lo, float4 hi) {
  return b.lo.x <= hi.x && b.lo.y <= hi.y && b.lo.z <= hi.z && b.hi.x >= lo.x && b.hi.y >= lo.y &&
         b.hi.z >= lo.z;
}
__device__ __forceinline__ int query_next(const Tree &t, float4 lo, float4 hi, int &node, int end = 0) {
  while (node > 0 && node != end) {
    Box b = t.nodes[node];
    if (overlaps(b, lo, hi)) {
      if (node < t.leaves) {
        node = 2 * node;
      } else {
        int found = t.ids[node - t.leaves];
        node = successor(node);
        return found;
      }
    } else {
      node = successor(node);
    }
  }
  return -1;
}

// Depth-first walk of a subtree without any stack in memory: the walk goes
// down four grandchildren at a time and ends in leaf buckets of 2^bucket_log
// leaves (bucket_log = the tuned bucket size or one more, so that the levels
// in between pair up), whose boxes are tested together (one to eight 128-byte
// lines). What is left to visit is one 4-bit mask of overlapping grandchildren
// per level, packed in one 64-bit word (16 levels), plus the bucket's mask of
// overlapping leaves; the node a popped mask bit stands for follows from the
// current node by shifts. The leaves reached and their order are those of the
// binary walk over the same subtree (a box lies inside its parent's), and the
// binary continuation after a leaf is still successor(leaf).
struct TrailWalk {
  uint64_t trail;   // level k (grandchildren of a level k-1 node) at bits [4k, 4k + 4)
  uint32_t bucket;  // overlapping leaves of the current bucket not yet returned
  int cur;          // node of the deepest expanded level (kc)
  int kc;           // its level (-1: the subtree root)
  int root, lvls;   // subtree root; 4-ary levels between it and the bucket roots
  int blog;         // log2 of the bucket size

  // plan the levels of the subtree at `r` with buckets of about 2^want leaves
  __device__ __forceinline__ void plan(const Tree &t, int r, int want) {
    int D = 31 - __clz(t.leaves), h = D - (31 - __clz(r));
    blog = h <= want ? h : want + ((h - want) & 1);
    lvls = (h - blog) >> 1;
    root = r;
    trail = 0;
    bucket = 0;
  }
  __device__ __forceinline__ uint32_t mask4(const Tree &t, int n, float4 lo, float4 hi) const {
    const Box *g = t.nodes + 4 * n;
    Box g0 = g[0], g1 = g[1], g2 = g[2], g3 = g[3];
    return (uint32_t)overlaps(g0, lo, hi) | (uint32_t)overlaps(g1, lo, hi) << 1 |
           (uint32_t)overlaps(g2, lo, hi) << 2 | (uint32_t)overlaps(g3, lo, hi) << 3;
  }
  __device__ __forceinline__ uint32_t leaves_mask(const Tree &t, int b, float4 lo, float4 hi) const {
    const Box *g = t.nodes + ((long)b << blog);
    uint32_t m = 0;
    for (int j = 0; j < (1 << blog); j += 4) {
      Box g0 = g[j], g1 = g[j + 1], g2 = g[j + 2], g3 = g[j + 3];
      m |= ((uint32_t)overlaps(g0, lo, hi) | (uint32_t)overlaps(g1, lo, hi) << 1 |
            (uint32_t)overlaps(g2, lo, hi) << 2 | (uint32_t)overlaps(g3, lo, hi) << 3) << j;
    }
    return m;
  }
  __device__ __forceinline__ uint32_t small_mask(const Tree &t, int b, float4 lo, float4 hi) const {
    // buckets of one or two leaves (subtrees of height 0 or 1)
    const Box *g = t.nodes + ((long)b << blog);
    uint32_t m = overlaps(g[0], lo, hi);
    if (blog == 1) m |= (uint32_t)overlaps(g[1], lo, hi) << 1;
    return m;
  }
  __device__ __forceinline__ uint32_t bucket_mask(const Tree &t, int b, float4 lo, float4 hi) const {
    return blog >= 2 ? leaves_mask(t, b, lo, hi) : small_mask(t, b, lo, hi);
  }
  // Overlap mask of `count` consecutive boxes, read four at a time (the node
  // array has four boxes of slack past its end for groups of one or two).
  __device__ __forceinline__ uint32_t group_mask(const Box *g, int count, float4 lo, float4 hi) const {
    uint32_t m = 0;
    for (int j = 0; j < count; j += 4) {
      Box g0 = g[j], g1 = g[j + 1], g2 = g[j + 2], g3 = g[j + 3];
      m |= ((uint32_t)overlaps(g0, lo, hi) | (uint32_t)overlaps(g1, lo, hi) << 1 |
            (uint32_t)overlaps(g2, lo, hi) << 2 | (uint32_t)overlaps(g3, lo, hi) << 3) << j;
    }
    return count < 4 ? m & ((1u << count) - 1) : m;
  }
  // every leaf of the subtree at r
  __device__ void start(const Tree &t, int r, int want, float4 lo, float4 hi) {
    plan(t, r, want);
    kc = -1;
    cur = r;
    if (lvls == 0) {
      kc = 0;
      bucket = bucket_mask(t, r, lo, hi);
    } else {
      trail = mask4(t, r, lo, hi);
    }
  }
  // the leaves of the subtree at r from leaf node `first` on (first inside it)
  __device__ void start_at(const Tree &t, int r, int first, int want, float4 lo, float4 hi) {
    plan(t, r, want);
    int b = first >> blog;  // its bucket root, at level lvls - 1 (or r itself)
    for (int k = 0; k < lvls; ++k) {
      int node = b >> (2 * (lvls - 1 - k));  // the path's node at level k
      int idx = node & 3;
      uint32_t later = mask4(t, node >> 2, lo, hi) & (0xEu << idx) & 0xFu;  // its right siblings
      trail |= (uint64_t)later << (4 * k);
    }
    cur = b;
    kc = lvls - 1;
    uint32_t m = bucket_mask(t, b, lo, hi);
    int from = first - (b << blog);
    bucket = m & (0xFFFFFFFFu << from);
  }
  // heap node of the next overlapping leaf; -1 when done
  __device__ int next_node(const Tree &t, float4 lo, float4 hi) {
    while (true) {
      if (bucket) {
        int j = __ffs(bucket) - 1;
        bucket &= bucket - 1;
        return (cur << blog) + j;
      }
      if (!trail) return -1;
      int lvl = (63 - __clzll((long long)trail)) >> 2;  // deepest level with a pending node
      int idx = __ffs((uint32_t)(trail >> (4 * lvl)) & 0xFu) - 1;
      trail &= ~(1ull << (4 * lvl + idx));
      int parent = lvl == 0 ? root : cur >> (2 * (kc - lvl + 1));
      cur = 4 * parent + idx;
      kc = lvl;
      // one read for every lane of a wave, whether it expands an internal
      // node (four grandchildren) or reaches a bucket of leaves
      const bool bottom = kc == lvls - 1;
      uint32_t m = group_mask(t.nodes + (bottom ? ((long)cur << blog) : 4L * cur), bottom ? 1 << blog : 4, lo, hi);
      if (bottom)
        bucket = m;
      else
        trail |= (uint64_t)m << (4 * (kc + 1));
    }
  }
};

// ---------------------------------------------------------------------------
// Wave64 aggregation. Same-address atomics serialise in L2, so counters and
// pool cursors are bumped once per wave.
__device__ __forceinline__ uint64_t lanemask_lt() {
  int lane = threadIdx.x & 63;
  return lane ? (~0ull >> (64 - lane)) : 0ull;
}
// Claim one slot per lane with want == true (works under divergence: only
// active lanes vote). Returns the slot, or -1.
__device__ __forceinline__ long long wave_claim(unsigned long long *counter, bool want) {
  uint64_t m = __ballot(want);
  if (!m) return -1;
  int lane = threadIdx.x & 63, leader = __ffsll((unsigned long long)m) - 1;
  unsigned long long base = 0;
  if (lane == leader) base = atomicAdd(counter, (unsigned long long)__popcll(m));
  base = __shfl(base, leader, 64);
  return want ? (long long)(base + __popcll(m & lanemask_lt())) : -1;
}
__device__ __forceinline__ int wave_claim32(int *counter, bool want) {
  uint64_t m = __ballot(want);
  if (!m) return -1;
  int lane = threadIdx.x & 63, leader = __ffsll((unsigned long long)m) - 1;
  int base = 0;
  if (lane == leader) base = atomicAdd(counter, __popcll(m));
  base = __shfl(base, leader, 64);
  return want ? base + __popcll(m & lanemask_lt()) : -1;
}
// Full-wave reductions: every lane of the wave must be active.
__device__ __forceinline__ long long wave_sum(long long v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}
__device__ __forceinline__ int wave_max(int v) {
  for (int off = 32; off > 0; off >>= 1) v = max(v, __shfl_xor(v, off, 64));
  return v;
}
// Exclusive prefix sum across the full wave; total in `total`.
__device__ __forceinline__ int wave_prefix(int v, int &total) {
  int lane = threadIdx.x & 63, x = v;
  for (int off = 1; off < 64; off <<= 1) {
    int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  total = __shfl(x, 63, 64);
  return x - v;
}
__device__ __forceinline__ void wave_add_metric(long long *metric, long long v) {
  v = wave_sum(v);
  if ((threadIdx.x & 63) == 0 && v) atomicAdd((unsigned long long *)metric, (unsigned long long)v);
}

// Same traversal, returning the leaf index (position in Morton order).
__device__ __forceinline__ int query_leaf(const Tree &t, float4 lo, float4 hi, int &node, int end = 0) {
  while (node > 0 && node != end) {
    Box b = t.nodes[node];
    if (overlaps(b, lo, hi)) {
      if (node < t.leaves) {
        node = 2 * node;
      } else {
        int leaf = node - t.leaves;
        node = successor(node);
        return leaf;
      }
    } else {
      node = successor(node);
    }
  }
  return -1;
}

// clean (frontier.py:45-71 = kernels.py:146-176): in-place cleanup of a
// strided ring (vertex j at xy[j * stride]); returns the new count.
__device__ inline int clean_slot(double2 *xy, int stride, int n) {
  if (n > MAXV) return n;
  int count = 0;
  double2 last = make_double2(0, 0);
  for (int j = 0; j < n; ++j) {
    double2 x = xy[(long)j * stride];
    if (count == 0 || dlength(make_double2(x.x - last.x, x.y - last.y)) > EPS * 0.01) {
      xy[(long)count * stride] = x;
      last = x;
      count += 1;
    }
  }
  if (count > 1) {
    double2 first = xy[0];
    if (dlength(make_double2(first.x - last.x, first.y - last.y)) <= EPS * 0.01) count -= 1;
  }
  if (count < 3) return 0;
  double2 origin = xy[0];
  double area2 = 0.0, perimeter = 0.0;
  for (int j = 0; j < count; ++j) {
    double2 a = xy[(long)j * stride], b = xy[(long)((j + 1) % count) * stride];
    area2 += cross3(origin, a, b);
    perimeter += fabs(b.x - a.x) + fabs(b.y - a.y);
  }
  if (area2 <= EPS * perimeter) return 0;
  return count;
}

// Padded binary32 box of a ring evaluated with its depth plane
// (warp_peel/kernels.py:235-250). Conversions round to nearest.
__device__ __forceinline__ void pad_box(double lx, double ly, double lz, double hx, double hy, double hz,
                                        float4 &lo, float4 &hi) {
  double pz = 2.0e-7 * fmax(1.0, fmax(fabs(lz), fabs(hz)));
  lo = make_float4((float)(lx - 2.0e-7), (float)(ly - 2.0e-7), (float)(lz - pz), 0.f);
  hi = make_float4((float)(hx + 2.0e-7), (float)(hy + 2.0e-7), (float)(hz + pz), 0.f);
}

}  // namespace peel
