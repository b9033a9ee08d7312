// Stage 3: shared ray ranks for the finished patches.
//
// Reference semantics (file:line in alecjacobson/exact-depth-peeling):
//  samples          warp_peel/event_kernels.py:28-48
//  exact groups     warp_peel/event_kernels.py:51-92 (group_queries)
//  contains_flat    warp_peel/event_kernels.py:131-148 (half-open ownership)
//  store_hit        warp_peel/event_kernels.py:151-169
//  ray_hits         warp_peel/event_kernels.py:196-245
//  heavy_ray_hits   warp_peel/event_kernels.py:248-298
//  hit order        warp_peel/events.py:703-718 (stable source, depth, group)
//  assign_ranks     warp_peel/event_kernels.py:342-385
//
// Samples with identical coordinates share one ray group; the group leader is
// the smallest patch index holding that sample (hash table + atomicMin instead
// of the reference's two float64 radix sorts and a binary search).
#include "workspace.hpp"

namespace peel {

__device__ __forceinline__ int npolys(const Ws &w) {
  int p = load_i(w.status + ST_POLYS);
  return p < w.cap.regions ? p : w.cap.regions;
}

__device__ __forceinline__ uint64_t point_hash(double2 x) {
  return mix64(dbits(x.x) ^ mix64(dbits(x.y) + 0x9e3779b97f4a7c15ull));
}

__global__ void k_rank_reset(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  for (int j = i; j < w.qhash_size; j += gridDim.x * blockDim.x) w.qhash[j] = -1;
  if (i == 0) {
    w.control[C_HEAVY] = 0;
  }
}

// Canonical sample of the representative patch: centroid summed from the
// lexicographically smallest vertex (rotation invariant).
__global__ void k_samples(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int P = npolys(w);
  if (i >= P || w.status[ST_ERR] != 0) return;
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
  __threadfence();
  int mask = w.qhash_size - 1;
  int slot = (int)(point_hash(x) & (uint64_t)mask);
  for (int probe = 0; probe < w.qhash_size; ++probe) {
    int cur = atomicCAS(&w.qhash[slot], -1, i);
    if (cur < 0) return;
    __threadfence();
    if (same(w.samples[cur], x)) {
      atomicMin(&w.qhash[slot], i);
      return;
    }
    slot = (slot + 1) & mask;
  }
}

__global__ void k_groups(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int P = npolys(w);
  if (i >= P || w.status[ST_ERR] != 0) return;
  double2 x = w.samples[i];
  int mask = w.qhash_size - 1;
  int slot = (int)(point_hash(x) & (uint64_t)mask);
  int leader = i;
  for (int probe = 0; probe < w.qhash_size; ++probe) {
    int cur = w.qhash[slot];
    if (cur < 0) break;
    if (same(w.samples[cur], x)) {
      leader = cur;
      break;
    }
    slot = (slot + 1) & mask;
  }
  w.groups[i] = leader;
}

// Half-open point-in-convex test against the flat source ring.
__device__ __forceinline__ bool contains_flat(const Ws &w, int source, double2 x) {
  int n = w.src_n[source];
  bool inside = n > 0;
  for (int j = 0; j < n && inside; ++j) {
    double2 a = w.src_xy[j * w.nf + source];
    double2 b = w.src_xy[((j + 1) % n) * w.nf + source];
    double s = cross3(a, b, x);
    bool owned = b.y > a.y || (b.y == a.y && b.x < a.x);
    if (s < 0.0 || (s == 0.0 && !owned)) inside = false;
  }
  return inside;
}

__device__ __forceinline__ void store_hit(const Ws &w, int source, int group, double2 x) {
  unsigned long long slot = atomicAdd((unsigned long long *)&w.metrics[M_RAY_HITS], 1ull);
  if (slot < (unsigned long long)w.cap.hits) {
    w.hit_depth[slot] = value(w.scores[source], x);
    w.hit_src[slot] = source;
    w.hit_group[slot] = group;
  } else {
    set_error(w.status, ERR_HITS);
  }
}

__device__ __forceinline__ void ray_box(double2 x, float4 &lo, float4 &hi) {
  lo = make_float4((float)x.x - 2.0e-7f, (float)x.y - 2.0e-7f, -1.0e30f, 0.f);
  hi = make_float4((float)x.x + 2.0e-7f, (float)x.y + 2.0e-7f, 1.0e30f, 0.f);
}

__global__ void k_ray_hits(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int P = npolys(w);
  if (i >= P || w.status[ST_ERR] != 0) return;
  if (w.groups[i] != i) return;
  double2 x = w.samples[i];
  Tree tree = w.source_tree.tree();
  float4 lo, hi;
  ray_box(x, lo, hi);
  int node = 1, count = 0, visits = 0;
  for (int step = 0; step < 32; ++step) {
    int source = query_next(tree, lo, hi, node);
    if (source < 0) break;
    visits += 1;
    if (contains_flat(w, source, x)) {
      store_hit(w, source, i, x);
      count += 1;
    }
    if (node == 0) break;
  }
  atomicAdd(&w.group_counts[i], count);
  atomicAdd((unsigned long long *)&w.metrics[M_RAY_GROUPS], 1ull);
  atomicAdd((unsigned long long *)&w.metrics[M_RAY_TESTS], (unsigned long long)visits);
  if (node > 0) {
    // the unvisited DFS suffix starts at the first leaf below the cursor
    while (node < tree.leaves) node *= 2;
    int slot = atomicAdd(w.control + C_HEAVY, 1);
    atomicAdd((unsigned long long *)&w.metrics[M_HEAVY], 1ull);
    w.heavy[slot] = make_int2(i, node);
  }
}

__global__ void k_heavy_hits(Ws w, int workers) {
  if (w.status[ST_ERR] != 0) return;
  Tree tree = w.source_tree.tree();
  int jobs = load_i(w.control + C_HEAVY);
  int parts = min(8192, tree.leaves);
  while (parts > 16 && (long long)parts * jobs > 65536) parts /= 2;
  int width = tree.leaves / parts;
  long total = (long)jobs * parts;
  long long visits = 0;
  for (long task = blockIdx.x * blockDim.x + threadIdx.x; task < total; task += workers) {
    int2 job = w.heavy[task / parts];
    int part = (int)(task % parts);
    int first = tree.leaves + part * width;
    if (job.y >= first + width) continue;
    int root = parts + part;
    int node = job.y > first ? job.y : root;
    int end = successor(root);
    int i = job.x;
    double2 x = w.samples[i];
    float4 lo, hi;
    ray_box(x, lo, hi);
    int count = 0;
    while (node > 0 && node != end) {
      int source = query_next(tree, lo, hi, node, end);
      if (source < 0) break;
      visits += 1;
      if (contains_flat(w, source, x)) {
        store_hit(w, source, i, x);
        count += 1;
      }
    }
    if (count) atomicAdd(&w.group_counts[i], count);
  }
  if (visits) atomicAdd((unsigned long long *)&w.metrics[M_RAY_TESTS], (unsigned long long)visits);
}

enum { C_NHITS = 4 };

__global__ void k_check_hits(Ws w) {
  long long h = w.metrics[M_RAY_HITS];
  if (h > w.cap.hits) set_error(w.status, ERR_HITS);
  w.control[C_NHITS] = (int)(h < w.cap.hits ? h : w.cap.hits);
}

// Stable LSD stages: source, then depth, then group => (group, depth, source).
__global__ void k_hit_keys(Ws w, int stage) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int H = load_i(w.control + C_NHITS);
  if (i >= H) return;
  int h = stage == 0 ? i : w.hids[i];
  if (stage == 0) w.hids[i] = i;
  uint64_t key;
  if (stage == 0)
    key = (uint32_t)w.hit_src[h];
  else if (stage == 1)
    key = dkey(w.hit_depth[h]);
  else
    key = (uint32_t)w.hit_group[h];
  w.hkeys[i] = key;
}

__global__ void k_assign(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int P = npolys(w);
  if (i >= P || w.status[ST_ERR] != 0) return;
  int source = w.out_src[w.reps[i]];
  double2 x = w.samples[i];
  double z = value(w.scores[source], x);
  int g = w.groups[i];
  int low = 0, high = w.group_counts[g];
  int base = w.group_offsets[g];
  while (low < high) {
    int mid = (low + high) / 2;
    int h = w.hids[base + mid];
    int hs = w.hit_src[h];
    double hz = value(w.scores[hs], x);
    if (hz < z || (hz == z && hs < source))
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
  } else {
    atomicMax(w.status + ST_LAYERS, low + 1);
  }
}

static inline dim3 grid_for(long n, int block = 256) {
  long g = (n + block - 1) / block;
  return dim3((unsigned)(g < 1 ? 1 : g));
}

void enqueue_connect(const Ws &w, hipStream_t s);

void enqueue_rank(const Ws &w, hipStream_t s) {
  long qgrid = ((long)w.qhash_size + 255) / 256;
  if (qgrid > 2048) qgrid = 2048;
  hipLaunchKernelGGL(k_rank_reset, dim3((unsigned)qgrid), dim3(256), 0, s, w);
  if (w.propagate_ranks) enqueue_connect(w, s);
  hipLaunchKernelGGL(k_samples, grid_for(w.cap.regions), dim3(256), 0, s, w);
  hipLaunchKernelGGL(k_groups, grid_for(w.cap.regions), dim3(256), 0, s, w);
  hipLaunchKernelGGL(k_ray_hits, grid_for(w.cap.regions), dim3(256), 0, s, w);
  hipLaunchKernelGGL(k_heavy_hits, dim3(32), dim3(256), 0, s, w, 32 * 256);
  hipLaunchKernelGGL(k_check_hits, dim3(1), dim3(1), 0, s, w);
  scan_exclusive(s, w.group_counts, w.group_offsets, w.status + ST_POLYS, 0, w.cap.regions,
                 nullptr, w.scan);
  const int *nh = w.control + C_NHITS;
  int src_bits = 1, group_bits = 1;
  while ((1L << src_bits) < (long)w.nf) ++src_bits;
  while ((1L << group_bits) < (long)w.cap.regions) ++group_bits;
  for (int stage = 0; stage < 3; ++stage) {
    hipLaunchKernelGGL(k_hit_keys, grid_for(w.cap.hits), dim3(256), 0, s, w, stage);
    int bits = stage == 0 ? src_bits : (stage == 1 ? 64 : group_bits);
    radix_sort(s, w.hkeys, w.hids, w.hkeys_alt, w.hids_alt, nh, 0, w.cap.hits, 0, bits, w.sort);
  }
  hipLaunchKernelGGL(k_assign, grid_for(w.cap.regions), dim3(256), 0, s, w);
}

}  // namespace peel
