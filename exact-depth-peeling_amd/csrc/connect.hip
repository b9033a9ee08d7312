// Stage 3a: rank propagation across certified patch adjacencies.
//
// Reference semantics (warp_peel/components.py, alecjacobson/exact-depth-peeling):
//  connection condition  components.py:32-52 (splits > 0 or an internal edge)
//  patch edge owners     components.py:80-120 (topological owner key)
//  hash matching         components.py:123-159 (<= 64 probes, misses counted)
//  certified connect     components.py:162-243 (midpoint contour query,
//                        lock-free min-index union-find)
//  roots / identity      components.py:246-271
//  convex coalescing     components.py:274-395 (propose, merge, live_counts,
//                        compact, restore; merge_regions=True)
//
// Patch edges are tasks (polygon, edge) at the scanned ring offsets. An edge's
// owner is the welded key of the internal mesh edge it lies on, or 2^63 |
// source for a source-local cut, so only exact reversed twins on the same
// topological edge can match. A match is accepted only if no contour at the
// shared edge's midpoint is nearer than or level with the patch (padded), and
// then the two patches join one component whose representative casts the
// single ray for all of them.
#include <algorithm>

#include "workspace.hpp"

namespace peel {

__device__ __forceinline__ int npolys_c(const Ws &w) {
  int p = load_i(w.status + ST_POLYS);
  return p < w.cap.regions ? p : w.cap.regions;
}

__device__ __forceinline__ bool connecting(const Ws &w) {
  return w.status[ST_ERR] == 0 && load_i(w.control + C_CONNECT) != 0;
}

__device__ __forceinline__ int pedges(const Ws &w) {
  int total = load_i(w.control + C_PEDGES);
  return total < w.cap.vertices ? total : w.cap.vertices;
}
__device__ __forceinline__ int eslots_live(const Ws &w) { return live_table(pedges(w), w.ehash2_size); }

// Check the patch-edge total of the ring-offset scan, then clear the live part
// of the reversed-edge table and the partner slots.
__global__ void k_eslots_reset(Ws w) {
  if (blockIdx.x == 0 && threadIdx.x == 0 && w.control[C_PEDGES] > w.cap.vertices) set_error(w.status, ERR_VERTS);
  if (!connecting(w)) return;
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int size = eslots_live(w), n = pedges(w);
  for (int j = i; j < size; j += gridDim.x * blockDim.x) w.eslots[j] = -1;
  for (int j = i; j < n; j += gridDim.x * blockDim.x) w.partners[j] = -1;
}

// edge_owners (components.py:80-120)
__device__ void edge_owners_one(const Ws &w, int i);

__global__ void k_edge_owners(Ws w) {
  if (!connecting(w)) return;
  ITEM_LOOP(i, npolys_c(w)) edge_owners_one(w, i);
}

__device__ void edge_owners_one(const Ws &w, int i) {
  int n = w.out_n[i];
  int base = w.voffsets[i];
  if (base + n > w.cap.vertices) return;
  int source = w.out_src[i];
  long R = w.cap.regions;
  // The source's internal edges, in the order they are tried, with their
  // endpoints and owner keys, read once for all the patch's edges (the edge
  // stores below would otherwise make every edge read them again). A clipped
  // source has none; an unclipped one is a triangle (at most three).
  double2 eu[3], ev[3];
  unsigned long long ekey[3];
  int cand = 0;
  if (w.clipped[source] == 0) {
    const int qn = w.src_n[source];
    const bool back = w.facing[source] < 0;
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      if (j >= qn) continue;
      const int edge = back ? (4 - j) % 3 : j;
      if (w.masks[3 * source + edge] != 0) continue;
      const double2 u = w.src_xy[(long)j * w.nf + source], v = w.src_xy[(long)((j + 1) % qn) * w.nf + source];
      const int ca = w.canonical[w.faces[3 * source + edge]];
      const int cb = w.canonical[w.faces[3 * source + (edge + 1) % 3]];
      const unsigned long long key = ((unsigned long long)(uint32_t)min(ca, cb) << 32) | (uint32_t)max(ca, cb);
#pragma unroll
      for (int c = 0; c < 3; ++c)
        if (c == cand) {
          eu[c] = u;
          ev[c] = v;
          ekey[c] = key;
        }
      ++cand;
    }
  }
  for (int k = 0; k < n; ++k) {
    double2 a = w.out_xy[k * R + i], b = w.out_xy[((k + 1) % n) * R + i];
    unsigned long long owner = 0x8000000000000000ull | (unsigned long long)(uint32_t)source;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      if (c >= cand) break;
      const double2 u = eu[c], v = ev[c];
      // identity checks also avoid cancellation at shared endpoints
      if ((same(a, u) || same(a, v) || cross3(u, v, a) == 0.0) &&
          (same(b, u) || same(b, v) || cross3(u, v, b) == 0.0)) {
        owner = ekey[c];
        break;
      }
    }
    int slot = base + k;
    w.owners[slot] = owner;
    w.starts[slot] = a;
    w.ends[slot] = b;
    w.tasks[slot] = make_int2(i, k);
  }
}

__device__ __forceinline__ uint64_t point_key(double2 x) {
  // quantisation only places the key in the table; full coordinates decide
  return mix64((uint64_t)(long long)(x.x * 1.0e12)) ^ mix64((uint64_t)(long long)(x.y * 1.0e12) + 19ull);
}

// match_edges (components.py:123-159)
__device__ void match_one(const Ws &w, int i, int n);

__global__ __launch_bounds__(256) void k_match_edges(Ws w) {
  if (!connecting(w)) return;
  int n = pedges(w);
  WAVE_LOOP(base, n) match_one(w, base + threadIdx.x, n);
}

__device__ void match_one(const Ws &w, int i, int n) {
  bool active = i < n;
  bool matched = false, missed = false;
  if (active) {
    double2 a = w.starts[i], b = w.ends[i];
    if (!same(a, b)) {
      unsigned long long owner = w.owners[i];
      uint64_t key = mix64(owner) ^ point_key(a) ^ point_key(b);
      int mask = eslots_live(w) - 1;
      int slot = (int)(key & (uint64_t)mask);
      missed = true;
      for (int attempt = 0; attempt < 64; ++attempt) {
        // read first: only an empty-looking slot is claimed with an atomic
        int other = load_i(&w.eslots[slot]);
        if (other < 0) other = atomicCAS(&w.eslots[slot], -1, i);
        if (other < 0) {
          missed = false;
          break;
        }
        // the occupant's owner, task and endpoints were written by
        // k_edge_owners, a previous kernel: no fence is needed to read them
        if (w.owners[other] == owner && w.tasks[i].x != w.tasks[other].x && same(a, w.ends[other]) &&
            same(b, w.starts[other])) {
          w.partners[i] = other;
          w.partners[other] = i;
          matched = true;
          missed = false;
          break;
        }
        slot = (slot + 1) & mask;
      }
    }
  }
  metric_add(w, M_MATCHED, matched);
  metric_add(w, M_HASH_MISSES, missed);
}

__device__ int find_root(int *parents, int i) {
  int node = i;
  while (true) {
    int parent = load_i(parents + node);
    if (parent == node) break;
    int grand = load_i(parents + parent);
    // path halving; a node already pointing at its root needs no atomic
    if (grand != parent) atomicMin(parents + node, grand);
    node = grand;
  }
  return node;
}

__device__ void unite(int *parents, int a, int b) {
  int ra = find_root(parents, a), rb = find_root(parents, b);
  while (ra != rb) {
    int high = max(ra, rb), low = min(ra, rb);
    int old = atomicMin(parents + high, low);
    if (old == high) break;
    ra = find_root(parents, old);
    rb = find_root(parents, low);
  }
}

// connect (components.py:162-243)
__device__ void connect_one(const Ws &w, int i, int n);

__global__ __launch_bounds__(256) void k_connect(Ws w) {
  if (!connecting(w)) return;
  int n = pedges(w);
  WAVE_LOOP(base, n) connect_one(w, base + threadIdx.x, n);
}

__device__ void connect_one(const Ws &w, int i, int n) {
  bool active = i < n;
  int other = active ? w.partners[i] : -1;
  active = active && other > i;
  long long visits = 0;
  bool linked = false;
  if (active) {
    int2 task = w.tasks[i];
    double2 a = w.starts[i], b = w.ends[i];
    if (dlength(make_double2(b.x - a.x, b.y - a.y)) > 16.0 * EPS) {
      double2 x = make_double2((a.x + b.x) * 0.5, (a.y + b.y) * 0.5);
      double z = value(w.scores[w.out_src[task.x]], x);
      double pad = 2.0e-7 * fmax(1.0, fabs(z));
      float4 lo = make_float4((float)x.x - 2.0e-7f, (float)x.y - 2.0e-7f, -1.0e30f, 0.f);
      float4 hi = make_float4((float)x.x + 2.0e-7f, (float)x.y + 2.0e-7f, (float)(z + pad), 0.f);
      Tree tree = w.contour_tree.tree();
      TrailWalk walk;
      walk.start(tree, 1, w.bucket_log, lo, hi);
      bool safe = true;
      while (true) {
        int leaf_node = walk.next_node(tree, lo, hi);
        if (leaf_node < 0) break;
        visits += 1;
        const int c = leaf_node - tree.leaves;  // the segment's record in leaf order
        double2 u = w.cl_a[c], v = w.cl_b[c];
        double2 d = make_double2(v.x - u.x, v.y - u.y);
        double length = dlength(d);
        if (length > 0.0) {
          double t = ((x.x - u.x) * d.x + (x.y - u.y) * d.y) / (d.x * d.x + d.y * d.y);
          if (fabs(cross3(u, v, x)) <= EPS * length && t >= -EPS / length && t <= 1.0 + EPS / length) {
            double2 nr = w.cl_near[c];
            double n0 = nr.x, n1 = nr.y;
            double depth = n0 + t * (n1 - n0);
            if (depth <= z + pad) {
              safe = false;
              break;
            }
          }
        }
      }
      if (safe) {
        unite(w.parents, task.x, w.tasks[other].x);
        linked = true;
      }
    }
  }
  metric_add(w, M_CONNECTION_TESTS, visits);
  metric_add(w, M_CONNECTIONS, linked);
}

// roots (components.py:246-258)
__global__ __launch_bounds__(256) void k_roots(Ws w) {
  if (!connecting(w)) return;
  int n = npolys_c(w);
  long long roots = 0;
  ITEM_LOOP(i, n) {
    int r = find_root(w.parents, i);
    w.reps[i] = r;
    roots += r == i;
  }
  metric_add(w, M_COMPONENTS, roots);
}

static inline dim3 grid_for(long n, int block = 256) {
  long g = (n + block - 1) / block;
  return dim3((unsigned)(g < 1 ? 1 : g));
}

// After the rank stage's reset kernel, which also resets the connectivity.
void enqueue_connect(const Ws &w, hipStream_t s) {
  // ring offsets of the patches: an exclusive scan of their vertex counts
  scan_exclusive(s, w.out_n, w.voffsets, w.status + ST_POLYS, 0, w.cap.regions, w.control + C_PEDGES, w.scan);
  long eg = ((long)std::max(w.ehash2_size, w.cap.vertices) + 255) / 256;
  hipLaunchKernelGGL(k_eslots_reset, dim3((unsigned)(eg > 1024 ? 1024 : eg)), dim3(256), 0, s, w);
  hipLaunchKernelGGL(k_edge_owners, stride_grid(w.cap.regions), dim3(256), 0, s, w);
  hipLaunchKernelGGL(k_match_edges, stride_grid(w.cap.vertices), dim3(256), 0, s, w);
  if (w.propagate_ranks) {
    hipLaunchKernelGGL(k_connect, stride_grid(w.cap.vertices), dim3(256), 0, s, w);
    hipLaunchKernelGGL(k_roots, stride_grid(w.cap.regions), dim3(256), 0, s, w);
  }
}

// ---------------------------------------------------------------------------
// Convex coalescing (components.py:274-395), after ranks: two matched patch
// edges of the same source and layer propose joining their patches when both
// junctions stay convex; each patch keeps its smallest proposal, a pair whose
// proposals agree is joined (collinear vertices dropped), and the surviving
// patches are compacted.
__device__ __forceinline__ double2 ring_vertex(const Ws &w, int p, int j) {
  return w.out_xy[(long)j * w.cap.regions + p];
}

__device__ __forceinline__ bool merge_pair(const Ws &w, int i, int &pa, int &pb, int &ka, int &kb) {
  int other = w.partners[i];
  if (other <= i) return false;
  int2 ta = w.tasks[i], tb = w.tasks[other];
  pa = ta.x, pb = tb.x, ka = ta.y, kb = tb.y;
  return true;
}

__global__ void k_propose(Ws w) {
  if (!connecting(w)) return;
  ITEM_LOOP(i, pedges(w)) {
    int pa, pb, ka, kb;
    if (!merge_pair(w, i, pa, pb, ka, kb)) continue;
    if (w.out_src[pa] != w.out_src[pb] || w.out_layer[pa] != w.out_layer[pb]) continue;
    int na = w.out_n[pa], nb = w.out_n[pb];
    if (na < 3 || nb < 3 || na + nb - 2 > MAXV) continue;
    double2 u = ring_vertex(w, pa, ka), v = ring_vertex(w, pb, kb);
    if (cross3(ring_vertex(w, pa, (ka + na - 1) % na), u, ring_vertex(w, pb, (kb + 2) % nb)) >= 0.0 &&
        cross3(ring_vertex(w, pb, (kb + nb - 1) % nb), v, ring_vertex(w, pa, (ka + 2) % na)) >= 0.0) {
      atomicMin(w.choices + pa, i);
      atomicMin(w.choices + pb, i);
    }
  }
}

__global__ void k_merge(Ws w) {
  if (!connecting(w)) return;
  long long merged = 0;
  ITEM_LOOP(i, pedges(w)) {
    int pa, pb, ka, kb;
    if (!merge_pair(w, i, pa, pb, ka, kb)) continue;
    if (w.choices[pa] != i || w.choices[pb] != i) continue;
    int na = w.out_n[pa], nb = w.out_n[pb];
    // a's ring from the vertex after the shared edge, then b's vertices
    // between the shared edge's ends
    double2 joined[MAXV];
    int m = 0;
    for (int j = 0; j < na; ++j) joined[m++] = ring_vertex(w, pa, (ka + 1 + j) % na);
    for (int j = 0; j < nb - 2; ++j) joined[m++] = ring_vertex(w, pb, (kb + 2 + j) % nb);
    int q = 0;
    for (int j = 0; j < m; ++j) {
      double2 prev = joined[(j + m - 1) % m], x = joined[j], next = joined[(j + 1) % m];
      if (cross3(prev, x, next) != 0.0) w.out_xy[(long)q++ * w.cap.regions + pa] = x;
    }
    // fewer than three vertices left cannot happen for two convex patches of
    // positive area; the reference keeps both patches unchanged then
    if (q >= 3) {
      w.out_n[pa] = q;
      w.out_n[pb] = 0;
      merged += 1;
    }
  }
  metric_add(w, M_MERGED, merged);
}

__global__ void k_live_counts(Ws w) {
  bool on = connecting(w);
  ITEM_LOOP(i, npolys_c(w)) w.vcounts[i] = on ? (w.out_n[i] > 0) : 1;
}

__global__ void k_compact(Ws w) {
  if (!connecting(w)) return;
  long R = w.cap.regions;
  ITEM_LOOP(i, npolys_c(w)) {
    int n = w.out_n[i];
    if (n == 0) continue;
    int slot = w.voffsets[i];
    for (int j = 0; j < n; ++j) w.merge_xy[(long)j * R + slot] = w.out_xy[(long)j * R + i];
    w.merge_n[slot] = n;
    w.merge_src[slot] = w.out_src[i];
    w.merge_layer[slot] = w.out_layer[i];
  }
}

__global__ void k_restore(Ws w) {
  if (!connecting(w)) return;
  long R = w.cap.regions;
  int live = load_i(w.control + C_LIVE);
  ITEM_LOOP(i, live) {
    int n = w.merge_n[i];
    for (int j = 0; j < n; ++j) w.out_xy[(long)j * R + i] = w.merge_xy[(long)j * R + i];
    w.out_n[i] = n;
    w.out_src[i] = w.merge_src[i];
    w.out_layer[i] = w.merge_layer[i];
  }
}

__global__ void k_live_total(Ws w) {
  if (connecting(w) && w.status[ST_ERR] == 0) w.status[ST_POLYS] = w.control[C_LIVE];
}

void enqueue_merge(const Ws &w, hipStream_t s) {
  hipLaunchKernelGGL(k_propose, stride_grid(w.cap.vertices), dim3(256), 0, s, w);
  hipLaunchKernelGGL(k_merge, stride_grid(w.cap.vertices), dim3(256), 0, s, w);
  hipLaunchKernelGGL(k_live_counts, stride_grid(w.cap.regions), dim3(256), 0, s, w);
  scan_exclusive(s, w.vcounts, w.voffsets, w.status + ST_POLYS, 0, w.cap.regions, w.control + C_LIVE,
                 w.scan);
  hipLaunchKernelGGL(k_compact, stride_grid(w.cap.regions), dim3(256), 0, s, w);
  hipLaunchKernelGGL(k_restore, stride_grid(w.cap.regions), dim3(256), 0, s, w);
  hipLaunchKernelGGL(k_live_total, dim3(1), dim3(1), 0, s, w);
}

}  // namespace peel
