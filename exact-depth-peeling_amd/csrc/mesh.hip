// Stage 4: boundary noding, centroid fans and world lifting.
//
// Reference semantics (file:line in alecjacobson/exact-depth-peeling):
//  vertex_counts / scatter  warp_peel/kernels.py:941-972
//  mesh_tasks               warp_peel/event_kernels.py:504-516
//  point_bounds + layered   warp_peel/kernels.py:975-991, events.py:744-754
//  lift                     warp_peel/kernels.py:994-1006
//  mesh_one_edge            warp_peel/kernels.py:1011-1124
//
// One lane per polygon edge: a point-hierarchy query collects same-layer
// boundary vertices strictly inside the edge (t in (EPS, 1-EPS), distance
// <= EPS * length), sorts their parameters, and fans the noded edge from the
// polygon centroid. Parameters live in a small register window; an edge with
// more than the window goes through a per-lane spill area in HBM.
#include "workspace.hpp"

namespace peel {

__device__ __forceinline__ int npolys_m(const Ws &w) {
  int p = load_i(w.status + ST_POLYS);
  return p < w.cap.regions ? p : w.cap.regions;
}

__global__ void k_vertex_counts(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.cap.regions) return;
  int c = 0;
  if (i < npolys_m(w) && w.status[ST_ERR] == 0) c = w.out_n[i];
  w.vcounts[i] = c;
}

__global__ void k_scatter(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npolys_m(w) || w.status[ST_ERR] != 0) return;
  int n = w.out_n[i];
  int base = w.voffsets[i];
  if (n == 0) return;
  if (base + n > w.cap.vertices) {
    set_error(w.status, ERR_VERTS);
    return;
  }
  int layer = w.out_layer[i];
  long R = w.cap.regions;
  for (int j = 0; j < n; ++j) {
    w.bverts[base + j] = w.out_xy[j * R + i];
    w.blayers[base + j] = layer;
    w.tasks[base + j] = make_int2(i, j);
  }
}

// Boundary vertices repeat once per incident polygon; only the first copy of
// each exact (layer, x, y) enters the point hierarchy. Duplicates would give
// the same edge parameter, which the fan loop skips anyway, so the output is
// unchanged and each edge query sees each distinct vertex once.
__device__ __forceinline__ uint64_t bvert_hash(double2 x, int layer) {
  return mix64(dbits(x.x) ^ mix64(dbits(x.y) + 0x9e3779b97f4a7c15ull) ^
               mix64((uint64_t)(uint32_t)layer + 0x3c6ef372fe94f82aull));
}

__global__ void k_phash_reset(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  for (int j = i; j < w.phash_size; j += gridDim.x * blockDim.x) w.phash[j] = -1;
}

__global__ void k_point_dedupe(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= load_i(w.status + ST_BVERTS) || w.status[ST_ERR] != 0) return;
  double2 x = w.bverts[i];
  int layer = w.blayers[i];
  int mask = w.phash_size - 1;
  int slot = (int)(bvert_hash(x, layer) & (uint64_t)mask);
  for (int probe = 0; probe < w.phash_size; ++probe) {
    int cur = load_i(&w.phash[slot]);
    if (cur < 0) {
      cur = atomicCAS(&w.phash[slot], -1, i);
      if (cur < 0) return;
    }
    if (same(w.bverts[cur], x) && w.blayers[cur] == layer) {
      if (i < cur) atomicMin(&w.phash[slot], i);
      return;
    }
    slot = (slot + 1) & mask;
  }
}

__device__ __forceinline__ bool first_copy(const Ws &w, int i) {
  double2 x = w.bverts[i];
  int layer = w.blayers[i];
  int mask = w.phash_size - 1;
  int slot = (int)(bvert_hash(x, layer) & (uint64_t)mask);
  for (int probe = 0; probe < w.phash_size; ++probe) {
    int cur = w.phash[slot];
    if (cur < 0) return true;
    if (same(w.bverts[cur], x) && w.blayers[cur] == layer) return cur == i;
    slot = (slot + 1) & mask;
  }
  return true;
}

__global__ void k_point_flags(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.cap.vertices) return;
  w.pflag[i] = i < load_i(w.status + ST_BVERTS) && w.status[ST_ERR] == 0 && first_copy(w, i);
}

// Unique boundary vertices, compacted, with point boxes (z = layer) for the
// layered hierarchy (point_bounds, kernels.py:975-991).
__global__ void k_point_compact(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.cap.vertices || !w.pflag[i]) return;
  int k = w.poff[i];
  double2 p = w.bverts[i];
  float l = (float)w.blayers[i];
  w.upts[k] = p;
  Box b;
  b.lo = make_float4((float)p.x - 2.0e-7f, (float)p.y - 2.0e-7f, l, 0.f);
  b.hi = make_float4((float)p.x + 2.0e-7f, (float)p.y + 2.0e-7f, l, 0.f);
  w.point_boxes[k] = b;
}

__global__ void k_point_count(Ws w) { w.metrics[M_UNIQUE_BVERTS] = w.control[C_NUP]; }

__device__ __forceinline__ void lift(const Ws &w, double2 x, double4 sc, double *out) {
  const double *s = w.settings;
  double u = s[0] + x.x * (s[2] - s[0]);
  double v = s[1] + x.y * (s[3] - s[1]);
  double z = value(sc, x);
  if (s[4] != 0.0) {
    z = -1.0 / z;
    u *= z;
    v *= z;
  }
  const double *m = w.inverse;
  for (int r = 0; r < 3; ++r) out[r] = m[4 * r] * u + m[4 * r + 1] * v + m[4 * r + 2] * z + m[4 * r + 3] * 1.0;
}

constexpr int EDGE_MAX = 128;

// Heapsort in place (kernels.py:1066-1088).
__device__ void heapsort(double *params, int count) {
  int start = count / 2 - 1, end = count - 1;
  while (end > 0) {
    int root = 0;
    if (start >= 0) {
      root = start;
      start -= 1;
    } else {
      double t = params[end];
      params[end] = params[0];
      params[0] = t;
      end -= 1;
    }
    while (root * 2 + 1 <= end) {
      int child = root * 2 + 1;
      if (child + 1 <= end && params[child] < params[child + 1]) child += 1;
      if (params[root] < params[child]) {
        double t = params[root];
        params[root] = params[child];
        params[child] = t;
        root = child;
      } else {
        break;
      }
    }
  }
}

// One lane per polygon edge (kernels.py:1011-1124). Phase 1 gathers and sorts
// the edge parameters and counts the fan triangles; one atomic per wave then
// reserves the wave's triangle slots; phase 2 writes uv and lifted corners.
__global__ __launch_bounds__(256) void k_mesh_edges(Ws w) {
  int j = blockIdx.x * blockDim.x + threadIdx.x;
  bool valid = j < load_i(w.status + ST_BVERTS) && w.status[ST_ERR] == 0;
  int i = 0, edge = 0, n = 0;
  if (valid) {
    int2 task = w.tasks[j];
    i = task.x;
    edge = task.y;
    n = w.out_n[i];
    valid = edge < n;
  }
  long R = w.cap.regions;
  double2 a = make_double2(0, 0), b = a, d = a, center = a;
  double length2 = 0.0;
  int layer = 0;
  double params[EDGE_MAX];
  int count = 0;
  if (valid) {
    a = w.out_xy[edge * R + i];
    b = w.out_xy[((edge + 1) % n) * R + i];
    d = make_double2(b.x - a.x, b.y - a.y);
    length2 = d.x * d.x + d.y * d.y;
    for (int k = 0; k < n; ++k) {
      double2 p = w.out_xy[k * R + i];
      center.x += p.x / (double)n;
      center.y += p.y / (double)n;
    }
    layer = w.out_layer[i];
    float4 lo = make_float4((float)fmin(a.x, b.x) - 2.0e-7f, (float)fmin(a.y, b.y) - 2.0e-7f, (float)layer, 0.f);
    float4 hi = make_float4((float)fmax(a.x, b.x) + 2.0e-7f, (float)fmax(a.y, b.y) + 2.0e-7f, (float)layer, 0.f);
    Tree tree = w.point_tree.tree();
    int node = 1;
    double len = sqrt(length2);
    while (node > 0 && length2 > 0.0) {
      int k = query_next(tree, lo, hi, node);
      if (k < 0) break;
      double2 x = w.upts[k];
      double t = ((x.x - a.x) * d.x + (x.y - a.y) * d.y) / length2;
      if (t > EPS && t < 1.0 - EPS && fabs(cross3(a, b, x)) <= EPS * len) {
        if (count < w.cap.edge_vertices) {
          params[count++] = t;
        } else {
          count += 1;
          break;
        }
      }
    }
    if (count > w.cap.edge_vertices) {
      set_error(w.status, ERR_EDGE);
      valid = false;
    }
  }
  int top = wave_max(valid ? count : 0);
  if ((threadIdx.x & 63) == 0 && top > 0)
    atomicMax((unsigned long long *)&w.metrics[M_EDGE_MAX], (unsigned long long)top);
  if (valid) heapsort(params, count);
  // phase 1: count emitted triangles
  int ntri = 0;
  if (valid) {
    double previous = 0.0;
    double2 cursor = a;
    for (int step = 0; step <= count; ++step) {
      double following = step < count ? params[step] : 1.0;
      if (following <= previous + EPS) continue;
      double2 next = make_double2(a.x + following * d.x, a.y + following * d.y);
      if (fabs(cross3(center, cursor, next)) > EPS * EPS) ntri += 1;
      cursor = next;
      previous = following;
    }
  }
  int total;
  int offset = wave_prefix(ntri, total);
  int base = 0;
  if ((threadIdx.x & 63) == 0 && total > 0) base = atomicAdd(w.status + ST_TRIS, total);
  base = __shfl(base, 0, 64);
  if (!valid || ntri == 0) return;
  // phase 2: emit
  int source = w.out_src[i];
  double4 sc = w.scores[source];
  double previous = 0.0;
  double2 cursor = a;
  int slot = base + offset;
  for (int step = 0; step <= count; ++step) {
    double following = step < count ? params[step] : 1.0;
    if (following <= previous + EPS) continue;
    double2 next = make_double2(a.x + following * d.x, a.y + following * d.y);
    if (fabs(cross3(center, cursor, next)) > EPS * EPS) {
      if (slot < w.cap.triangles) {
        double2 tri[3] = {center, cursor, next};
        double *uv = w.tri_uv + 6L * slot;
        double *world = w.tri_world + 9L * slot;
        for (int c = 0; c < 3; ++c) {
          uv[2 * c] = tri[c].x;
          uv[2 * c + 1] = tri[c].y;
          lift(w, tri[c], sc, world + 3 * c);
        }
        w.tri_src[slot] = source;
        w.tri_layer[slot] = layer;
      } else {
        set_error(w.status, ERR_TRIS);
      }
      slot += 1;
    }
    cursor = next;
    previous = following;
  }
}

static inline dim3 grid_for(long n, int block = 256) {
  long g = (n + block - 1) / block;
  return dim3((unsigned)(g < 1 ? 1 : g));
}

__global__ void k_check_vertices(Ws w) {
  int total = w.control[C_NBV];
  if (total > w.cap.vertices) set_error(w.status, ERR_VERTS);
  w.status[ST_BVERTS] = total > w.cap.vertices ? 0 : total;
}

void enqueue_mesh(const Ws &w, hipStream_t s, Timeline *t) {
  hipLaunchKernelGGL(k_vertex_counts, grid_for(w.cap.regions), dim3(256), 0, s, w);
  scan_exclusive(s, w.vcounts, w.voffsets, nullptr, w.cap.regions, w.cap.regions,
                 w.control + C_NBV, w.scan);
  hipLaunchKernelGGL(k_check_vertices, dim3(1), dim3(1), 0, s, w);
  hipLaunchKernelGGL(k_scatter, grid_for(w.cap.regions), dim3(256), 0, s, w);
  {
    long g = ((long)w.phash_size + 255) / 256;
    hipLaunchKernelGGL(k_phash_reset, dim3((unsigned)(g > 2048 ? 2048 : g)), dim3(256), 0, s, w);
  }
  hipLaunchKernelGGL(k_point_dedupe, grid_for(w.cap.vertices), dim3(256), 0, s, w);
  mark(t, s, "scatter_vertices");
  hipLaunchKernelGGL(k_point_flags, grid_for(w.cap.vertices), dim3(256), 0, s, w);
  scan_exclusive(s, w.pflag, w.poff, nullptr, w.cap.vertices, w.cap.vertices, w.control + C_NUP,
                 w.scan);
  hipLaunchKernelGGL(k_point_compact, grid_for(w.cap.vertices), dim3(256), 0, s, w);
  hipLaunchKernelGGL(k_point_count, dim3(1), dim3(1), 0, s, w);
  Hierarchy pt = w.point_tree;
  build_hierarchy(s, w.point_boxes, w.control + C_NUP, 0, w.cap.vertices, true, pt, w.sort);
  mark(t, s, "point_hierarchy");
  hipLaunchKernelGGL(k_mesh_edges, grid_for(w.cap.vertices), dim3(256), 0, s, w);
  mark(t, s, "k_mesh_edges");
}

}  // namespace peel
