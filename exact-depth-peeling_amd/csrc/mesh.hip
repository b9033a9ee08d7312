// Stage 4: boundary noding, centroid fans and world lifting.
//
// Reference semantics (file:line in alecjacobson/exact-depth-peeling):
//  vertex_counts / scatter  warp_peel/kernels.py:941-972
//  mesh_tasks               warp_peel/event_kernels.py:504-516
//  point_bounds + layered   warp_peel/kernels.py:975-991, events.py:744-754
//                           (replaced by a per-layer cell grid of vertices)
//  lift                     warp_peel/kernels.py:994-1006
//  mesh_one_edge            warp_peel/kernels.py:1011-1124
//
// One lane per polygon edge: a walk over the grid cells along the edge
// collects same-layer boundary vertices strictly inside the edge (t in
// (EPS, 1-EPS), distance <= EPS * length), sorts their parameters, and fans
// the noded edge from the polygon centroid. Up to 8 parameters stay in
// registers; an edge with more, or a long one, is walked by a whole wave.
#include "workspace.hpp"

namespace peel {

__device__ __forceinline__ int npolys_m(const Ws &w) {
  int p = load_i(w.status + ST_POLYS);
  return p < w.cap.regions ? p : w.cap.regions;
}

__device__ void scatter_one(const Ws &w, int i);

// After the exclusive scan of the polygons' vertex counts (their ring
// offsets, total in control word `total_word`): check and publish the
// boundary-vertex total, clear the live part of the point table for
// k_point_dedupe, and scatter every ring's vertices.
__global__ void k_scatter(Ws w, int total_word) {
  int total = load_i(w.control + total_word);
  int live = total > w.cap.vertices ? 0 : total;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    if (total > w.cap.vertices) set_error(w.status, ERR_VERTS);
    w.status[ST_BVERTS] = live;
  }
  int size = live_table(live, w.phash_size);
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < size; j += gridDim.x * blockDim.x) w.phash[j] = -1;
  if (w.status[ST_ERR] != 0) return;
  ITEM_LOOP(i, npolys_m(w)) scatter_one(w, i);
}

__device__ void scatter_one(const Ws &w, int i) {
  int n = w.out_n[i];
  int base = w.voffsets[i];
  if (n == 0) return;
  if (base + n > w.cap.vertices) {
    set_error(w.status, ERR_VERTS);
    return;
  }
  int layer = w.out_layer[i];
  long R = w.cap.regions;
  double2 center = make_double2(0.0, 0.0);
  // eight vertices are loaded before any is stored (a store between two loads
  // would make each load wait for the one before)
  for (int j0 = 0; j0 < n; j0 += 8) {
    double2 v[8];
#pragma unroll
    for (int r = 0; r < 8; ++r)
      if (j0 + r < n) v[r] = w.out_xy[(j0 + r) * R + i];
#pragma unroll
    for (int r = 0; r < 8; ++r) {
      const int j = j0 + r;
      if (j >= n) continue;
      center.x += v[r].x / (double)n;
      center.y += v[r].y / (double)n;
      w.bverts[base + j] = v[r];
      w.blayers[base + j] = layer;
      w.tasks[base + j] = make_int2(i, j | n << 16);
    }
  }
  w.centers[i] = center;
}

// Boundary vertices repeat once per incident polygon; only the first copy of
// each exact (layer, x, y) enters the vertex grid. Duplicates would give
// the same edge parameter, which the fan loop skips anyway, so the output is
// unchanged and each edge query sees each distinct vertex once. The
// edge_vertices capacity therefore bounds distinct vertices per edge, where
// the reference bounds copies (kernels.py:1051-1057): long contour cuts on
// the 65,536-face torus knot carry about 90 distinct vertices, about twice as
// many copies, which a bound of 128 copies would reject.
__device__ __forceinline__ uint64_t bvert_hash(double2 x, int layer) {
  return mix64(dbits(x.x) ^ mix64(dbits(x.y) + 0x9e3779b97f4a7c15ull) ^
               mix64((uint64_t)(uint32_t)layer + 0x3c6ef372fe94f82aull));
}

__device__ __forceinline__ int phash_live(const Ws &w) {
  return live_table(load_i(w.status + ST_BVERTS), w.phash_size);
}

__device__ void dedupe_one(const Ws &w, int i);
__device__ int grid_side(const Ws &w);
__device__ __forceinline__ long cells_live(const Ws &w, int G);

// Deduplicate the boundary vertices per layer; beside it, plan the per-layer
// cell grid from the vertex total and clear its live cells' counts.
__global__ void k_point_dedupe(Ws w) {
  int G = grid_side(w);
  long cells = cells_live(w, G);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    w.control[C_GRID] = G;
    w.control[C_CELLS] = (int)cells;
  }
  for (long j = (long)blockIdx.x * blockDim.x + threadIdx.x; j < cells; j += (long)gridDim.x * blockDim.x)
    w.cell_start[j] = 0;
  if (w.status[ST_ERR] != 0) return;
  ITEM_LOOP(i, load_i(w.status + ST_BVERTS)) dedupe_one(w, i);
}

__device__ void dedupe_one(const Ws &w, int i) {
  double2 x = w.bverts[i];
  int layer = w.blayers[i];
  int size = phash_live(w), mask = size - 1;
  int slot = (int)(bvert_hash(x, layer) & (uint64_t)mask);
  for (int probe = 0; probe < size; ++probe) {
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
  int size = phash_live(w), mask = size - 1;
  int slot = (int)(bvert_hash(x, layer) & (uint64_t)mask);
  for (int probe = 0; probe < size; ++probe) {
    int cur = w.phash[slot];
    if (cur < 0) return true;
    if (same(w.bverts[cur], x) && w.blayers[cur] == layer) return cur == i;
    slot = (slot + 1) & mask;
  }
  return true;
}

// ---------------------------------------------------------------------------
// Cell grid of the distinct boundary vertices, per layer: G x G cells over the
// unit window (G a power of two chosen on the device), cells numbered column
// by column, and the vertices stored cell by cell in that order (compressed
// rows: a count per cell, an exclusive scan, a fill). The cells of one column
// between two rows therefore hold one contiguous run of vertices, so an edge
// query reads one run per column its segment passes through (instead of a
// layered hierarchy over vertex boxes, which a long diagonal edge meets over
// its whole bounding box). layers * G^2 stays within the allocated cells.
__device__ __forceinline__ long cell_slot(int layer, int cx, int cy, int G) {
  return ((long)layer * G + cx) * G + cy;
}

__device__ __forceinline__ int cell_index(double v, int G) {
  double c = floor(v * (double)G);
  return c < 0.0 ? 0 : (c > (double)(G - 1) ? G - 1 : (int)c);
}

// Cells per axis of each layer's grid: about two boundary vertices per cell,
// within the cell pool.
__device__ int grid_side(const Ws &w) {
  int layers = max(1, w.status[ST_LAYERS]);
  double per_layer = (double)load_i(w.status + ST_BVERTS) / (2.0 * (double)layers);
  int G = 8;
  while (G < 4096 && (double)G * (double)G < per_layer) G *= 2;
  while (G > 1 && (long)layers * G * G > w.cell_capacity) G /= 2;
  return G;
}

__device__ __forceinline__ long cells_live(const Ws &w, int G) {
  return (long)max(1, w.status[ST_LAYERS]) * G * G;
}

// Count every first copy of a (layer, point) in its cell (and remember the
// cell); an exclusive scan of the counts gives each cell's first row, and
// k_cell_fill stores the points cell by cell.
__global__ void k_cell_count(Ws w) {
  long long unique = 0;
  if (w.status[ST_ERR] == 0) {
    int G = load_i(w.control + C_GRID);
    ITEM_LOOP(i, load_i(w.status + ST_BVERTS)) {
      int slot = -1;
      if (first_copy(w, i)) {
        double2 x = w.bverts[i];
        slot = (int)cell_slot(w.blayers[i], cell_index(x.x, G), cell_index(x.y, G), G);
        atomicAdd(w.cell_start + slot, 1);
        unique += 1;
      }
      w.cell_of[i] = slot;
    }
  }
  metric_add(w, M_UNIQUE_BVERTS, unique);
}

// Each cell's first row advances past its points as they are stored, so a
// filled cell c ends at cell_start[c] and starts where cell c - 1 ends.
__global__ void k_cell_fill(Ws w) {
  if (w.status[ST_ERR] != 0) return;
  ITEM_LOOP(i, load_i(w.status + ST_BVERTS)) {
    int slot = w.cell_of[i];
    if (slot >= 0) w.cell_xy[atomicAdd(w.cell_start + slot, 1)] = w.bverts[i];
  }
}

// Visit every distinct vertex of `layer` in a cell that can hold a point within
// PAD of the segment p0-p1: each column the padded x-range touches and, in it,
// each row that the segment's y-range over the padded column (itself padded)
// touches. A vertex accepted by the edge test lies within EPS of the segment,
// far inside PAD, so it is visited; each cell, and so each vertex, once.
// A column's rows are one run of vertices, read SCAN_BATCH at a time with the
// loads issued together (one memory latency per batch, not per vertex).
constexpr double NODE_PAD = 2.0e-7;
constexpr int SCAN_BATCH = 4;

template <class Visit>
__device__ __forceinline__ bool scan_run(const Ws &w, int b, int e, Visit &visit) {
  for (int k = b; k < e; k += SCAN_BATCH) {
    double2 x[SCAN_BATCH];
    // past the run's end the batch repeats its last vertex, which is skipped
#pragma unroll
    for (int r = 0; r < SCAN_BATCH; ++r) x[r] = w.cell_xy[min(k + r, e - 1)];
#pragma unroll
    for (int r = 0; r < SCAN_BATCH; ++r)
      if (k + r < e && !visit(x[r])) return false;
  }
  return true;
}

// The columns a walk over segment p0-p1 passes through, and in column cx the
// run of cell_xy rows [b, e) of the rows it touches there.
struct SegmentCells {
  double2 p0;
  double xa, xb, ya, yb, dx, dy;
  int layer, G, cx0, cx1;
  __device__ SegmentCells(double2 a, double2 b, int layer_, int G_) : p0(a), layer(layer_), G(G_) {
    xa = fmin(a.x, b.x);
    xb = fmax(a.x, b.x);
    ya = fmin(a.y, b.y);
    yb = fmax(a.y, b.y);
    dx = b.x - a.x;
    dy = b.y - a.y;
    cx0 = cell_index(xa - NODE_PAD, G);
    cx1 = cell_index(xb + NODE_PAD, G);
  }
  __device__ int2 run(const Ws &w, int cx) const {
    double xl = fmax(xa, (double)cx / G - NODE_PAD), xr = fmin(xb, (double)(cx + 1) / G + NODE_PAD);
    double ylo = ya, yhi = yb;
    if (dx != 0.0 && xl <= xr) {
      double y0 = p0.y + (xl - p0.x) * (dy / dx), y1 = p0.y + (xr - p0.x) * (dy / dx);
      ylo = fmax(ya, fmin(y0, y1));
      yhi = fmin(yb, fmax(y0, y1));
    }
    int cy0 = cell_index(ylo - NODE_PAD, G), cy1 = cell_index(yhi + NODE_PAD, G);
    long first = cell_slot(layer, cx, cy0, G), last = cell_slot(layer, cx, cy1, G);
    return make_int2(first > 0 ? w.cell_start[first - 1] : 0, w.cell_start[last]);
  }
};

template <class Visit>
__device__ __forceinline__ void walk_segment(const Ws &w, double2 p0, double2 p1, int layer, int G, Visit visit) {
  const SegmentCells cells(p0, p1, layer, G);
  for (int cx = cells.cx0; cx <= cells.cx1; ++cx) {
    int2 r = cells.run(w, cx);
    if (!scan_run(w, r.x, r.y, visit)) return;
  }
}

// Cells of the segment's padded bounding box, a bound on the cells walked.
__device__ __forceinline__ int box_cells(double2 a, double2 b, int G) {
  int cols = cell_index(fmax(a.x, b.x) + NODE_PAD, G) - cell_index(fmin(a.x, b.x) - NODE_PAD, G) + 1;
  int rows = cell_index(fmax(a.y, b.y) + NODE_PAD, G) - cell_index(fmin(a.y, b.y) - NODE_PAD, G) + 1;
  return cols + rows;
}

// Window point -> world point: the window's range and the inverse camera are
// read once per kernel into registers (a triangle's stores between two lifts
// would otherwise make every lift read them again, one memory latency each).
struct Lifter {
  double s0, s1, s2, s3, m[12];
  bool persp;
  __device__ explicit Lifter(const Ws &w) {
    const double *s = w.settings;
    s0 = s[0];
    s1 = s[1];
    s2 = s[2];
    s3 = s[3];
    persp = s[4] != 0.0;
#pragma unroll
    for (int k = 0; k < 12; ++k) m[k] = w.inverse[k];
  }
  __device__ void operator()(double2 x, double4 sc, double *out) const {
    double u = s0 + x.x * (s2 - s0);
    double v = s1 + x.y * (s3 - s1);
    double z = value(sc, x);
    if (persp) {
      z = -1.0 / z;
      u *= z;
      v *= z;
    }
#pragma unroll
    for (int r = 0; r < 3; ++r) out[r] = m[4 * r] * u + m[4 * r + 1] * v + m[4 * r + 2] * z + m[4 * r + 3] * 1.0;
  }
};

constexpr int EDGE_MAX = 128;

// Parameters of the same-layer vertices strictly inside edge (a, b):
// t in (EPS, 1 - EPS) and distance <= EPS * |b - a| (kernels.py:1040-1064).
// The distance test, a few products, goes first: almost every vertex a walk
// reads is off the line, and t (a division) is formed only for the others.
struct EdgeQuery {
  double2 a, b, d;
  double length2, len;
  __device__ bool accept(double2 x, double &t) const {
    if (!(fabs(cross3(a, b, x)) <= EPS * len)) return false;
    t = ((x.x - a.x) * d.x + (x.y - a.y) * d.y) / length2;
    return t > EPS && t < 1.0 - EPS;
  }
};

constexpr int REG_PARAMS = 8;

// Sorting network on 8 registers (Batcher odd-even merge), static indices only.
__device__ __forceinline__ void cswap(double &x, double &y) {
  double lo = fmin(x, y), hi = fmax(x, y);
  x = lo;
  y = hi;
}
__device__ __forceinline__ void sort8(double (&p)[REG_PARAMS]) {
  cswap(p[0], p[1]); cswap(p[2], p[3]); cswap(p[4], p[5]); cswap(p[6], p[7]);
  cswap(p[0], p[2]); cswap(p[1], p[3]); cswap(p[4], p[6]); cswap(p[5], p[7]);
  cswap(p[1], p[2]); cswap(p[5], p[6]); cswap(p[0], p[4]); cswap(p[3], p[7]);
  cswap(p[1], p[5]); cswap(p[2], p[6]);
  cswap(p[1], p[4]); cswap(p[3], p[6]);
  cswap(p[2], p[4]); cswap(p[3], p[5]);
  cswap(p[3], p[4]);
}

// Fan walk over the sorted parameters: counts (emit == false) or writes
// the triangles starting at `slot`.
template <class Param>
__device__ __forceinline__ int fan(const Ws &w, const Lifter &lift, const EdgeQuery &q, double2 center,
                                   int count, Param param, bool emit, int slot, int source, int layer,
                                   double4 sc) {
  double previous = 0.0;
  double2 cursor = q.a;
  int ntri = 0;
  for (int step = 0; step <= count; ++step) {
    double following = step < count ? param(step) : 1.0;
    if (following <= previous + EPS) continue;
    double2 next = make_double2(q.a.x + following * q.d.x, q.a.y + following * q.d.y);
    if (fabs(cross3(center, cursor, next)) > EPS * EPS) {
      if (emit) {
        int k = slot + ntri;
        if (k < w.cap.triangles) {
          double2 tri[3] = {center, cursor, next};
          double *uv = w.tri_uv + 6L * k;
          double *world = w.tri_world + 9L * k;
          for (int c = 0; c < 3; ++c) {
            uv[2 * c] = tri[c].x;
            uv[2 * c + 1] = tri[c].y;
            lift(tri[c], sc, world + 3 * c);
          }
          w.tri_src[k] = source;
          w.tri_layer[k] = layer;
        } else {
          set_error(w.status, ERR_TRIS);
        }
      }
      ntri += 1;
    }
    cursor = next;
    previous = following;
  }
  return ntri;
}

// Edge j of the scattered rings: its polygon's vertices are bverts[j - edge,
// j - edge + n), and k_scatter left the polygon's centroid (summed in ring
// order) in centers.
__device__ __forceinline__ bool edge_setup(const Ws &w, int j, EdgeQuery &q, double2 &center, int &i,
                                           int &layer) {
  int2 task = w.tasks[j];
  i = task.x;
  const int edge = task.y & 0xFFFF, n = task.y >> 16;
  if (edge >= n) return false;
  q.a = w.bverts[j];
  q.b = w.bverts[edge + 1 < n ? j + 1 : j - edge];
  q.d = make_double2(q.b.x - q.a.x, q.b.y - q.a.y);
  q.length2 = q.d.x * q.d.x + q.d.y * q.d.y;
  q.len = sqrt(q.length2);
  center = w.centers[i];
  layer = w.blayers[j];
  return true;
}

// One lane per polygon edge (kernels.py:1011-1124). Up to REG_PARAMS edge
// parameters stay in registers (static indices, sorting network); an edge with
// more is queued for k_mesh_crowded. Phase 1 counts the fan triangles, one
// atomic per wave reserves the wave's slots, phase 2 writes uv and lifted
// world corners.
__device__ void mesh_edge_one(const Ws &w, int j, int n);

// Four waves per SIMD (at most 128 VGPRs): the walk is bound by memory
// latency, so residency sets the throughput.
__global__ __launch_bounds__(256, 4) void k_mesh_edges(Ws w) {
  // the whole block leaves or stays (mesh_edge_one synchronises the block;
  // an error bit set by another block meanwhile is seen by one read only)
  __shared__ int failed;
  if (threadIdx.x == 0) failed = load_i(w.status + ST_ERR);
  __syncthreads();
  if (failed) return;
  int n = load_i(w.status + ST_BVERTS);
  WAVE_LOOP(base, n) mesh_edge_one(w, base + threadIdx.x, n);
}

__device__ void mesh_edge_one(const Ws &w, int j, int n) {
  bool valid = j < n;
  EdgeQuery q;
  double2 center = make_double2(0, 0);
  int i = 0, layer = 0, count = 0;
  bool crowded = false;
  double p[REG_PARAMS];
#pragma unroll
  for (int k = 0; k < REG_PARAMS; ++k) p[k] = 2.0;
  if (valid) valid = edge_setup(w, j, q, center, i, layer);
  // the polygon's source is read now; its wait falls after the walk
  const int source = valid ? w.out_src[i] : 0;
  if (valid && q.length2 > 0.0) {
    const int G = load_i(w.control + C_GRID);
    const SegmentCells cells(q.a, q.b, layer, G);
    // long edges go straight to the wave-parallel path
    crowded = box_cells(q.a, q.b, G) > w.long_edge_cells;
    auto visit = [&](double2 x) {
      double t;
      if (q.accept(x, t)) {
#pragma unroll
        for (int r = 0; r < REG_PARAMS; ++r) p[r] = (r == count) ? t : p[r];
        count += 1;
        if (count > REG_PARAMS) {
          crowded = true;
          return false;
        }
      }
      return true;
    };
    for (int cx = cells.cx0; !crowded && cx <= cells.cx1; ++cx) {
      int2 r = cells.run(w, cx);
      if (!scan_run(w, r.x, r.y, visit)) break;
    }
    sort8(p);
  }
  // read here, after the walk has released its registers
  const Lifter lift(w);
  int slot = wave_claim32(w.control + C_CROWDED, crowded);
  if (crowded) {
    if (slot < w.cap.vertices) w.crowded[slot] = j;
    valid = false;
  }
  int top = wave_max(valid ? count : 0);
  if (valid && count > w.cap.edge_vertices) {
    set_error(w.status, ERR_EDGE);
    valid = false;
  }
  metric_max(w, M_EDGE_MAX, top);
  auto reg = [&](int k) {
    double v = p[0];
#pragma unroll
    for (int r = 1; r < REG_PARAMS; ++r) v = (k == r) ? p[r] : v;
    return v;
  };
  double4 sc = valid ? w.scores[source] : make_double4(0, 0, 0, 0);
  int ntri = valid ? fan(w, lift, q, center, count, reg, false, 0, 0, 0, sc) : 0;
  // one slot claim per block (returning atomics on one word serialise)
  __shared__ int wave_tris[4], block_base;
  const int wave = threadIdx.x >> 6;
  int total;
  int offset = wave_prefix(ntri, total);
  if ((threadIdx.x & 63) == 0) wave_tris[wave] = total;
  __syncthreads();
  int all = 0;
  for (int k = 0; k < 4; ++k) {
    if (k < wave) offset += wave_tris[k];
    all += wave_tris[k];
  }
  if (threadIdx.x == 0) block_base = all ? atomicAdd(w.status + ST_TRIS, all) : 0;
  __syncthreads();
  int base = block_base;
  __syncthreads();  // the next edge of this block reuses the words
  if (!valid || ntri == 0) return;
  fan(w, lift, q, center, count, reg, true, base + offset, source, layer, sc);
}

// Crowded edges, one wave each, the whole edge read by all its lanes: the
// walk's column runs (one per column, up to 64 columns at a time, a column per
// lane) are laid end to end, and the lanes take the rows of that sequence 64 at
// a time, so a dense cell or a long edge is spread over the wave. Accepted
// parameters are appended to a list in LDS and ranked into sorted order there;
// the fan is then built from the sorted list. Up to edge_vertices parameters;
// more sets error bit 128 (kernels.py:1051-1057).
__global__ __launch_bounds__(256) void k_mesh_crowded(Ws w) {
  __shared__ double found_lds[4][EDGE_MAX];  // accepted parameters, in the order found
  __shared__ double lds[4][EDGE_MAX];        // the same, sorted
  __shared__ int kept_lds[4][EDGE_MAX + 2];  // fan steps the skip rule keeps, then their count
  __shared__ int run_row[4][64], run_at[4][64];  // first row of each lane's run, its place in the sequence
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double *list = lds[wave], *found = found_lds[wave];
  int *row0 = run_row[wave], *at = run_at[wave];
  int n = load_i(w.control + C_CROWDED);
  if (n > w.cap.vertices) n = w.cap.vertices;
  const Lifter lift(w);
  int waves = gridDim.x * 4;
  for (int t = blockIdx.x * 4 + wave; t < n; t += waves) {
    int j = w.crowded[t];
    EdgeQuery q;
    double2 center;
    int i, layer;
    if (!edge_setup(w, j, q, center, i, layer)) continue;  // wave-uniform
    const int G = load_i(w.control + C_GRID);
    int total = 0;
    if (q.length2 > 0.0) {
      const SegmentCells cells(q.a, q.b, layer, G);
      for (int c0 = cells.cx0; c0 <= cells.cx1; c0 += 64) {
        int cx = c0 + lane, b = 0, e = 0;
        if (cx <= cells.cx1) {
          int2 r = cells.run(w, cx);
          b = r.x;
          e = r.y;
        }
        int rows;
        int place = wave_prefix(max(e - b, 0), rows);
        row0[lane] = b;
        at[lane] = place;
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
        __builtin_amdgcn_wave_barrier();
        for (int f0 = 0; f0 < rows; f0 += 64) {
          int f = f0 + lane;
          bool ok = false;
          double tt = 0.0;
          if (f < rows) {
            // the last run starting at or before f holds it (empty runs share
            // their successor's place)
            int r = 0;
#pragma unroll
            for (int step = 32; step > 0; step >>= 1)
              if (at[r + step] <= f) r += step;
            ok = q.accept(w.cell_xy[row0[r] + (f - at[r])], tt);
          }
          uint64_t m = __ballot(ok);
          int pos = total + __popcll(m & lanemask_lt());
          if (ok && pos < EDGE_MAX) found[pos] = tt;
          total += __popcll(m);
        }
        __builtin_amdgcn_wave_barrier();  // the run table is rewritten for the next 64 columns
      }
    }
    metric_max(w, M_EDGE_MAX, total);
    if (total > w.cap.edge_vertices) {
      if (lane == 0) set_error(w.status, ERR_EDGE);
      continue;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    // rank sort: each parameter's place is the number of parameters before it
    // in (value, order found)
    for (int s = lane; s < total; s += 64) {
      double v = found[s];
      int rank = 0;
      for (int k = 0; k < total; ++k) {
        double u = found[k];
        rank += (u < v || (u == v && k < s)) ? 1 : 0;
      }
      list[rank] = v;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    int source = w.out_src[i];
    double4 sc = w.scores[source];
    // The fan steps run across the lanes when no two consecutive parameters
    // (0 and 1 included) lie within 2 EPS: then the sequential skip rule keeps
    // every step, and step s joins the points at parameters s-1 and s, computed
    // exactly as the sequential walk computes them. Otherwise lane 0 applies
    // the skip rule alone (a few LDS reads per step) and the lanes emit the
    // kept steps, each joining the points of the previous kept parameter and
    // its own, as the sequential walk does.
    const int steps = total + 1;
    bool tight = false;
    for (int st = lane; st < steps; st += 64) {
      double prev = st == 0 ? 0.0 : list[st - 1], cur = st < total ? list[st] : 1.0;
      tight = tight || cur <= prev + 2.0 * EPS;
    }
    // kept steps: all of them, or those lane 0's pass of the skip rule keeps
    const bool any_tight = __ballot(tight) != 0;
    int *kept = kept_lds[wave];
    if (any_tight && lane == 0) {
      double previous = 0.0;
      int m = 0;
      for (int st = 0; st < steps; ++st) {
        double following = st < total ? list[st] : 1.0;
        if (following <= previous + EPS) continue;
        kept[m++] = st;
        previous = following;
      }
      kept[EDGE_MAX + 1] = m;
    }
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    const int m = any_tight ? kept[EDGE_MAX + 1] : steps;
    for (int round = 0; round < m; round += 64) {
      int j = round + lane;
      bool emit = false;
      double2 cursor = q.a, next = q.a;
      if (j < m) {
        int st = any_tight ? kept[j] : j;
        double cur = st < total ? list[st] : 1.0;
        next = make_double2(q.a.x + cur * q.d.x, q.a.y + cur * q.d.y);
        if (j > 0) {
          int ps = any_tight ? kept[j - 1] : j - 1;
          double prev = ps < total ? list[ps] : 1.0;
          cursor = make_double2(q.a.x + prev * q.d.x, q.a.y + prev * q.d.y);
        }
        emit = fabs(cross3(center, cursor, next)) > EPS * EPS;
      }
      int slot = wave_claim32(w.status + ST_TRIS, emit);
      if (emit) {
        if (slot < w.cap.triangles) {
          double2 tri[3] = {center, cursor, next};
          double *uv = w.tri_uv + 6L * slot;
          double *world = w.tri_world + 9L * slot;
          for (int c = 0; c < 3; ++c) {
            uv[2 * c] = tri[c].x;
            uv[2 * c + 1] = tri[c].y;
            lift(tri[c], sc, world + 3 * c);
          }
          w.tri_src[slot] = source;
          w.tri_layer[slot] = layer;
        } else {
          set_error(w.status, ERR_TRIS);
        }
      }
    }
    __builtin_amdgcn_wave_barrier();
  }
}

static inline dim3 grid_for(long n, int block = 256) {
  long g = (n + block - 1) / block;
  return dim3((unsigned)(g < 1 ? 1 : g));
}

// The metric rows folded into `metrics` (sums, and the maximum for the
// largest edge candidate count), after every kernel that adds to them.
__global__ void k_metrics_fold(Ws w) {
  const int m = threadIdx.x;
  if (m >= M_COUNT) return;
  const bool maximum = m == M_EDGE_MAX;
  long long acc = 0;
#pragma unroll 8
  for (int r = 0; r < METRIC_ROWS; ++r) {
    long long x = w.metric_rows[(long)r * METRIC_STRIDE + m];
    acc = maximum ? max(acc, x) : acc + x;
  }
  w.metrics[m] = maximum ? max(w.metrics[m], acc) : w.metrics[m] + acc;
}

void enqueue_mesh(const Ws &w, hipStream_t s, Timeline *t) {
  // ring offsets: an exclusive scan of the polygons' vertex counts. The
  // connectivity stage scanned the same counts into the same offsets; with
  // every layer kept (no ring is emptied by the layer limit) and no merge,
  // the rings are unchanged since, and its offsets and total serve here.
  const bool scanned = w.route != ROUTE_LAYERED && w.propagate_ranks && !w.merge_regions && w.all_layers;
  if (!scanned)
    scan_exclusive(s, w.out_n, w.voffsets, w.status + ST_POLYS, 0, w.cap.regions, w.control + C_NBV, w.scan);
  hipLaunchKernelGGL(k_scatter, stride_grid(w.cap.regions), dim3(256), 0, s, w, scanned ? (int)C_PEDGES : (int)C_NBV);
  hipLaunchKernelGGL(k_point_dedupe, stride_grid(w.cap.vertices), dim3(256), 0, s, w);
  mark(t, s, "scatter_vertices");
  hipLaunchKernelGGL(k_cell_count, stride_grid(w.cap.vertices), dim3(256), 0, s, w);
  scan_exclusive(s, w.cell_start, w.cell_start, w.control + C_CELLS, 0, (int)w.cell_capacity, nullptr, w.scan);
  hipLaunchKernelGGL(k_cell_fill, stride_grid(w.cap.vertices), dim3(256), 0, s, w);
  mark(t, s, "vertex_grid");
  hipLaunchKernelGGL(k_mesh_edges, stride_grid(w.cap.vertices), dim3(256), 0, s, w);
  mark(t, s, "k_mesh_edges");
  hipLaunchKernelGGL(k_mesh_crowded, dim3(1024), dim3(256), 0, s, w);
  mark(t, s, "k_mesh_crowded");
  hipLaunchKernelGGL(k_metrics_fold, dim3(1), dim3(64), 0, s, w);
}

}  // namespace peel
