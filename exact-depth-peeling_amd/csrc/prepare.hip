// Stage 1: reset, input validation, exact welding, projection, contour
// topology, contour segment grouping and the two static hierarchies.
//
// Reference semantics (file:line in alecjacobson/exact-depth-peeling):
//  reset            warp_peel/kernels.py:253-266
//  validate_input   warp_peel/kernels.py:309-350
//  weld             warp_peel/kernels.py:269-306 (stable-sort first equal vertex)
//  project          warp_peel/kernels.py:353-468
//  edge topology    warp_peel/kernels.py:471-536
//  contour segments warp_peel/event_kernels.py:388-501
//
// Welding and the two groupings use open-addressing hash tables instead of the
// reference's chains of stable radix sorts: the representative of a group is
// the minimum index (atomicMin), which is exactly "first equal element of a
// stable sort that starts from the identity permutation", in one pass.
#include "workspace.hpp"

namespace peel {

__global__ void k_reset(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < 16) w.status[i] = 0;
  if (i < M_COUNT) w.metrics[i] = 0;
  for (int j = i; j < METRIC_ROWS * METRIC_STRIDE; j += gridDim.x * blockDim.x) w.metric_rows[j] = 0;
  if (i < C_COUNT) w.control[i] = 0;
  for (int j = i; j < w.vhash_size; j += gridDim.x * blockDim.x) w.vhash[j] = -1;
  for (int j = i; j < w.ehash_size; j += gridDim.x * blockDim.x) w.etable[j] = EdgeSlot{~0ull, 0ull};
  for (int j = i; j < w.shash_size; j += gridDim.x * blockDim.x) w.shash[j] = -1;
  for (int j = i; j < 2 * w.nseg; j += gridDim.x * blockDim.x) w.seg_near[j] = ~0ull;
}

__global__ void k_validate(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool bad = false;
  if (i < w.nv) {
    for (int j = 0; j < 3; ++j) bad |= !isfinite(w.vertices[3 * i + j]);
  }
  if (i < w.nf) {
    for (int j = 0; j < 3; ++j) {
      int f = w.faces[3 * i + j];
      bad |= f < 0 || f >= w.nv;
    }
  }
  if (i == 0) {
    const double *m = w.camera;
    for (int k = 0; k < 16; ++k) bad |= !isfinite(m[k]);
    for (int k = 0; k < 6; ++k) bad |= !isfinite(w.settings[k]);
    // determinant by cofactor expansion along the last row
    double det = 0.0;
    {
      double a[4][4];
      for (int r = 0; r < 4; ++r)
        for (int c = 0; c < 4; ++c) a[r][c] = m[4 * r + c];
      for (int c = 0; c < 4; ++c) {
        double minor[3][3];
        for (int r = 0; r < 3; ++r) {
          int cc = 0;
          for (int k = 0; k < 4; ++k)
            if (k != c) minor[r][cc++] = a[r][k];
        }
        double d3 = minor[0][0] * (minor[1][1] * minor[2][2] - minor[1][2] * minor[2][1]) -
                    minor[0][1] * (minor[1][0] * minor[2][2] - minor[1][2] * minor[2][0]) +
                    minor[0][2] * (minor[1][0] * minor[2][1] - minor[1][1] * minor[2][0]);
        double sign = ((3 + c) & 1) ? -1.0 : 1.0;
        det += sign * a[3][c] * d3;
      }
    }
    bad |= det == 0.0 || m[12] != 0.0 || m[13] != 0.0 || m[14] != 0.0 || m[15] != 1.0;
    const double *s = w.settings;
    bad |= s[2] <= s[0] || s[3] <= s[1] || s[5] <= 0.0 || (s[4] != 0.0 && s[4] != 1.0);
    // inverse affine camera for lifting (rows 0..2 of [R t; 0 1]^-1)
    double R[3][3], inv[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) R[r][c] = m[4 * r + c];
    double d = R[0][0] * (R[1][1] * R[2][2] - R[1][2] * R[2][1]) -
               R[0][1] * (R[1][0] * R[2][2] - R[1][2] * R[2][0]) +
               R[0][2] * (R[1][0] * R[2][1] - R[1][1] * R[2][0]);
    inv[0][0] = (R[1][1] * R[2][2] - R[1][2] * R[2][1]) / d;
    inv[0][1] = (R[0][2] * R[2][1] - R[0][1] * R[2][2]) / d;
    inv[0][2] = (R[0][1] * R[1][2] - R[0][2] * R[1][1]) / d;
    inv[1][0] = (R[1][2] * R[2][0] - R[1][0] * R[2][2]) / d;
    inv[1][1] = (R[0][0] * R[2][2] - R[0][2] * R[2][0]) / d;
    inv[1][2] = (R[0][2] * R[1][0] - R[0][0] * R[1][2]) / d;
    inv[2][0] = (R[1][0] * R[2][1] - R[1][1] * R[2][0]) / d;
    inv[2][1] = (R[0][1] * R[2][0] - R[0][0] * R[2][1]) / d;
    inv[2][2] = (R[0][0] * R[1][1] - R[0][1] * R[1][0]) / d;
    for (int r = 0; r < 3; ++r) {
      double t = -(inv[r][0] * m[3] + inv[r][1] * m[7] + inv[r][2] * m[11]);
      for (int c = 0; c < 3; ++c) w.inverse[4 * r + c] = inv[r][c];
      w.inverse[4 * r + 3] = t;
    }
    w.inverse[12] = w.inverse[13] = w.inverse[14] = 0.0;
    w.inverse[15] = 1.0;
  }
  if (bad) set_error(w.status, ERR_INPUT);
}

// ---------------------------------------------------------------------------
// Exact welding (warp_peel/kernels.py:290-306): canonical[v] = smallest index
// of a vertex with equal coordinates (+0 == -0).
__device__ __forceinline__ uint64_t vertex_hash(double x, double y, double z) {
  return mix64(dbits(x) ^ mix64(dbits(y) + 0x9e3779b97f4a7c15ull) ^
               mix64(dbits(z) + 0x3c6ef372fe94f82aull));
}

__device__ __forceinline__ bool same_vertex(const double *v, int a, int b) {
  return v[3 * a] == v[3 * b] && v[3 * a + 1] == v[3 * b + 1] && v[3 * a + 2] == v[3 * b + 2];
}

__global__ void k_weld_insert(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.nv) return;
  const double *v = w.vertices;
  int mask = w.vhash_size - 1;
  int slot = (int)(vertex_hash(v[3 * i], v[3 * i + 1], v[3 * i + 2]) & (uint64_t)mask);
  for (int probe = 0; probe < w.vhash_size; ++probe) {
    int cur = load_i(&w.vhash[slot]);
    if (cur < 0) {
      cur = atomicCAS(&w.vhash[slot], -1, i);
      if (cur < 0) return;
    }
    if (same_vertex(v, cur, i)) {
      if (i < cur) atomicMin(&w.vhash[slot], i);
      return;
    }
    slot = (slot + 1) & mask;
  }
}

__global__ void k_weld_resolve(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.nv) return;
  const double *v = w.vertices;
  int mask = w.vhash_size - 1;
  int slot = (int)(vertex_hash(v[3 * i], v[3 * i + 1], v[3 * i + 2]) & (uint64_t)mask);
  int result = i;
  for (int probe = 0; probe < w.vhash_size; ++probe) {
    int cur = w.vhash[slot];
    if (cur < 0) break;
    if (same_vertex(v, cur, i)) {
      result = cur;
      break;
    }
    slot = (slot + 1) & mask;
  }
  w.canonical[i] = result;
}

// ---------------------------------------------------------------------------
// Ring helpers for projection (rings of at most 8 vertices, kept in registers
// by full unrolling). Semantics of warp_peel/kernels.py:146-232.
struct Ring8 {
  double2 v[SRCV + 1];
  int n;
};

__device__ __forceinline__ void ring_append(Ring8 &r, double2 x) {
  if (r.n < SRCV + 1) r.v[r.n] = x;
  r.n += 1;
}

__device__ void ring_clean(Ring8 &p) {
  // warp_peel/kernels.py:146-176
  if (p.n > SRCV + 1) return;
  Ring8 q;
  q.n = 0;
  for (int j = 0; j < p.n; ++j) {
    double2 x = p.v[j];
    if (q.n == 0 || dlength(make_double2(x.x - q.v[q.n - 1].x, x.y - q.v[q.n - 1].y)) > EPS * 0.01)
      ring_append(q, x);
  }
  if (q.n > 1 && dlength(make_double2(q.v[0].x - q.v[q.n - 1].x, q.v[0].y - q.v[q.n - 1].y)) <= EPS * 0.01)
    q.n -= 1;
  if (q.n < 3) {
    q.n = 0;
  } else {
    double2 origin = q.v[0];
    double area2 = 0.0, perimeter = 0.0;
    for (int j = 0; j < q.n; ++j) {
      double2 a = q.v[j], b = q.v[(j + 1) % q.n];
      area2 += cross3(origin, a, b);
      perimeter += fabs(b.x - a.x) + fabs(b.y - a.y);
    }
    if (area2 <= EPS * perimeter) q.n = 0;
  }
  p = q;
}

// Keep the part with s(x) >= 0 (inside of split(), warp_peel/kernels.py:184-203).
__device__ void ring_keep(Ring8 &p, double sx, double sy, double sc) {
  Ring8 in;
  in.n = 0;
  for (int j = 0; j < p.n && j < SRCV + 1; ++j) {
    double2 a = p.v[j], b = p.v[(j + 1) % p.n];
    double da = sx * (a.x - 0.0) + sy * (a.y - 0.0) + sc;
    double db = sx * (b.x - 0.0) + sy * (b.y - 0.0) + sc;
    if (da >= 0.0) ring_append(in, a);
    if ((da > 0.0 && db < 0.0) || (da < 0.0 && db > 0.0)) {
      double t = da / (da - db);
      ring_append(in, make_double2(a.x + t * (b.x - a.x), a.y + t * (b.y - a.y)));
    }
  }
  ring_clean(in);
  p = in;
}

// The common case of k_project: a triangle in front of the near plane whose
// projection lies inside the window. Every clip of the general path then
// keeps the ring as it is, and its cleaning reduces to the checks below, so
// this computes the same values with the three vertices in registers (the
// general path's run-time indexed rings live in scratch memory). Returns
// false, having written nothing, when a vertex leaves the window.
__device__ __forceinline__ bool project_unclipped(const Ws &w, int i, const double (&c)[3][3], bool persp) {
  const double *s = w.settings;
  double2 v[3];
  double depth[3];
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    double invz = 1.0;
    if (persp) invz = 1.0 / c[j][2];
    v[j] = make_double2((c[j][0] * invz - s[0]) / (s[2] - s[0]), (c[j][1] * invz - s[1]) / (s[3] - s[1]));
    depth[j] = persp ? -invz : c[j][2];
    if (!(v[j].x >= 0.0 && v[j].x <= 1.0 && v[j].y >= 0.0 && v[j].y <= 1.0)) return false;
  }
  const int nf = w.nf;
  double4 score = make_double4(0, 0, 0, 0);
  double det = cross3(v[0], v[1], v[2]);
  int face = 0, n = 0;
  double2 u[3] = {v[0], v[1], v[2]};
  if (det != 0.0) {
    double2 pa = make_double2(v[1].x - v[0].x, v[1].y - v[0].y);
    double2 pb = make_double2(v[2].x - v[0].x, v[2].y - v[0].y);
    double s0 = ((depth[1] - depth[0]) * pb.y - (depth[2] - depth[0]) * pa.y) / det;
    double s1 = (pa.x * (depth[2] - depth[0]) - pb.x * (depth[1] - depth[0])) / det;
    score = make_double4(s0, s1, depth[0] - s0 * v[0].x - s1 * v[0].y, 0.0);
    face = 1;
    if (det < 0.0) {
      face = -1;
      u[0] = v[2];
      u[2] = v[0];
    }
    // ring_clean of a triangle: consecutive and closing vertices apart, then
    // the area test, summed in ring order as ring_clean sums it
    const double gap = EPS * 0.01;
    bool apart = dlength(make_double2(u[1].x - u[0].x, u[1].y - u[0].y)) > gap &&
                 dlength(make_double2(u[2].x - u[1].x, u[2].y - u[1].y)) > gap &&
                 !(dlength(make_double2(u[0].x - u[2].x, u[0].y - u[2].y)) <= gap);
    if (apart) {
      double area2 = 0.0, perimeter = 0.0;
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        double2 a = u[j], b = u[(j + 1) % 3];
        area2 += cross3(u[0], a, b);
        perimeter += fabs(b.x - a.x) + fabs(b.y - a.y);
      }
      if (!(area2 <= EPS * perimeter)) n = 3;
    }
  }
  if (!isfinite(score.x) || !isfinite(score.y) || !isfinite(score.z) || fabs(score.x) > 1e28 ||
      fabs(score.y) > 1e28 || fabs(score.z) > 1e28)
    set_error(w.status, ERR_INPUT);
  w.facing[i] = face;
  w.clipped[i] = 0;
  w.scores[i] = score;
  w.src_n[i] = n;
  double lx = 1e30, ly = 1e30, lz = 1e30, hx = -1e30, hy = -1e30, hz = -1e30;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    if (j < n) {
      double2 x = u[j];
      w.src_xy[j * nf + i] = x;
      double z = value(score, x);
      lx = fmin(lx, x.x), ly = fmin(ly, x.y), lz = fmin(lz, z);
      hx = fmax(hx, x.x), hy = fmax(hy, x.y), hz = fmax(hz, z);
    }
  }
  Box b;
  pad_box(lx, ly, lz, hx, hy, hz, b.lo, b.hi);
  w.src_boxes[i] = b;
  return true;
}

__global__ void k_project(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.nf) return;
  int nf = w.nf;
  Box empty;
  empty.lo = make_float4(1e30f, 1e30f, 1e30f, 0.f);
  empty.hi = make_float4(-1e30f, -1e30f, -1e30f, 0.f);
  if (w.status[ST_ERR] != 0) {
    w.src_n[i] = 0;
    w.src_boxes[i] = empty;
    w.scores[i] = make_double4(0, 0, 0, 0);
    w.facing[i] = 0;
    w.clipped[i] = 0;
    return;
  }
  const double *m = w.camera;
  const double *s = w.settings;
  bool persp = s[4] != 0.0;
  double c[3][3];
  bool bad = false;
#pragma unroll
  for (int j = 0; j < 3; ++j) {
    int vi = w.faces[3 * i + j];
    double x = w.vertices[3 * vi], y = w.vertices[3 * vi + 1], z = w.vertices[3 * vi + 2];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      c[j][k] = m[4 * k] * x + m[4 * k + 1] * y + m[4 * k + 2] * z + m[4 * k + 3] * 1.0;
      bad |= !isfinite(c[j][k]);
    }
  }
  if (bad) set_error(w.status, ERR_INPUT);
  if ((!persp || (c[0][2] - s[5] >= 0.0 && c[1][2] - s[5] >= 0.0 && c[2][2] - s[5] >= 0.0)) &&
      project_unclipped(w, i, c, persp))
    return;
  // a face that the near plane or the window clips: rings of up to SRCV
  // vertices, indexed at run time
  double r[SRCV][3];
  int count = 3;
  for (int j = 0; j < 3; ++j)
    for (int k = 0; k < 3; ++k) r[j][k] = c[j][k];
  int was_clipped = 0;
  if (persp) {
    // perspective near clip (warp_peel/kernels.py:392-415)
    double near_ring[SRCV][3];
    count = 0;
    for (int j = 0; j < 3; ++j) {
      const double *a = r[j], *b = r[(j + 1) % 3];
      double da = a[2] - s[5], db = b[2] - s[5];
      if (da >= 0.0) {
        for (int k = 0; k < 3; ++k) near_ring[count][k] = a[k];
        count += 1;
      } else {
        was_clipped = 1;
      }
      if ((da > 0.0 && db < 0.0) || (da < 0.0 && db > 0.0)) {
        double t = da / (da - db);
        for (int k = 0; k < 3; ++k) near_ring[count][k] = a[k] + t * (b[k] - a[k]);
        count += 1;
      }
    }
    for (int j = 0; j < count; ++j)
      for (int k = 0; k < 3; ++k) r[j][k] = near_ring[j][k];
  }
  Ring8 p;
  p.n = 0;
  double depth[SRCV];
  for (int j = 0; j < count; ++j) {
    double invz = 1.0;
    if (persp) invz = 1.0 / r[j][2];
    double2 x = make_double2((r[j][0] * invz - s[0]) / (s[2] - s[0]),
                             (r[j][1] * invz - s[1]) / (s[3] - s[1]));
    ring_append(p, x);
    depth[j] = persp ? -invz : r[j][2];
    if (x.x < 0.0 || x.x > 1.0 || x.y < 0.0 || x.y > 1.0) was_clipped = 1;
  }
  double4 score = make_double4(0, 0, 0, 0);
  double det = 0.0;
  int index = 2;
  if (count >= 3) {
    for (int j = 2; j < count; ++j) {
      det = cross3(p.v[0], p.v[1], p.v[j]);
      if (det != 0.0) {
        index = j;
        break;
      }
    }
  }
  int face = 0;
  if (det != 0.0) {
    double2 pa = make_double2(p.v[1].x - p.v[0].x, p.v[1].y - p.v[0].y);
    double2 pb = make_double2(p.v[index].x - p.v[0].x, p.v[index].y - p.v[0].y);
    double s0 = ((depth[1] - depth[0]) * pb.y - (depth[index] - depth[0]) * pa.y) / det;
    double s1 = (pa.x * (depth[index] - depth[0]) - pb.x * (depth[1] - depth[0])) / det;
    score = make_double4(s0, s1, depth[0] - s0 * p.v[0].x - s1 * p.v[0].y, 0.0);
    face = 1;
    if (det < 0.0) {
      face = -1;
      Ring8 q;
      q.n = 0;
      for (int j = 0; j < p.n; ++j) ring_append(q, p.v[p.n - 1 - j]);
      p = q;
    }
    ring_keep(p, 1.0, 0.0, 0.0);
    ring_keep(p, -1.0, 0.0, 1.0);
    ring_keep(p, 0.0, 1.0, 0.0);
    ring_keep(p, 0.0, -1.0, 1.0);
  } else {
    p.n = 0;
  }
  if (!isfinite(score.x) || !isfinite(score.y) || !isfinite(score.z) || fabs(score.x) > 1e28 ||
      fabs(score.y) > 1e28 || fabs(score.z) > 1e28)
    set_error(w.status, ERR_INPUT);
  if (p.n > SRCV) {
    set_error(w.status, ERR_POLY);
    p.n = 0;
  }
  w.facing[i] = face;
  w.clipped[i] = was_clipped;
  w.scores[i] = score;
  w.src_n[i] = p.n;
  double lx = 1e30, ly = 1e30, lz = 1e30, hx = -1e30, hy = -1e30, hz = -1e30;
  for (int j = 0; j < p.n; ++j) {
    double2 x = p.v[j];
    w.src_xy[j * nf + i] = x;
    double z = value(score, x);
    lx = fmin(lx, x.x), ly = fmin(ly, x.y), lz = fmin(lz, z);
    hx = fmax(hx, x.x), hy = fmax(hy, x.y), hz = fmax(hz, z);
  }
  Box b;
  pad_box(lx, ly, lz, hx, hy, hz, b.lo, b.hi);
  w.src_boxes[i] = b;
}

// ---------------------------------------------------------------------------
// Undirected edge incidences (warp_peel/kernels.py:471-520): an edge is
// internal iff exactly two incidences with different start vertices and equal
// facing; everything else is a contour. Each incidence adds its direction and
// facing to its edge's slot, so deciding an incidence reads its face, its two
// canonical vertices and one slot (key and summary share a 16-byte record)
// instead of the other incidence's face, vertices and facing.
__device__ __forceinline__ unsigned long long edge_key(const Ws &w, int e, bool &from_min) {
  int f = e / 3, j = e % 3;
  int a = w.canonical[w.faces[3 * f + j]];
  int b = w.canonical[w.faces[3 * f + (j + 1) % 3]];
  from_min = a <= b;
  return ((unsigned long long)(uint32_t)min(a, b) << 32) | (uint32_t)max(a, b);
}

__global__ void k_edge_insert(Ws w) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 3 * w.nf || w.status[ST_ERR] != 0) return;
  bool from_min;
  unsigned long long key = edge_key(w, e, from_min);
  int face = w.facing[e / 3];
  unsigned long long add = 1ull | ((unsigned long long)from_min << EDGE_DIR_SHIFT) |
                           ((unsigned long long)(face > 0) << EDGE_FRONT_SHIFT) |
                           ((unsigned long long)(face < 0) << EDGE_BACK_SHIFT);
  int mask = w.ehash_size - 1;
  int slot = (int)(mix64(key) & (uint64_t)mask);
  for (int probe = 0; probe < w.ehash_size; ++probe) {
    unsigned long long cur = atomicCAS(&w.etable[slot].key, ~0ull, key);
    if (cur == ~0ull || cur == key) {
      atomicAdd(&w.etable[slot].info, add);
      return;
    }
    slot = (slot + 1) & mask;
  }
}

__global__ void k_edge_mark(Ws w) {
  int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= 3 * w.nf) return;
  if (w.status[ST_ERR] != 0) {
    w.masks[e] = 0;
    return;
  }
  bool from_min;
  unsigned long long key = edge_key(w, e, from_min);
  int mask = w.ehash_size - 1;
  int slot = (int)(mix64(key) & (uint64_t)mask);
  int contour = 1;
  for (int probe = 0; probe < w.ehash_size; ++probe) {
    EdgeSlot cur = w.etable[slot];
    if (cur.key == key) {
      unsigned long long info = cur.info;
      if ((uint32_t)info == 2u) {
        // two incidences: opposite directions (one starts at the smaller
        // vertex), and both front-, both back- or both zero-facing
        int dir = (int)((info >> EDGE_DIR_SHIFT) & 0xffff);
        int front = (int)((info >> EDGE_FRONT_SHIFT) & 0xff), back = (int)(info >> EDGE_BACK_SHIFT);
        bool same_facing = front == 2 || back == 2 || (front == 0 && back == 0);
        if (dir == 1 && same_facing) contour = 0;
      }
      break;
    }
    if (cur.key == ~0ull) break;
    slot = (slot + 1) & mask;
  }
  w.masks[e] = contour;
  if (!contour) w.control[C_INTERNAL] = 1;
}

__device__ __forceinline__ bool contour_edge(const Ws &w, int source, int edge) {
  // warp_peel/kernels.py:523-536
  if (w.clipped[source] != 0) return true;
  int j = edge;
  if (w.facing[source] < 0) j = (4 - edge) % 3;
  return w.masks[3 * source + j] != 0;
}

// ---------------------------------------------------------------------------
// Contour segments (warp_peel/event_kernels.py:388-501): one slot per source
// edge, canonical endpoint order, exact duplicate grouping with the minimum
// depth at each endpoint.
__device__ __forceinline__ uint64_t segment_hash(double2 a, double2 b) {
  return mix64(dbits(a.x) ^ mix64(dbits(a.y) + 0x9e3779b97f4a7c15ull) ^
               mix64(dbits(b.x) + 0x3c6ef372fe94f82aull) ^ mix64(dbits(b.y) + 0xdaa66d2c7ddf743full));
}

__global__ void k_segments(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.nseg) return;
  int source = i / 8, edge = i % 8;
  double2 a = make_double2(1e30, 1e30), b = a;
  bool active = false;
  if (source < w.nf && w.status[ST_ERR] == 0) {
    int n = w.src_n[source];
    if (edge < n && contour_edge(w, source, edge)) {
      a = w.src_xy[edge * w.nf + source];
      b = w.src_xy[((edge + 1) % n) * w.nf + source];
      if (lexless(b, a)) {
        double2 t = a;
        a = b;
        b = t;
      }
      active = true;
    }
  }
  w.seg_a[i] = a;
  w.seg_b[i] = b;
  w.seg_rep[i] = -1;
}

// Exact duplicate grouping: the minimum slot of equal segments owns the hash
// slot. The endpoints compared were written by k_segments (a previous
// kernel), so the probe needs no fences.
// Copies of one segment in a wave (the same starting hash slot and the same
// endpoints, e.g. the three edges shared by 65,536 identical footprints)
// act through their lowest lane, which also holds the smallest index: up to
// SEGMENT_ROUNDS such groups per wave, each found by one ballot. `leader` is
// the lane that acts for this lane (itself when ungrouped).
constexpr int SEGMENT_ROUNDS = 4;

__device__ __forceinline__ uint64_t segment_group(bool valid, int slot, double2 a, double2 b, int &leader,
                                                  int round_leader[SEGMENT_ROUNDS]) {
  const int lane = threadIdx.x & 63;
  leader = lane;
  uint64_t todo = __ballot(valid);
  uint64_t grouped = 0;
  for (int r = 0; r < SEGMENT_ROUNDS; ++r) {
    round_leader[r] = -1;
    if (!todo) continue;
    int l = __ffsll((unsigned long long)todo) - 1;
    int ls = __shfl(slot, l, 64);
    double2 la = make_double2(__shfl(a.x, l, 64), __shfl(a.y, l, 64));
    double2 lb = make_double2(__shfl(b.x, l, 64), __shfl(b.y, l, 64));
    bool peer = valid && ((grouped >> lane) & 1) == 0 && slot == ls && same(a, la) && same(b, lb);
    uint64_t peers = __ballot(peer);
    if (peer) leader = l;
    round_leader[r] = l;
    grouped |= peers;
    todo &= ~peers;
  }
  return grouped;
}

__global__ void k_segments_insert(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool valid = i < w.nseg;
  double2 a = valid ? w.seg_a[i] : make_double2(1e30, 0), b = valid ? w.seg_b[i] : make_double2(0, 0);
  valid = valid && a.x < 1e29;
  int mask = w.shash_size - 1;
  int slot = valid ? (int)(segment_hash(a, b) & (uint64_t)mask) : -1;
  int leader, rounds[SEGMENT_ROUNDS];
  segment_group(valid, slot, a, b, leader, rounds);
  if (!valid || leader != (int)(threadIdx.x & 63)) return;
  for (int probe = 0; probe < w.shash_size; ++probe) {
    int cur = load_i(&w.shash[slot]);
    if (cur < 0) {
      cur = atomicCAS(&w.shash[slot], -1, i);
      if (cur < 0) return;
    }
    double2 ca = w.seg_a[cur], cb = w.seg_b[cur];
    if (same(ca, a) && same(cb, b)) {
      if (i < cur) atomicMin(&w.shash[slot], i);
      return;
    }
    slot = (slot + 1) & mask;
  }
}

__global__ void k_segments_reduce(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  bool valid = i < w.nseg;
  double2 a = valid ? w.seg_a[i] : make_double2(1e30, 0), b = valid ? w.seg_b[i] : make_double2(0, 0);
  valid = valid && a.x < 1e29;
  int mask = w.shash_size - 1;
  int slot = valid ? (int)(segment_hash(a, b) & (uint64_t)mask) : -1;
  int leader, rounds[SEGMENT_ROUNDS];
  segment_group(valid, slot, a, b, leader, rounds);
  int rep = i;
  unsigned long long ka = ~0ull, kb = ~0ull;
  if (valid) {
    int probe_slot = slot;
    for (int probe = 0; probe < w.shash_size; ++probe) {
      int cur = w.shash[probe_slot];
      if (cur < 0) break;
      if (same(w.seg_a[cur], a) && same(w.seg_b[cur], b)) {
        rep = cur;
        break;
      }
      probe_slot = (probe_slot + 1) & mask;
    }
    w.seg_rep[i] = rep;
    double4 sc = w.scores[i / 8];
    ka = (unsigned long long)dkey(value(sc, a));
    kb = (unsigned long long)dkey(value(sc, b));
  }
  // each group's nearest endpoint depths, gathered on its leader
  const int lane = threadIdx.x & 63;
  for (int r = 0; r < SEGMENT_ROUNDS; ++r) {
    if (rounds[r] < 0) continue;  // wave-uniform
    bool peer = valid && leader == rounds[r];
    unsigned long long ma = peer ? ka : ~0ull, mb = peer ? kb : ~0ull;
    for (int off = 32; off > 0; off >>= 1) {
      ma = min(ma, (unsigned long long)__shfl_xor(ma, off, 64));
      mb = min(mb, (unsigned long long)__shfl_xor(mb, off, 64));
    }
    if (lane == rounds[r]) ka = ma, kb = mb;
  }
  if (valid && leader == lane) {
    if (__hip_atomic_load(&w.seg_near[rep], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > ka)
      atomicMin(&w.seg_near[rep], ka);
    if (__hip_atomic_load(&w.seg_near[w.nseg + rep], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) > kb)
      atomicMin(&w.seg_near[w.nseg + rep], kb);
  }
  metric_add(w, M_UNIQUE, valid && rep == i ? 1 : 0);
}

__global__ void k_segment_flags(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.nseg) return;
  w.seg_flag[i] = w.seg_rep[i] == i;
}

// Unique segments, compacted in slot order, with their padded binary32 boxes
// (edge_bounds, event_kernels.py:483-501).
__global__ void k_segment_compact(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.nseg || w.seg_rep[i] != i) return;
  int k = w.seg_off[i];
  double2 a = w.seg_a[i], c = w.seg_b[i];
  double za = unkey(w.seg_near[i]), zc = unkey(w.seg_near[w.nseg + i]);
  w.cseg_a[k] = a;
  w.cseg_b[k] = c;
  w.cnear[k] = za;
  w.cnear[w.nseg + k] = zc;
  Box b;
  pad_box(fmin(a.x, c.x), fmin(a.y, c.y), fmin(za, zc), fmax(a.x, c.x), fmax(a.y, c.y), fmax(za, zc),
          b.lo, b.hi);
  w.seg_boxes[k] = b;
}

static inline dim3 grid_for(long n, int block = 256) {
  long g = (n + block - 1) / block;
  return dim3((unsigned)(g < 1 ? 1 : g));
}

void enqueue_prepare(const Ws &w, hipStream_t s, Timeline *t) {
  long reset_n = w.vhash_size;
  if (w.ehash_size > reset_n) reset_n = w.ehash_size;
  if (w.shash_size > reset_n) reset_n = w.shash_size;
  if (2L * w.nseg > reset_n) reset_n = 2L * w.nseg;
  long reset_grid = (reset_n + 255) / 256;
  if (reset_grid > 2048) reset_grid = 2048;
  hipLaunchKernelGGL(k_reset, dim3((unsigned)reset_grid), dim3(256), 0, s, w);
  long nvf = w.nv > w.nf ? w.nv : w.nf;
  hipLaunchKernelGGL(k_validate, grid_for(nvf < 1 ? 1 : nvf), dim3(256), 0, s, w);
  if (w.nv > 0) {
    hipLaunchKernelGGL(k_weld_insert, grid_for(w.nv), dim3(256), 0, s, w);
    hipLaunchKernelGGL(k_weld_resolve, grid_for(w.nv), dim3(256), 0, s, w);
  }
  if (w.nf > 0) {
    hipLaunchKernelGGL(k_project, grid_for(w.nf), dim3(256), 0, s, w);
    hipLaunchKernelGGL(k_edge_insert, grid_for(3L * w.nf), dim3(256), 0, s, w);
    hipLaunchKernelGGL(k_edge_mark, grid_for(3L * w.nf), dim3(256), 0, s, w);
  }
  mark(t, s, "weld_project_topology");
  if (w.route == ROUTE_LAYERED) {
    Hierarchy st = w.source_tree;
    build_hierarchy(s, w.src_boxes, nullptr, w.nf, w.nf, false, st, w.sort);
    mark(t, s, "source_hierarchy");
    return;
  }
  hipLaunchKernelGGL(k_segments, grid_for(w.nseg), dim3(256), 0, s, w);
  hipLaunchKernelGGL(k_segments_insert, grid_for(w.nseg), dim3(256), 0, s, w);
  hipLaunchKernelGGL(k_segments_reduce, grid_for(w.nseg), dim3(256), 0, s, w);
  hipLaunchKernelGGL(k_segment_flags, grid_for(w.nseg), dim3(256), 0, s, w);
  scan_exclusive(s, w.seg_flag, w.seg_off, nullptr, w.nseg, w.nseg, w.control + C_NSEG, w.scan);
  hipLaunchKernelGGL(k_segment_compact, grid_for(w.nseg), dim3(256), 0, s, w);
  Hierarchy st = w.source_tree, ct = w.contour_tree;
  mark(t, s, "contour_segments");
  build_hierarchy(s, w.src_boxes, nullptr, w.nf, w.nf, false, st, w.sort);
  mark(t, s, "source_hierarchy");
  LeafGather records;
  records.a = w.cseg_a;
  records.b = w.cseg_b;
  records.near = w.cnear;
  records.near_stride = w.nseg;
  records.out_a = w.cl_a;
  records.out_b = w.cl_b;
  records.out_near = w.cl_near;
  build_hierarchy(s, w.seg_boxes, w.control + C_NSEG, 0, w.nseg, false, ct, w.sort, records);
  mark(t, s, "contour_hierarchy");
}

}  // namespace peel
