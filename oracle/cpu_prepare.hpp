// ORACLE — test infrastructure only (see expansion.hpp). Shared by the CPU
// restatements of the device routes (cpu_event.cpp, cpu_layered.cpp): ring
// helpers, the Morton hierarchy and the preparation stages, in the reference's
// binary64 operation order (compiled with -ffp-contract=off):
//  validate / weld / project / edge topology  warp_peel/kernels.py:269-536
//  Morton hierarchy + stackless traversal     warp_peel/solver.py:34-89, kernels.py:24-116
//  clean, split, front_segment                warp_peel/kernels.py:146-232, frontier.py:45-103
#pragma once
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

constexpr int MAXV = 32;
constexpr double EPS = 1.0e-10;

struct V2 {
  double x, y;
};
struct Box {
  float lo[3], hi[3];
};

inline double cross(V2 a, V2 b, V2 c) {
  double ux = b.x - a.x, uy = b.y - a.y, vx = c.x - a.x, vy = c.y - a.y;
  return ux * vy - uy * vx;
}
inline double value(const double *s, V2 x) { return s[0] * x.x + s[1] * x.y + s[2]; }
inline bool lexless(V2 a, V2 b) { return a.x < b.x || (a.x == b.x && a.y < b.y); }
inline bool same(V2 a, V2 b) { return a.x == b.x && a.y == b.y; }
inline double len(V2 d) { return std::sqrt(d.x * d.x + d.y * d.y); }

Box empty_box() {
  Box b;
  for (int k = 0; k < 3; ++k) b.lo[k] = 1e30f, b.hi[k] = -1e30f;
  return b;
}
Box pad_box(double lx, double ly, double lz, double hx, double hy, double hz) {
  double pz = 2.0e-7 * std::fmax(1.0, std::fmax(std::fabs(lz), std::fabs(hz)));
  Box b;
  b.lo[0] = (float)(lx - 2.0e-7), b.lo[1] = (float)(ly - 2.0e-7), b.lo[2] = (float)(lz - pz);
  b.hi[0] = (float)(hx + 2.0e-7), b.hi[1] = (float)(hy + 2.0e-7), b.hi[2] = (float)(hz + pz);
  return b;
}

// ---------------------------------------------------------------------------
// Morton-ordered implicit complete tree (solver.py:34-89, kernels.py:67-116).
struct Tree {
  std::vector<Box> nodes;  // [2 leaves]
  std::vector<int> ids;    // [leaves]
  int leaves = 1;
  void build(const std::vector<Box> &boxes, bool layered) {
    int n = (int)boxes.size();
    leaves = 1;
    while (leaves < n) leaves <<= 1;
    std::vector<uint64_t> keys(n);
    ids.assign(leaves, 0);
    for (int i = 0; i < n; ++i) {
      const Box &b = boxes[i];
      uint64_t key = ~0ull;
      if (b.lo[0] <= b.hi[0] && b.lo[1] <= b.hi[1]) {
        float cx = (b.lo[0] + b.hi[0]) * 0.5f, cy = (b.lo[1] + b.hi[1]) * 0.5f;
        cx = std::fmin(std::fmax(cx, 0.0f), 1.0f);
        cy = std::fmin(std::fmax(cy, 0.0f), 1.0f);
        uint32_t x = (uint32_t)(int)(cx * 32767.0f), y = (uint32_t)(int)(cy * 32767.0f);
        key = 0;
        for (int bit = 0; bit < 15; ++bit) {
          key |= (uint64_t)((x >> bit) & 1u) << (2 * bit);
          key |= (uint64_t)((y >> bit) & 1u) << (2 * bit + 1);
        }
        if (layered) key |= (uint64_t)(uint32_t)(int)b.lo[2] << 32;
      }
      keys[i] = layered ? key : (key & 0xffffffffull);
    }
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return keys[a] < keys[b]; });
    for (int i = 0; i < n; ++i) ids[i] = order[i];
    nodes.assign(2 * leaves, empty_box());
    for (int i = 0; i < n; ++i) nodes[leaves + i] = boxes[ids[i]];
    for (int i = leaves - 1; i >= 1; --i) {
      const Box &a = nodes[2 * i], &b = nodes[2 * i + 1];
      for (int k = 0; k < 3; ++k) {
        nodes[i].lo[k] = std::fmin(a.lo[k], b.lo[k]);
        nodes[i].hi[k] = std::fmax(a.hi[k], b.hi[k]);
      }
    }
  }
  static int successor(int node) {
    while (node > 0 && (node & 1) != 0) node >>= 1;
    if (node > 0) node += 1;
    return node;
  }
  // next overlapping leaf position in depth-first order; advances `node`
  int next_leaf(const float *lo, const float *hi, int &node, int end = 0) const {
    while (node > 0 && node != end) {
      const Box &b = nodes[node];
      bool overlap = b.lo[0] <= hi[0] && b.lo[1] <= hi[1] && b.lo[2] <= hi[2] && b.hi[0] >= lo[0] &&
                     b.hi[1] >= lo[1] && b.hi[2] >= lo[2];
      if (overlap) {
        if (node < leaves) {
          node = 2 * node;
        } else {
          int leaf = node - leaves;
          node = successor(node);
          return leaf;
        }
      } else {
        node = successor(node);
      }
    }
    return -1;
  }
};

// ---------------------------------------------------------------------------
// Ring helpers (kernels.py:146-232, frontier.py:45-78).
struct Ring {
  V2 v[MAXV];
  int n = 0;
  void append(V2 x) {
    if (n < MAXV) v[n] = x;
    n += 1;
  }
};

void clean(Ring &p) {
  if (p.n > MAXV) return;
  Ring q;
  for (int j = 0; j < p.n; ++j) {
    V2 x = p.v[j];
    if (q.n == 0 || len({x.x - q.v[q.n - 1].x, x.y - q.v[q.n - 1].y}) > EPS * 0.01) q.append(x);
  }
  if (q.n > 1 && len({q.v[0].x - q.v[q.n - 1].x, q.v[0].y - q.v[q.n - 1].y}) <= EPS * 0.01) q.n -= 1;
  if (q.n < 3) {
    q.n = 0;
  } else {
    V2 origin = q.v[0];
    double area2 = 0.0, perimeter = 0.0;
    for (int j = 0; j < q.n; ++j) {
      V2 a = q.v[j], b = q.v[(j + 1) % q.n];
      area2 += cross(origin, a, b);
      perimeter += std::fabs(b.x - a.x) + std::fabs(b.y - a.y);
    }
    if (area2 <= EPS * perimeter) q.n = 0;
  }
  p = q;
}

void keep_inside(Ring &p, double sx, double sy, double sc) {
  Ring in;
  for (int j = 0; j < p.n; ++j) {
    V2 a = p.v[j], b = p.v[(j + 1) % p.n];
    double da = sx * (a.x - 0.0) + sy * (a.y - 0.0) + sc;
    double db = sx * (b.x - 0.0) + sy * (b.y - 0.0) + sc;
    if (da >= 0.0) in.append(a);
    if ((da > 0.0 && db < 0.0) || (da < 0.0 && db > 0.0)) {
      double t = da / (da - db);
      in.append({a.x + t * (b.x - a.x), a.y + t * (b.y - a.y)});
    }
  }
  clean(in);
  p = in;
}

bool front_segment(const Ring &p, V2 a, V2 b, double za, double zb) {
  double low = 0.0, high = 1.0;
  for (int j = 0; j <= p.n; ++j) {
    double da = za, db = zb;
    if (j < p.n) {
      V2 u = p.v[j], v = p.v[(j + 1) % p.n];
      da = cross(u, v, a);
      db = cross(u, v, b);
    }
    if (da <= 0.0 && db <= 0.0) return false;
    if (da < 0.0) low = std::fmax(low, da / (da - db));
    if (db < 0.0) high = std::fmin(high, da / (da - db));
  }
  return high > low;
}


// Projected sources and their topology (kernels.py:269-536).
struct Prep {
  std::vector<int> canonical;
  std::vector<Ring> src;
  std::vector<double> score;  // [4 nf] depth plane s0, s1, s2 (and padding)
  std::vector<int> facing, clipped, masks;
  std::vector<Box> src_box;
  int error = 0;
  bool persp = false;
  // contour_edge (kernels.py:523-536): ring edge `edge` of source s
  bool contour_edge(int s, int edge) const {
    if (clipped[s]) return true;
    int j = facing[s] < 0 ? (4 - edge) % 3 : edge;
    return masks[3 * s + j] != 0;
  }
};

// Returns the error bits (0 on success); P holds every stage's output.
int prepare(const double *V, int nv, const int *F, int nf, const double *M, const double *S, Prep &P) {
  int error = 0;
  // ---- validate (kernels.py:309-350) ------------------------------------
  for (int i = 0; i < 3 * nv; ++i)
    if (!std::isfinite(V[i])) error |= 256;
  for (int i = 0; i < 3 * nf; ++i)
    if (F[i] < 0 || F[i] >= nv) error |= 256;
  if (error) return P.error = error;
  bool persp = S[4] != 0.0;
  P.persp = persp;
  // ---- weld: canonical = smallest index with equal coordinates ----------
  std::vector<int> &canonical = P.canonical;
  canonical.assign(nv, 0);
  {
    std::vector<int> order(nv);
    for (int i = 0; i < nv; ++i) order[i] = i;
    auto less = [&](int a, int b) {
      const double *p = V + 3 * a, *q = V + 3 * b;
      if (p[0] != q[0]) return p[0] < q[0];
      if (p[1] != q[1]) return p[1] < q[1];
      return p[2] < q[2];
    };
    std::stable_sort(order.begin(), order.end(), less);
    int first = 0;
    for (int k = 0; k < nv; ++k) {
      if (k > 0 && less(order[k - 1], order[k])) first = k;
      canonical[order[k]] = order[first];
    }
  }
  // ---- project (kernels.py:353-468) -------------------------------------
  std::vector<Ring> &src = P.src;
  std::vector<double> &score = P.score;
  std::vector<int> &facing = P.facing, &clipped = P.clipped;
  std::vector<Box> &src_box = P.src_box;
  src.assign(nf, Ring());
  score.assign(4 * std::max(nf, 1), 0.0);
  facing.assign(nf, 0), clipped.assign(nf, 0);
  src_box.assign(nf, Box());
#pragma omp parallel for schedule(static) reduction(| : error)
  for (int i = 0; i < nf; ++i) {
    double r[8][3];
    int count = 3;
    for (int j = 0; j < 3; ++j) {
      const double *w = V + 3 * F[3 * i + j];
      for (int k = 0; k < 3; ++k) {
        r[j][k] = M[4 * k] * w[0] + M[4 * k + 1] * w[1] + M[4 * k + 2] * w[2] + M[4 * k + 3] * 1.0;
        if (!std::isfinite(r[j][k])) error |= 256;
      }
    }
    int was_clipped = 0;
    if (persp) {
      double nr[8][3];
      count = 0;
      for (int j = 0; j < 3; ++j) {
        const double *a = r[j], *b = r[(j + 1) % 3];
        double da = a[2] - S[5], db = b[2] - S[5];
        if (da >= 0.0) {
          for (int k = 0; k < 3; ++k) nr[count][k] = a[k];
          count += 1;
        } else {
          was_clipped = 1;
        }
        if ((da > 0.0 && db < 0.0) || (da < 0.0 && db > 0.0)) {
          double t = da / (da - db);
          for (int k = 0; k < 3; ++k) nr[count][k] = a[k] + t * (b[k] - a[k]);
          count += 1;
        }
      }
      std::memcpy(r, nr, sizeof(r));
    }
    Ring p;
    double depth[8];
    for (int j = 0; j < count; ++j) {
      double invz = 1.0;
      if (persp) invz = 1.0 / r[j][2];
      V2 x{(r[j][0] * invz - S[0]) / (S[2] - S[0]), (r[j][1] * invz - S[1]) / (S[3] - S[1])};
      p.append(x);
      depth[j] = persp ? -invz : r[j][2];
      if (x.x < 0.0 || x.x > 1.0 || x.y < 0.0 || x.y > 1.0) was_clipped = 1;
    }
    double det = 0.0;
    int index = 2;
    if (count >= 3)
      for (int j = 2; j < count; ++j) {
        det = cross(p.v[0], p.v[1], p.v[j]);
        if (det != 0.0) {
          index = j;
          break;
        }
      }
    double *sc = &score[4 * i];
    if (det != 0.0) {
      V2 pa{p.v[1].x - p.v[0].x, p.v[1].y - p.v[0].y}, pb{p.v[index].x - p.v[0].x, p.v[index].y - p.v[0].y};
      double s0 = ((depth[1] - depth[0]) * pb.y - (depth[index] - depth[0]) * pa.y) / det;
      double s1 = (pa.x * (depth[index] - depth[0]) - pb.x * (depth[1] - depth[0])) / det;
      sc[0] = s0, sc[1] = s1, sc[2] = depth[0] - s0 * p.v[0].x - s1 * p.v[0].y;
      facing[i] = 1;
      if (det < 0.0) {
        facing[i] = -1;
        Ring q;
        for (int j = 0; j < p.n; ++j) q.append(p.v[p.n - 1 - j]);
        p = q;
      }
      keep_inside(p, 1.0, 0.0, 0.0);
      keep_inside(p, -1.0, 0.0, 1.0);
      keep_inside(p, 0.0, 1.0, 0.0);
      keep_inside(p, 0.0, -1.0, 1.0);
    } else {
      p.n = 0;
    }
    for (int k = 0; k < 3; ++k)
      if (!std::isfinite(sc[k]) || std::fabs(sc[k]) > 1e28) error |= 256;
    clipped[i] = was_clipped;
    src[i] = p;
    double lx = 1e30, ly = 1e30, lz = 1e30, hx = -1e30, hy = -1e30, hz = -1e30;
    for (int j = 0; j < p.n; ++j) {
      double z = value(sc, p.v[j]);
      lx = std::fmin(lx, p.v[j].x), ly = std::fmin(ly, p.v[j].y), lz = std::fmin(lz, z);
      hx = std::fmax(hx, p.v[j].x), hy = std::fmax(hy, p.v[j].y), hz = std::fmax(hz, z);
    }
    src_box[i] = pad_box(lx, ly, lz, hx, hy, hz);
  }
  if (error) return P.error = error;
  // ---- edge topology (kernels.py:471-520) --------------------------------
  std::vector<int> &masks = P.masks;
  masks.assign(3 * nf, 1);
  {
    std::vector<std::pair<uint64_t, int>> keys(3 * nf);
    for (int e = 0; e < 3 * nf; ++e) {
      int f = e / 3, j = e % 3;
      int a = canonical[F[3 * f + j]], b = canonical[F[3 * f + (j + 1) % 3]];
      keys[e] = {((uint64_t)(uint32_t)std::min(a, b) << 32) | (uint32_t)std::max(a, b), e};
    }
    std::stable_sort(keys.begin(), keys.end(),
                     [](const std::pair<uint64_t, int> &a, const std::pair<uint64_t, int> &b) {
                       return a.first < b.first;
                     });
    for (size_t i = 0; i < keys.size();) {
      size_t end = i + 1;
      while (end < keys.size() && keys[end].first == keys[i].first) ++end;
      bool contour = true;
      if (end - i == 2) {
        int e = keys[i].second, f = keys[i + 1].second;
        int a = canonical[F[3 * (e / 3) + e % 3]], b = canonical[F[3 * (f / 3) + f % 3]];
        if (a != b && facing[e / 3] == facing[f / 3]) contour = false;
      }
      for (size_t k = i; k < end; ++k) masks[keys[k].second] = contour;
      i = end;
    }
  }
  return P.error = 0;
}

}  // namespace
