// ORACLE — test infrastructure only (see expansion.hpp). Built into
// oracle/liboracle.so; used by tests/ as an independent partition reference and
// by bench.py as the timed CPU baseline. Never linked by the product.
//
// CPU FP64 restatement of the reference's event route, sequential with OpenMP
// over independent items, following the reference file by file:
//  validate / weld / project / edge topology  warp_peel/kernels.py:269-536
//  contour segments                           warp_peel/event_kernels.py:388-501
//  Morton hierarchy + stackless traversal     warp_peel/solver.py:34-89, kernels.py:24-116
//  subdivision (seed, clean, front_segment)   warp_peel/frontier.py:15-223
//  samples, exact groups, rays, hit order     warp_peel/event_kernels.py:28-385, events.py:703-718
//  noding, centroid fans, lift                warp_peel/kernels.py:941-1124
// Every binary64 expression is written in the reference's operation order and
// the file is compiled with -ffp-contract=off, so each operation rounds once.
// Patches are processed per seed with a worklist (each patch's cuts depend only
// on its own ring and hierarchy cursor), which yields the same final patch set
// as the reference's round-synchronous frontier.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <functional>
#include <map>
#include <array>
#include <tuple>
#include <unordered_map>
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

struct Out {
  std::vector<Ring> polys;
  std::vector<int> src, layer;
};

bool g_propagate = true;
bool g_share = true;

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

}  // namespace

extern "C" {

void oracle_cpu_event_options(int propagate_ranks, int share_regions) {
  g_propagate = propagate_ranks != 0;
  g_share = share_regions != 0;
}

// counts[0..13]: polygons, triangles, layers, splits, ray_groups, ray_hits,
// error bits, contour_tests, unique_contours, boundary vertices, rank
// components, rank connections, beam attempts, certified beams.
// seconds[0..3]: prepare, subdivide, rank, mesh.
int oracle_cpu_event_peel(const double *V, int nv, const int *F, int nf, const double *M,
                          const double *S, int layers_limit, int threads, double *poly_xy,
                          int *poly_n, int *poly_src, int *poly_layer, int cap_poly, double *tri_uv,
                          double *tri_world, int *tri_src, int *tri_layer, int cap_tri, long *counts,
                          double *seconds) {
#ifdef _OPENMP
  if (threads > 0) omp_set_num_threads(threads);
#endif
  for (int k = 0; k < 14; ++k) counts[k] = 0;
  int error = 0;
  double t0 = now();
  // ---- validate (kernels.py:309-350) ------------------------------------
  for (int i = 0; i < 3 * nv; ++i)
    if (!std::isfinite(V[i])) error |= 256;
  for (int i = 0; i < 3 * nf; ++i)
    if (F[i] < 0 || F[i] >= nv) error |= 256;
  if (error) {
    counts[6] = error;
    return 0;
  }
  bool persp = S[4] != 0.0;
  // ---- weld: canonical = smallest index with equal coordinates ----------
  std::vector<int> canonical(nv);
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
  std::vector<Ring> src(nf);
  std::vector<double> score(4 * std::max(nf, 1), 0.0);
  std::vector<int> facing(nf, 0), clipped(nf, 0);
  std::vector<Box> src_box(nf);
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
  if (error) {
    counts[6] = error;
    return 0;
  }
  // ---- edge topology (kernels.py:471-520) --------------------------------
  std::vector<int> masks(3 * nf, 1);
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
  auto contour_edge = [&](int s, int edge) {
    if (clipped[s]) return true;
    int j = facing[s] < 0 ? (4 - edge) % 3 : edge;
    return masks[3 * s + j] != 0;
  };
  // ---- contour segments (event_kernels.py:388-501) -----------------------
  std::vector<V2> seg_a, seg_b;
  std::vector<double> near0, near1;
  {
    std::map<std::tuple<double, double, double, double>, int> unique;
    for (int i = 0; i < 8 * nf; ++i) {
      int s = i / 8, edge = i % 8;
      const Ring &p = src[s];
      if (edge >= p.n || !contour_edge(s, edge)) continue;
      V2 a = p.v[edge], b = p.v[(edge + 1) % p.n];
      if (lexless(b, a)) std::swap(a, b);
      double za = value(&score[4 * s], a), zb = value(&score[4 * s], b);
      auto key = std::make_tuple(a.x == 0.0 ? 0.0 : a.x, a.y == 0.0 ? 0.0 : a.y, b.x == 0.0 ? 0.0 : b.x,
                                 b.y == 0.0 ? 0.0 : b.y);
      auto it = unique.find(key);
      if (it == unique.end()) {
        unique.emplace(key, (int)seg_a.size());
        seg_a.push_back(a), seg_b.push_back(b), near0.push_back(za), near1.push_back(zb);
      } else {
        int k = it->second;
        near0[k] = std::fmin(near0[k], za);
        near1[k] = std::fmin(near1[k], zb);
      }
    }
  }
  counts[8] = (long)seg_a.size();
  std::vector<Box> seg_box(seg_a.size());
  for (size_t k = 0; k < seg_a.size(); ++k) {
    V2 a = seg_a[k], c = seg_b[k];
    seg_box[k] = pad_box(std::fmin(a.x, c.x), std::fmin(a.y, c.y), std::fmin(near0[k], near1[k]),
                         std::fmax(a.x, c.x), std::fmax(a.y, c.y), std::fmax(near0[k], near1[k]));
  }
  Tree source_tree, contour_tree;
  source_tree.build(src_box, false);
  contour_tree.build(seg_box, false);
  double t1 = now();
  // ---- subdivision (frontier.py:106-223) ----------------------------------
  std::vector<Out> per(nf);
  long splits = 0, tests = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : splits, tests)
  for (int seed = 0; seed < nf; ++seed) {
    if (src[seed].n == 0) continue;
    const double *sc = &score[4 * seed];
    std::vector<std::pair<Ring, int>> work{{src[seed], 1}};
    while (!work.empty()) {
      Ring p = work.back().first;
      int node = work.back().second;
      work.pop_back();
      double lx = 1e30, ly = 1e30, lz = 1e30, hx = -1e30, hy = -1e30, hz = -1e30;
      for (int j = 0; j < p.n; ++j) {
        double z = value(sc, p.v[j]);
        lx = std::fmin(lx, p.v[j].x), ly = std::fmin(ly, p.v[j].y), lz = std::fmin(lz, z);
        hx = std::fmax(hx, p.v[j].x), hy = std::fmax(hy, p.v[j].y), hz = std::fmax(hz, z);
      }
      Box q = pad_box(lx, ly, lz, hx, hy, hz);
      q.lo[2] = -1e30f;
      bool done = true;
      while (true) {
        int c = contour_tree.next_leaf(q.lo, q.hi, node);
        if (c < 0) break;
        int k = contour_tree.ids[c];
        tests += 1;
        V2 a = seg_a[k], b = seg_b[k];
        bool positive = false, negative = false;
        for (int j = 0; j < p.n; ++j) {
          double side = cross(a, b, p.v[j]);
          positive |= side > 0.0;
          negative |= side < 0.0;
        }
        if (!(positive && negative)) continue;
        double za = value(sc, a) - near0[k], zb = value(sc, b) - near1[k];
        if (!front_segment(p, a, b, za, zb)) continue;
        Ring left, right;
        V2 d{b.x - a.x, b.y - a.y};
        for (int j = 0; j < p.n; ++j) {
          V2 u = p.v[j], v = p.v[(j + 1) % p.n];
          double du = -d.y * (u.x - a.x) + d.x * (u.y - a.y) + 0.0;
          double dv = -d.y * (v.x - a.x) + d.x * (v.y - a.y) + 0.0;
          if (du >= 0.0) left.append(u);
          if (du <= 0.0) right.append(u);
          if ((du > 0.0 && dv < 0.0) || (du < 0.0 && dv > 0.0)) {
            double t = du / (du - dv);
            V2 x{u.x + t * (v.x - u.x), u.y + t * (v.y - u.y)};
            left.append(x);
            right.append(x);
          }
        }
        clean(left);
        clean(right);
        if (left.n > MAXV || right.n > MAXV) {
#pragma omp atomic
          error |= 1;
          done = false;
          break;
        }
        if (left.n > 0 && right.n > 0) {
          splits += 1;
          work.push_back({right, node});
          work.push_back({left, node});
          done = false;
          break;
        }
        // collapsed child: the parent continues from the cursor
      }
      if (done) {
        per[seed].polys.push_back(p);
        per[seed].src.push_back(seed);
      }
    }
  }
  Out out;
  for (int s = 0; s < nf; ++s)
    for (size_t k = 0; k < per[s].polys.size(); ++k) {
      out.polys.push_back(per[s].polys[k]);
      out.src.push_back(s);
    }
  int P = (int)out.polys.size();
  out.layer.assign(P, 0);
  counts[3] = splits;
  counts[7] = tests;
  double t2 = now();
  // ---- rank propagation (components.py:15-271) -----------------------------
  std::vector<int> reps(P);
  for (int i = 0; i < P; ++i) reps[i] = i;
  bool any_internal = false;
  for (int e = 0; e < 3 * nf; ++e) any_internal |= masks[e] == 0;
  long components = g_propagate ? P : 0, connections = 0;  // identity: components.py:261-271
  if (g_propagate && (splits > 0 || any_internal)) {
    struct PE {
      uint64_t owner;
      V2 a, b;
      int poly;
    };
    std::vector<PE> edges;
    for (int i = 0; i < P; ++i) {
      const Ring &p = out.polys[i];
      int s = out.src[i];
      const Ring &q = src[s];
      for (int k = 0; k < p.n; ++k) {
        V2 a = p.v[k], b = p.v[(k + 1) % p.n];
        uint64_t owner = 0x8000000000000000ull | (uint64_t)(uint32_t)s;
        if (!clipped[s]) {
          for (int j = 0; j < q.n; ++j) {
            int edge = facing[s] < 0 ? (4 - j) % 3 : j;
            if (masks[3 * s + edge] == 0) {
              V2 u = q.v[j], v = q.v[(j + 1) % q.n];
              if ((same(a, u) || same(a, v) || cross(u, v, a) == 0.0) &&
                  (same(b, u) || same(b, v) || cross(u, v, b) == 0.0)) {
                int ca = canonical[F[3 * s + edge]], cb = canonical[F[3 * s + (edge + 1) % 3]];
                owner = ((uint64_t)(uint32_t)std::min(ca, cb) << 32) | (uint32_t)std::max(ca, cb);
                break;
              }
            }
          }
        }
        edges.push_back({owner, a, b, i});
      }
    }
    // exact reversed twins on the same owner (components.py:123-159)
    auto key = [](uint64_t o, V2 a, V2 b) {
      return std::make_tuple(o, a.x == 0.0 ? 0.0 : a.x, a.y == 0.0 ? 0.0 : a.y, b.x == 0.0 ? 0.0 : b.x,
                             b.y == 0.0 ? 0.0 : b.y);
    };
    std::map<std::tuple<uint64_t, double, double, double, double>, int> open;
    std::vector<std::pair<int, int>> pairs;
    for (int e = 0; e < (int)edges.size(); ++e) {
      const PE &x = edges[e];
      if (same(x.a, x.b)) continue;
      auto it = open.find(key(x.owner, x.b, x.a));
      if (it != open.end() && edges[it->second].poly != x.poly) {
        pairs.push_back({it->second, e});
        open.erase(it);
      } else {
        open.emplace(key(x.owner, x.a, x.b), e);
      }
    }
    std::vector<int> parent(P);
    for (int i = 0; i < P; ++i) parent[i] = i;
    std::function<int(int)> find = [&](int i) {
      while (parent[i] != i) i = parent[i] = parent[parent[i]];
      return i;
    };
    for (auto &pr : pairs) {
      const PE &x = edges[pr.first];
      if (len({x.b.x - x.a.x, x.b.y - x.a.y}) <= 16.0 * EPS) continue;
      V2 m{(x.a.x + x.b.x) * 0.5, (x.a.y + x.b.y) * 0.5};
      double z = value(&score[4 * out.src[x.poly]], m);
      double pad = 2.0e-7 * std::fmax(1.0, std::fabs(z));
      float lo[3] = {(float)m.x - 2.0e-7f, (float)m.y - 2.0e-7f, -1.0e30f};
      float hi[3] = {(float)m.x + 2.0e-7f, (float)m.y + 2.0e-7f, (float)(z + pad)};
      int node = 1;
      bool safe = true;
      while (node > 0) {
        int leaf = contour_tree.next_leaf(lo, hi, node);
        if (leaf < 0) break;
        int c = contour_tree.ids[leaf];
        V2 u = seg_a[c], v = seg_b[c];
        V2 d{v.x - u.x, v.y - u.y};
        double length = len(d);
        if (length > 0.0) {
          double t = ((m.x - u.x) * d.x + (m.y - u.y) * d.y) / (d.x * d.x + d.y * d.y);
          if (std::fabs(cross(u, v, m)) <= EPS * length && t >= -EPS / length && t <= 1.0 + EPS / length) {
            double depth = near0[c] + t * (near1[c] - near0[c]);
            if (depth <= z + pad) {
              safe = false;
              break;
            }
          }
        }
      }
      if (!safe) continue;
      connections += 1;
      int ra = find(x.poly), rb = find(edges[pr.second].poly);
      if (ra != rb) parent[std::max(ra, rb)] = std::min(ra, rb);
    }
    components = 0;
    for (int i = 0; i < P; ++i) {
      reps[i] = find(i);
      components += reps[i] == i;
    }
  }
  counts[10] = components;
  counts[11] = connections;
  // ---- ranks (event_kernels.py:28-385) ------------------------------------
  std::vector<V2> samples(P);
  for (int i = 0; i < P; ++i) {
    const Ring &p = out.polys[reps[i]];
    int anchor = 0;
    for (int j = 1; j < p.n; ++j)
      if (lexless(p.v[j], p.v[anchor])) anchor = j;
    V2 x{0.0, 0.0};
    for (int j = 0; j < p.n; ++j) {
      V2 a = p.v[(anchor + j) % p.n];
      x.x += a.x / (double)p.n;
      x.y += a.y / (double)p.n;
    }
    samples[i] = x;
  }
  std::vector<int> group(P);
  {
    std::map<std::pair<double, double>, int> first;
    for (int i = 0; i < P; ++i) {
      auto key = std::make_pair(samples[i].x == 0.0 ? 0.0 : samples[i].x, samples[i].y == 0.0 ? 0.0 : samples[i].y);
      auto it = first.emplace(key, i).first;
      group[i] = it->second;
    }
  }
  // shared rays over certified cell rectangles (beams.py:18-207): a cell of
  // the 32 x 32 window grid whose samples are not all identical shares one
  // ray when every source whose padded box meets the samples' bounding
  // rectangle strictly contains its four corners
  if (g_share) {
    const int side = 32, cells = side * side;
    std::vector<double> bx0(cells, 1e300), by0(cells, 1e300), bx1(cells, -1e300), by1(cells, -1e300);
    std::vector<int> leader(cells, -1), cell_of(P);
    for (int i = 0; i < P; ++i) {
      V2 x = samples[i];
      int cx = (int)std::min(std::max(x.x * side, 0.0), (double)(side - 1));
      int cy = (int)std::min(std::max(x.y * side, 0.0), (double)(side - 1));
      int c = cx * side + cy;
      cell_of[i] = c;
      bx0[c] = std::min(bx0[c], x.x);
      by0[c] = std::min(by0[c], x.y);
      bx1[c] = std::max(bx1[c], x.x);
      by1[c] = std::max(by1[c], x.y);
      if (leader[c] < 0) leader[c] = i;
    }
    std::vector<char> valid(cells, 0);
    for (int c = 0; c < cells; ++c) {
      if (leader[c] < 0 || (bx0[c] == bx1[c] && by0[c] == by1[c])) continue;
      counts[12] += 1;
      V2 corner[4] = {{bx0[c], by0[c]}, {bx1[c], by0[c]}, {bx0[c], by1[c]}, {bx1[c], by1[c]}};
      float lo[3] = {(float)bx0[c] - 2.0e-7f, (float)by0[c] - 2.0e-7f, -1.0e30f};
      float hi[3] = {(float)bx1[c] + 2.0e-7f, (float)by1[c] + 2.0e-7f, 1.0e30f};
      bool ok = true;
      int node = 1;
      while (ok && node > 0) {
        int leaf = source_tree.next_leaf(lo, hi, node);
        if (leaf < 0) break;
        const Ring &r = src[source_tree.ids[leaf]];
        ok = r.n > 0;
        for (int j = 0; j < r.n && ok; ++j) {
          V2 a = r.v[j], b = r.v[(j + 1) % r.n];
          double margin = 1.0e-10 * std::sqrt((b.x - a.x) * (b.x - a.x) + (b.y - a.y) * (b.y - a.y));
          for (int k = 0; k < 4 && ok; ++k)
            if (cross(a, b, corner[k]) <= margin) ok = false;
        }
      }
      valid[c] = ok;
      counts[13] += ok;
    }
    for (int i = 0; i < P; ++i)
      if (valid[cell_of[i]]) group[i] = leader[cell_of[i]];
  }
  // flat source records in leaf order
  struct Hit {
    double depth;
    int source;
  };
  std::vector<std::vector<Hit>> hits(P);
  long ray_groups = 0, ray_hits = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(+ : ray_groups, ray_hits)
  for (int i = 0; i < P; ++i) {
    if (group[i] != i) continue;
    ray_groups += 1;
    V2 x = samples[i];
    float lo[3] = {(float)x.x - 2.0e-7f, (float)x.y - 2.0e-7f, -1.0e30f};
    float hi[3] = {(float)x.x + 2.0e-7f, (float)x.y + 2.0e-7f, 1.0e30f};
    int node = 1;
    while (node > 0) {
      int leaf = source_tree.next_leaf(lo, hi, node);
      if (leaf < 0) break;
      int s = source_tree.ids[leaf];
      const Ring &r = src[s];
      bool inside = r.n > 0;
      for (int j = 0; j < r.n && inside; ++j) {
        V2 a = r.v[j], b = r.v[(j + 1) % r.n];
        double c = cross(a, b, x);
        bool owned = b.y > a.y || (b.y == a.y && b.x < a.x);
        if (c < 0.0 || (c == 0.0 && !owned)) inside = false;
      }
      if (inside) hits[i].push_back({value(&score[4 * s], x), s});
    }
    std::sort(hits[i].begin(), hits[i].end(), [](const Hit &a, const Hit &b) {
      return a.depth < b.depth || (a.depth == b.depth && a.source < b.source);
    });
    ray_hits += (long)hits[i].size();
  }
  counts[4] = ray_groups;
  counts[5] = ray_hits;
  int layers = 0, limit = layers_limit > 0 ? layers_limit : (1 << 30);
  for (int i = 0; i < P; ++i) {
    int s = out.src[reps[i]];
    V2 x = samples[i];
    double z = value(&score[4 * s], x);
    const auto &h = hits[group[i]];
    int low = 0, high = (int)h.size();
    while (low < high) {
      int mid = (low + high) / 2;
      double hz = value(&score[4 * h[mid].source], x);
      if (hz < z || (hz == z && h[mid].source < s))
        low = mid + 1;
      else
        high = mid;
    }
    out.layer[i] = low;
    if (low >= limit)
      out.polys[i].n = 0;
    else
      layers = std::max(layers, low + 1);
  }
  counts[2] = layers;
  double t3 = now();
  // ---- meshing (kernels.py:941-1124) ---------------------------------------
  std::vector<V2> upts;
  std::vector<Box> pboxes;
  {
    std::map<std::tuple<int, double, double>, int> seen;
    for (int i = 0; i < P; ++i) {
      const Ring &p = out.polys[i];
      for (int j = 0; j < p.n; ++j) {
        auto key = std::make_tuple(out.layer[i], p.v[j].x == 0.0 ? 0.0 : p.v[j].x, p.v[j].y == 0.0 ? 0.0 : p.v[j].y);
        if (seen.emplace(key, (int)upts.size()).second) {
          upts.push_back(p.v[j]);
          Box b;
          float l = (float)out.layer[i];
          b.lo[0] = (float)p.v[j].x - 2.0e-7f, b.lo[1] = (float)p.v[j].y - 2.0e-7f, b.lo[2] = l;
          b.hi[0] = (float)p.v[j].x + 2.0e-7f, b.hi[1] = (float)p.v[j].y + 2.0e-7f, b.hi[2] = l;
          pboxes.push_back(b);
        }
      }
    }
  }
  counts[9] = (long)upts.size();
  Tree point_tree;
  point_tree.build(pboxes, true);
  // inverse affine camera (rows of [R t; 0 1]^-1)
  double inv[12];
  {
    double R[3][3], I[3][3];
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) R[r][c] = M[4 * r + c];
    double d = R[0][0] * (R[1][1] * R[2][2] - R[1][2] * R[2][1]) - R[0][1] * (R[1][0] * R[2][2] - R[1][2] * R[2][0]) +
               R[0][2] * (R[1][0] * R[2][1] - R[1][1] * R[2][0]);
    I[0][0] = (R[1][1] * R[2][2] - R[1][2] * R[2][1]) / d;
    I[0][1] = (R[0][2] * R[2][1] - R[0][1] * R[2][2]) / d;
    I[0][2] = (R[0][1] * R[1][2] - R[0][2] * R[1][1]) / d;
    I[1][0] = (R[1][2] * R[2][0] - R[1][0] * R[2][2]) / d;
    I[1][1] = (R[0][0] * R[2][2] - R[0][2] * R[2][0]) / d;
    I[1][2] = (R[0][2] * R[1][0] - R[0][0] * R[1][2]) / d;
    I[2][0] = (R[1][0] * R[2][1] - R[1][1] * R[2][0]) / d;
    I[2][1] = (R[0][1] * R[2][0] - R[0][0] * R[2][1]) / d;
    I[2][2] = (R[0][0] * R[1][1] - R[0][1] * R[1][0]) / d;
    for (int r = 0; r < 3; ++r) {
      double t = -(I[r][0] * M[3] + I[r][1] * M[7] + I[r][2] * M[11]);
      for (int c = 0; c < 3; ++c) inv[4 * r + c] = I[r][c];
      inv[4 * r + 3] = t;
    }
  }
  auto lift = [&](V2 x, const double *sc, double *o) {
    double u = S[0] + x.x * (S[2] - S[0]);
    double v = S[1] + x.y * (S[3] - S[1]);
    double z = value(sc, x);
    if (S[4] != 0.0) {
      z = -1.0 / z;
      u *= z;
      v *= z;
    }
    for (int r = 0; r < 3; ++r) o[r] = inv[4 * r] * u + inv[4 * r + 1] * v + inv[4 * r + 2] * z + inv[4 * r + 3] * 1.0;
  };
  std::vector<std::vector<std::pair<int, std::array<V2, 3>>>> tris(P);
#pragma omp parallel for schedule(dynamic, 64) reduction(| : error)
  for (int i = 0; i < P; ++i) {
    const Ring &p = out.polys[i];
    V2 center{0.0, 0.0};
    for (int k = 0; k < p.n; ++k) {
      center.x += p.v[k].x / (double)p.n;
      center.y += p.v[k].y / (double)p.n;
    }
    for (int edge = 0; edge < p.n; ++edge) {
      V2 a = p.v[edge], b = p.v[(edge + 1) % p.n];
      V2 d{b.x - a.x, b.y - a.y};
      double length2 = d.x * d.x + d.y * d.y, l = std::sqrt(length2);
      float lo[3] = {(float)std::fmin(a.x, b.x) - 2.0e-7f, (float)std::fmin(a.y, b.y) - 2.0e-7f, (float)out.layer[i]};
      float hi[3] = {(float)std::fmax(a.x, b.x) + 2.0e-7f, (float)std::fmax(a.y, b.y) + 2.0e-7f, (float)out.layer[i]};
      std::vector<double> params;
      int node = 1;
      while (node > 0 && length2 > 0.0) {
        int leaf = point_tree.next_leaf(lo, hi, node);
        if (leaf < 0) break;
        V2 x = upts[point_tree.ids[leaf]];
        double t = ((x.x - a.x) * d.x + (x.y - a.y) * d.y) / length2;
        if (t > EPS && t < 1.0 - EPS && std::fabs(cross(a, b, x)) <= EPS * l) params.push_back(t);
      }
      if (params.size() > 128) error |= 128;
      std::sort(params.begin(), params.end());
      double previous = 0.0;
      V2 cursor = a;
      for (size_t step = 0; step <= params.size(); ++step) {
        double following = step < params.size() ? params[step] : 1.0;
        if (following <= previous + EPS) continue;
        V2 next{a.x + following * d.x, a.y + following * d.y};
        if (std::fabs(cross(center, cursor, next)) > EPS * EPS) tris[i].push_back({i, {center, cursor, next}});
        cursor = next;
        previous = following;
      }
    }
  }
  long T = 0;
  for (int i = 0; i < P; ++i) {
    for (auto &t : tris[i]) {
      if (T < cap_tri) {
        const double *sc = &score[4 * out.src[i]];
        for (int c = 0; c < 3; ++c) {
          tri_uv[6 * T + 2 * c] = t.second[c].x;
          tri_uv[6 * T + 2 * c + 1] = t.second[c].y;
          lift(t.second[c], sc, tri_world + 9 * T + 3 * c);
        }
        tri_src[T] = out.src[i];
        tri_layer[T] = out.layer[i];
      }
      ++T;
    }
  }
  for (int i = 0; i < P && i < cap_poly; ++i) {
    poly_n[i] = out.polys[i].n;
    poly_src[i] = out.src[i];
    poly_layer[i] = out.layer[i];
    for (int j = 0; j < out.polys[i].n; ++j) {
      poly_xy[(long)i * 64 + 2 * j] = out.polys[i].v[j].x;
      poly_xy[(long)i * 64 + 2 * j + 1] = out.polys[i].v[j].y;
    }
  }
  if (P > cap_poly) error |= 8;
  if (T > cap_tri) error |= 32;
  counts[0] = P;
  counts[1] = T;
  counts[6] = error;
  double t4 = now();
  seconds[0] = t1 - t0;
  seconds[1] = t2 - t1;
  seconds[2] = t3 - t2;
  seconds[3] = t4 - t3;
  return 0;
}

}  // extern "C"
