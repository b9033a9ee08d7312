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

#include "cpu_prepare.hpp"

namespace {

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
  Prep prep;
  if (int bits = prepare(V, nv, F, nf, M, S, prep)) {
    counts[6] = bits;
    return 0;
  }
  const std::vector<int> &canonical = prep.canonical;
  const std::vector<Ring> &src = prep.src;
  const std::vector<double> &score = prep.score;
  const std::vector<int> &facing = prep.facing, &clipped = prep.clipped, &masks = prep.masks;
  const std::vector<Box> &src_box = prep.src_box;
  auto contour_edge = [&](int s, int edge) { return prep.contour_edge(s, edge); };
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
