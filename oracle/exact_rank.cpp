// ORACLE — test infrastructure only (see expansion.hpp). Built into
// oracle/liboracle.so by oracle/Makefile; loaded by tests/, the smoke check and
// bench.py's CPU-baseline leg, never by the product path.
//
// Exact per-point depth ranks, straight from the definition of the layers
// (geometric_depth_peeling_design.md:99-122): at an image point x, the rank of
// a covering source s is the number of covering sources strictly nearer than s
// at x. The reference's exact solver decides the same predicates with CGAL
// rationals (src/bvh_peel.cpp:60-79 side_sign/score_sign, ray_order :583-599,
// scene semantics src/camera.cpp:74-229); its arrangement oracle records the
// resulting front-to-back order per cell (src/arrangement_oracle.cpp:55-69).
// Here every predicate is evaluated exactly on the binary64 inputs that the
// device route receives, so no rounding of the scene enters the oracle:
//
//  * camera coordinates are exact expansions of M * [x y z 1];
//  * orthographic coverage: strict orientation of x against the projected
//    triangle edges; depth: barycentric interpolation of camera z, compared
//    by cross-multiplication (no division);
//  * perspective coverage: the ray d = (u, v, 1) lies strictly inside the cone
//    over the camera-space triangle and hits it at camera z >= near (the near
//    clip of src/camera.cpp:120-133); depth: camera z of the hit,
//    z = det[A,B,C] / (det[B,C,d] + det[C,A,d] + det[A,B,d]);
//  * nearer means smaller camera z, i.e. smaller depth score (orthographic
//    score = z, perspective score = -1/z; src/camera.cpp:138-141);
//  * equal depths are ordered by source index (the device tie rule of
//    warp_peel/event_kernels.py:342-385) and flagged.
//
// Query points are given in normalised window coordinates q in [0,1]^2; the
// image point is x = window_min + q * (window_max - window_min), evaluated
// exactly. Points on a source edge, the near plane or the window border are
// flagged ambiguous.
#include "expansion.hpp"

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <utility>
#include <vector>

using namespace oracle;

namespace {

enum : int {
  FLAG_EDGE = 1,    // the point lies on an edge / near plane of this source
  FLAG_TIE = 2,     // equal depth with another covering source
  FLAG_WINDOW = 4,  // point outside or on the border of the window
};

struct Vec3 {
  Exp x, y, z;
};

struct Source {
  Vec3 v[3];             // exact camera-space vertices
  double lo[2], hi[2];   // conservative normalised bounds (for the grid only)
  bool empty = true;     // no positive-area footprint at all
};

struct Scene {
  bool perspective = false;
  double near = 0.01;
  double w[4];
  std::vector<Source> sources;
  // uniform grid over [0,1]^2 for candidate lookup
  int g = 1;
  std::vector<int> cell_start, cell_items;
};

inline Exp orient2(const Exp &ax, const Exp &ay, const Exp &bx, const Exp &by, const Exp &px,
                   const Exp &py) {
  return sub(mul(sub(bx, ax), sub(py, ay)), mul(sub(by, ay), sub(px, ax)));
}

inline Exp det3(const Vec3 &a, const Vec3 &b, const Vec3 &c) {
  // a . (b x c)
  Exp cx = sub(mul(b.y, c.z), mul(b.z, c.y));
  Exp cy = sub(mul(b.z, c.x), mul(b.x, c.z));
  Exp cz = sub(mul(b.x, c.y), mul(b.y, c.x));
  return add(add(mul(a.x, cx), mul(a.y, cy)), mul(a.z, cz));
}

// Depth of a covering source at the query point, as the exact fraction num/den
// (den > 0 after normalisation of signs). Orthographic: z = N / D with
// N = sum(barycentric areas * z), D = orient(a, b, c). Perspective: z = D / S.
struct Hit {
  int source;
  Exp num, den;  // depth = num / den, den > 0
  double approx;
};

int compare_depth(const Hit &a, const Hit &b) {
  // sign(a.num / a.den - b.num / b.den) with positive denominators
  return sub(mul(a.num, b.den), mul(b.num, a.den)).sign();
}

void build_grid(Scene &s) {
  int n = (int)s.sources.size();
  int g = 1;
  while (g * g < n / 2 && g < 1024) g *= 2;
  s.g = g;
  std::vector<int> counts(g * g + 1, 0);
  auto range = [&](const Source &src, int &x0, int &x1, int &y0, int &y1) {
    x0 = std::max(0, std::min(g - 1, (int)std::floor(src.lo[0] * g)));
    x1 = std::max(0, std::min(g - 1, (int)std::floor(src.hi[0] * g)));
    y0 = std::max(0, std::min(g - 1, (int)std::floor(src.lo[1] * g)));
    y1 = std::max(0, std::min(g - 1, (int)std::floor(src.hi[1] * g)));
  };
  for (auto &src : s.sources) {
    if (src.empty) continue;
    int x0, x1, y0, y1;
    range(src, x0, x1, y0, y1);
    for (int y = y0; y <= y1; ++y)
      for (int x = x0; x <= x1; ++x) counts[y * g + x + 1]++;
  }
  for (int i = 0; i < g * g; ++i) counts[i + 1] += counts[i];
  s.cell_start = counts;
  s.cell_items.assign(counts[g * g], 0);
  std::vector<int> fill(counts.begin(), counts.end() - 1);
  for (int i = 0; i < n; ++i) {
    auto &src = s.sources[i];
    if (src.empty) continue;
    int x0, x1, y0, y1;
    range(src, x0, x1, y0, y1);
    for (int y = y0; y <= y1; ++y)
      for (int x = x0; x <= x1; ++x) s.cell_items[fill[y * g + x]++] = i;
  }
}

Scene make_scene(const double *V, int nv, const int *F, int nf, const double *M,
                 const double *settings) {
  Scene s;
  s.perspective = settings[4] != 0.0;
  s.near = settings[5];
  for (int k = 0; k < 4; ++k) s.w[k] = settings[k];
  s.sources.resize(nf);
  double sx = s.w[2] - s.w[0], sy = s.w[3] - s.w[1];
  for (int f = 0; f < nf; ++f) {
    Source &src = s.sources[f];
    double cam[3][3];
    bool ok = true;
    for (int j = 0; j < 3; ++j) {
      int vi = F[3 * f + j];
      if (vi < 0 || vi >= nv) {
        ok = false;
        break;
      }
      const double *p = V + 3 * vi;
      Exp *dst[3] = {&src.v[j].x, &src.v[j].y, &src.v[j].z};
      for (int r = 0; r < 3; ++r) {
        Exp e(M[4 * r + 3]);
        for (int k = 0; k < 3; ++k) {
          double hi, lo;
          two_prod(M[4 * r + k], p[k], hi, lo);
          e = add(e, two(hi, lo));
        }
        *dst[r] = e;
        cam[j][r] = e.approx();
      }
    }
    if (!ok) continue;
    // Conservative footprint in normalised coordinates for grid lookup only.
    double lo[2] = {1e300, 1e300}, hi[2] = {-1e300, -1e300};
    bool any = false;
    if (!s.perspective) {
      for (int j = 0; j < 3; ++j) {
        double u = (cam[j][0] - s.w[0]) / sx, v = (cam[j][1] - s.w[1]) / sy;
        lo[0] = std::min(lo[0], u), hi[0] = std::max(hi[0], u);
        lo[1] = std::min(lo[1], v), hi[1] = std::max(hi[1], v);
      }
      any = true;
    } else {
      // near-clipped ring in floating point, with a generous margin below
      double ring[4][3];
      int n = 0;
      double nearm = s.near * (1 - 1e-9);
      for (int j = 0; j < 3; ++j) {
        const double *a = cam[j], *b = cam[(j + 1) % 3];
        double da = a[2] - nearm, db = b[2] - nearm;
        if (da >= 0) {
          std::memcpy(ring[n++], a, sizeof(double) * 3);
        }
        if ((da > 0 && db < 0) || (da < 0 && db > 0)) {
          double t = da / (da - db);
          for (int k = 0; k < 3; ++k) ring[n][k] = a[k] + t * (b[k] - a[k]);
          ++n;
        }
      }
      for (int j = 0; j < n; ++j) {
        double z = std::max(ring[j][2], nearm);
        double u = (ring[j][0] / z - s.w[0]) / sx, v = (ring[j][1] / z - s.w[1]) / sy;
        lo[0] = std::min(lo[0], u), hi[0] = std::max(hi[0], u);
        lo[1] = std::min(lo[1], v), hi[1] = std::max(hi[1], v);
        any = true;
      }
    }
    if (!any) continue;
    for (int k = 0; k < 2; ++k) {
      double pad = 1e-7 * (1.0 + std::fabs(lo[k]) + std::fabs(hi[k]));
      src.lo[k] = lo[k] - pad;
      src.hi[k] = hi[k] + pad;
    }
    if (src.hi[0] < 0 || src.hi[1] < 0 || src.lo[0] > 1 || src.lo[1] > 1) continue;
    src.empty = false;
  }
  build_grid(s);
  return s;
}

// Coverage of image point (ux, uy) by source f. Returns 1 (strictly covered),
// 0 (not covered) or -1 (on the boundary: ambiguous). Fills the depth fraction.
int cover(const Scene &s, const Source &src, int f, const Exp &ux, const Exp &uy, Hit &hit) {
  hit.source = f;
  if (!s.perspective) {
    const Vec3 &a = src.v[0], &b = src.v[1], &c = src.v[2];
    Exp D = orient2(a.x, a.y, b.x, b.y, c.x, c.y);
    int sd = D.sign();
    if (sd == 0) return 0;  // zero projected area contributes no coverage
    Exp e0 = orient2(b.x, b.y, c.x, c.y, ux, uy);  // barycentric weight of a
    Exp e1 = orient2(c.x, c.y, a.x, a.y, ux, uy);  // of b
    Exp e2 = orient2(a.x, a.y, b.x, b.y, ux, uy);  // of c
    int s0 = e0.sign() * sd, s1 = e1.sign() * sd, s2 = e2.sign() * sd;
    if (s0 < 0 || s1 < 0 || s2 < 0) return 0;
    if (s0 == 0 || s1 == 0 || s2 == 0) return -1;
    Exp N = add(add(mul(e0, a.z), mul(e1, b.z)), mul(e2, c.z));
    if (sd < 0) {
      N = neg(N);
      D = neg(D);
    }
    hit.num = N;
    hit.den = D;
    hit.approx = N.approx() / D.approx();
    return 1;
  }
  const Vec3 &A = src.v[0], &B = src.v[1], &C = src.v[2];
  Vec3 d{ux, uy, Exp(1.0)};
  Exp D = det3(A, B, C);
  int sd = D.sign();
  if (sd == 0) return 0;  // plane through the camera centre: edge-on
  Exp ga = det3(B, C, d), gb = det3(C, A, d), gc = det3(A, B, d);
  int s0 = ga.sign() * sd, s1 = gb.sign() * sd, s2 = gc.sign() * sd;
  if (s0 < 0 || s1 < 0 || s2 < 0) return 0;
  if (s0 == 0 || s1 == 0 || s2 == 0) return -1;
  Exp S = add(add(ga, gb), gc);  // same sign as D; hit z = D / S > 0
  if (sd < 0) {
    D = neg(D);
    S = neg(S);
  }
  // z >= near  <=>  D - near * S >= 0 (S > 0)
  int clip = sub(D, scale(S, s.near)).sign();
  if (clip < 0) return 0;
  if (clip == 0) return -1;
  hit.num = D;
  hit.den = S;
  hit.approx = D.approx() / S.approx();
  return 1;
}

}  // namespace

extern "C" {

// Exact ranks at query points. points: nq x 2 normalised coordinates.
// Output is CSR: offsets[nq + 1]; for entry k in [offsets[i], offsets[i+1]):
// sources[k] in front-to-back order, ranks[k] = k - offsets[i], flags[k].
// point_flags[i] is the OR of its entries' flags (plus FLAG_WINDOW / FLAG_EDGE
// for a boundary contact with a source not counted as covering).
// Returns the total number of entries; if it exceeds cap, nothing beyond cap is
// written and the caller retries with a larger buffer. Negative on bad input.
long oracle_exact_ranks(const double *vertices, int nv, const int *faces, int nf,
                        const double *camera, const double *settings, const double *points,
                        int nq, long *offsets, int *sources, int *ranks, int *flags,
                        int *point_flags, long cap) {
  if (nv < 0 || nf < 0 || nq < 0) return -1;
  Scene s = make_scene(vertices, nv, faces, nf, camera, settings);
  double sx0 = s.w[0], sy0 = s.w[1];
  Exp wx = two(s.w[2], -s.w[0]), wy = two(s.w[3], -s.w[1]);
  // Points are independent: each thread ranks a share of them into its own
  // lists; the CSR is then written in point order.
  std::vector<std::vector<std::pair<int, int>>> lists(nq);
#pragma omp parallel
  {
    std::vector<Hit> hits;
#pragma omp for schedule(dynamic, 64)
    for (int i = 0; i < nq; ++i) {
      double qx = points[2 * i], qy = points[2 * i + 1];
      int pflag = 0;
      hits.clear();
      if (!(qx > 0 && qx < 1 && qy > 0 && qy < 1)) {
        pflag |= FLAG_WINDOW;
      } else {
        Exp ux = add(Exp(sx0), scale(wx, qx));
        Exp uy = add(Exp(sy0), scale(wy, qy));
        int gx = std::min(s.g - 1, (int)(qx * s.g)), gy = std::min(s.g - 1, (int)(qy * s.g));
        int cell = gy * s.g + gx;
        for (int k = s.cell_start[cell]; k < s.cell_start[cell + 1]; ++k) {
          int f = s.cell_items[k];
          const Source &src = s.sources[f];
          if (qx < src.lo[0] || qx > src.hi[0] || qy < src.lo[1] || qy > src.hi[1]) continue;
          Hit h;
          int c = cover(s, src, f, ux, uy, h);
          if (c < 0) pflag |= FLAG_EDGE;
          if (c > 0) hits.push_back(std::move(h));
        }
      }
      // Exact front-to-back order; equal depths by source index.
      std::vector<int> order(hits.size());
      for (size_t k = 0; k < hits.size(); ++k) order[k] = (int)k;
      std::vector<int> tie(hits.size(), 0);
      std::sort(order.begin(), order.end(), [&](int a, int b) {
        const Hit &ha = hits[a], &hb = hits[b];
        double da = ha.approx, db = hb.approx;
        double scale_ab = std::fabs(da) + std::fabs(db);
        if (std::isfinite(da) && std::isfinite(db) && std::fabs(da - db) > 1e-6 * scale_ab + 1e-300)
          return da < db;
        int c = compare_depth(ha, hb);
        if (c == 0) {
          tie[a] = tie[b] = 1;
          return ha.source < hb.source;
        }
        return c < 0;
      });
      // std::sort may not compare every adjacent pair; confirm ties explicitly.
      for (size_t k = 0; k + 1 < order.size(); ++k)
        if (compare_depth(hits[order[k]], hits[order[k + 1]]) == 0) tie[order[k]] = tie[order[k + 1]] = 1;
      auto &out = lists[i];
      out.reserve(order.size());
      for (size_t k = 0; k < order.size(); ++k) {
        int f = tie[order[k]] ? FLAG_TIE : 0;
        pflag |= f;
        out.emplace_back(hits[order[k]].source, f);
      }
      point_flags[i] = pflag;
    }
  }
  long total = 0;
  for (int i = 0; i < nq; ++i) {
    offsets[i] = total;
    for (size_t k = 0; k < lists[i].size(); ++k) {
      if (total < cap) {
        sources[total] = lists[i][k].first;
        ranks[total] = (int)k;
        flags[total] = lists[i][k].second;
      }
      ++total;
    }
  }
  offsets[nq] = total;
  return total;
}

}  // extern "C"
