// ORACLE — test infrastructure only (see expansion.hpp). Built into
// oracle/liboracle.so; bench.py times it as the CPU baseline and tests/ check
// it against the exact ranks. Never linked by the product.
//
// CPU restatement of the reference's fast CPU method, `fast-bvh-peel
// --resolved-mesh` (src/fast_peel.cpp, src/peel_common.h): projected geometry
// on a fixed-point grid of 2^bits cells per axis with exact integer
// predicates, binary64 depth planes and constructions.
//  projection + topology (untimed setup)     fast_peel.cpp:159-318, peel_common.h:94-133
//  clean (hull repair, ribbon removal)       fast_peel.cpp:48-105
//  split, separated, cut (tie band)          fast_peel.cpp:349-443
//  FastPolicy: boxes, contour guards, rays   fast_peel.cpp:480-605
//  median-split hierarchy, near-first query
//  with rank-subtree retirement              peel_common.h:20-92, 172-200
//  classification, per-layer scheduler       peel_common.h:135-170, 344-460
//  conforming boundaries, centroid fans      peel_common.h:203-258, fast_peel.cpp:445-478
// The timed scope is the reference's `algorithm_seconds`: classification and
// every layer's hierarchy, clipping and meshing, after projection and
// topology (fast_peel.cpp:653-660). Threads: OpenMP, dynamic schedules as the
// reference's.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <limits>
#include <map>
#include <numeric>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

using Clock = std::chrono::steady_clock;
double since(Clock::time_point t) { return std::chrono::duration<double>(Clock::now() - t).count(); }
constexpr double EPS64 = std::numeric_limits<double>::epsilon();

struct GP {
  int32_t x = 0, y = 0;
};
inline bool operator==(GP a, GP b) { return a.x == b.x && a.y == b.y; }
inline bool operator!=(GP a, GP b) { return !(a == b); }
inline bool operator<(GP a, GP b) { return a.x == b.x ? a.y < b.y : a.x < b.x; }

// Exact orientation on the grid: coordinates lie in [0, 2^30], so the
// products fit in int64.
inline int64_t orient(GP a, GP b, GP p) {
  return int64_t(b.x - a.x) * (p.y - a.y) - int64_t(b.y - a.y) * (p.x - a.x);
}

// A ring with inline storage for the common case and heap storage beyond it.
class Ring {
 public:
  static constexpr int INLINE = 12;
  Ring() = default;
  Ring(const Ring &o) { *this = o; }
  Ring &operator=(const Ring &o) {
    if (this == &o) return *this;
    n_ = 0;
    heap_.clear();
    for (int i = 0; i < o.size(); ++i) push(o[i]);
    return *this;
  }
  Ring(Ring &&o) noexcept { *this = std::move(o); }
  Ring &operator=(Ring &&o) noexcept {
    n_ = o.n_;
    std::memcpy(inl_, o.inl_, sizeof(inl_));
    heap_ = std::move(o.heap_);
    o.n_ = 0;
    return *this;
  }
  int size() const { return n_; }
  bool empty() const { return n_ == 0; }
  GP *data() { return n_ > INLINE ? heap_.data() : inl_; }
  const GP *data() const { return n_ > INLINE ? heap_.data() : inl_; }
  GP &operator[](int i) { return data()[i]; }
  const GP &operator[](int i) const { return data()[i]; }
  void clear() {
    n_ = 0;
    heap_.clear();
  }
  void push(GP p) {
    if (n_ < INLINE) {
      inl_[n_++] = p;
      return;
    }
    if (n_ == INLINE) heap_.assign(inl_, inl_ + INLINE);
    heap_.push_back(p);
    ++n_;
  }
  void resize(int n) {  // shrink only
    if (n >= n_) return;
    if (n_ > INLINE && n <= INLINE) std::memcpy(inl_, heap_.data(), sizeof(GP) * n), heap_.clear();
    else if (n > INLINE) heap_.resize(n);
    n_ = n;
  }
  void reverse() { std::reverse(data(), data() + n_); }

 private:
  int n_ = 0;
  GP inl_[INLINE];
  std::vector<GP> heap_;
};

int64_t twice_area(const Ring &r) {
  int64_t a = 0;
  for (int j = 1; j + 1 < r.size(); ++j) a += orient(r[0], r[j], r[j + 1]);
  return a;
}

struct Tally {
  uint64_t candidates = 0, cuts = 0, repairs = 0, collapsed = 0, ties = 0, occluded = 0;
};

// clean (fast_peel.cpp:48-105): drop repeated points, repair a non-convex
// ring by its convex hull, and empty rings of at most one grid cell of area
// or thinner than two cells (L1) across some edge.
void clean(Ring &r, Tally &t) {
  int m = 0;
  for (int i = 0; i < r.size(); ++i)
    if (m == 0 || r[i] != r[m - 1]) r[m++] = r[i];
  r.resize(m);
  if (r.size() > 1 && r[0] == r[r.size() - 1]) r.resize(r.size() - 1);
  if (r.size() < 3) {
    if (!r.empty()) ++t.collapsed;
    r.clear();
    return;
  }
  const int n = r.size();
  bool convex = true;
  for (int i = 0; i < n && convex; ++i) convex = orient(r[i], r[(i + 1) % n], r[(i + 2) % n]) >= 0;
  if (!convex) {
    ++t.repairs;
    std::vector<GP> pts(r.data(), r.data() + n);
    std::sort(pts.begin(), pts.end());
    pts.erase(std::unique(pts.begin(), pts.end()), pts.end());
    std::vector<GP> hull;
    for (GP p : pts) {  // lower chain
      while (hull.size() > 1 && orient(hull[hull.size() - 2], hull.back(), p) <= 0) hull.pop_back();
      hull.push_back(p);
    }
    size_t lower = hull.size();
    for (int i = (int)pts.size() - 2; i >= 0; --i) {  // upper chain
      while (hull.size() > lower && orient(hull[hull.size() - 2], hull.back(), pts[i]) <= 0) hull.pop_back();
      hull.push_back(pts[i]);
    }
    hull.pop_back();
    r.clear();
    for (GP p : hull) r.push(p);
  }
  if (r.size() < 3 || twice_area(r) <= 2) {
    ++t.collapsed;
    r.clear();
    return;
  }
  for (int i = 0; i < r.size(); ++i) {
    GP a = r[i], b = r[(i + 1) % r.size()];
    int64_t width = 2 * (int64_t(std::abs(b.x - a.x)) + std::abs(b.y - a.y));
    if (width == 0) continue;
    int64_t reach = 0;
    for (int k = 0; k < r.size(); ++k) reach = std::max(reach, orient(a, b, r[k]));
    if (reach <= width) {
      ++t.collapsed;
      r.clear();
      return;
    }
  }
}

struct Box {
  int32_t lo[2] = {INT32_MAX, INT32_MAX}, hi[2] = {INT32_MIN, INT32_MIN};
  double near = INFINITY, far = -INFINITY;
  int64_t centre(int axis) const { return int64_t(lo[axis]) + hi[axis]; }
  void add(GP p) {
    lo[0] = std::min(lo[0], p.x), lo[1] = std::min(lo[1], p.y);
    hi[0] = std::max(hi[0], p.x), hi[1] = std::max(hi[1], p.y);
  }
  void add(const Box &b) {
    for (int k = 0; k < 2; ++k) lo[k] = std::min(lo[k], b.lo[k]), hi[k] = std::max(hi[k], b.hi[k]);
    near = std::min(near, b.near), far = std::max(far, b.far);
  }
};
inline bool overlap(const Box &a, const Box &b) {
  return a.lo[0] <= b.hi[0] && b.lo[0] <= a.hi[0] && a.lo[1] <= b.hi[1] && b.lo[1] <= a.hi[1];
}
Box bound(const Ring &r) {
  Box b;
  for (int i = 0; i < r.size(); ++i) b.add(r[i]);
  return b;
}

// split (fast_peel.cpp:349-376): the parts of a ring where eval >= 0 and
// <= 0, new vertices rounded to the nearest grid point, both cleaned.
template <class Eval>
void split(const Ring &r, Eval eval, Ring &in, Ring &out, Tally &t) {
  double d[64];
  std::vector<double> big;
  double *dv = d;
  if (r.size() > 64) big.resize(r.size()), dv = big.data();
  bool pos = false, neg = false;
  for (int i = 0; i < r.size(); ++i) {
    dv[i] = eval(r[i]);
    pos |= dv[i] > 0;
    neg |= dv[i] < 0;
  }
  in.clear();
  out.clear();
  if (!neg) {
    in = r;
    return;
  }
  if (!pos) {
    out = r;
    return;
  }
  for (int i = 0; i < r.size(); ++i) {
    int j = (i + 1) % r.size();
    if (dv[i] >= 0) in.push(r[i]);
    if (dv[i] <= 0) out.push(r[i]);
    if ((dv[i] > 0 && dv[j] < 0) || (dv[i] < 0 && dv[j] > 0)) {
      double s = dv[i] / (dv[i] - dv[j]);
      GP p{int32_t(std::llround(r[i].x + s * (r[j].x - r[i].x))), int32_t(std::llround(r[i].y + s * (r[j].y - r[i].y)))};
      in.push(p);
      out.push(p);
    }
  }
  clean(in, t);
  clean(out, t);
}

// some edge of a has no vertex of b strictly on its left
bool separated(const Ring &a, const Ring &b) {
  for (int i = 0; i < a.size(); ++i) {
    GP u = a[i], v = a[(i + 1) % a.size()];
    bool left = false;
    for (int k = 0; k < b.size() && !left; ++k) left = orient(u, v, b[k]) > 0;
    if (!left) return true;
  }
  return false;
}

struct Source {
  int id;          // external id (input face index)
  double score[3];  // depth plane over grid coordinates
  Ring ring;
};

inline double depth_at(const double *s, GP p) { return s[0] * p.x + s[1] * p.y + s[2]; }

struct Patch {
  int source;
  Ring ring;
};

struct Work {
  std::vector<Ring> visible, hidden;
  Tally tally;
};

// cut (fast_peel.cpp:380-443): remove from the visible pieces the part that
// `cand` covers nearer, inside a binary64 tie band.
void cut(Work &w, const Ring &cand, const Source &target, const Source &other) {
  Box cb = bound(cand);
  double d[3];
  for (int k = 0; k < 3; ++k) d[k] = target.score[k] - other.score[k];
  double band = 32 * EPS64 * std::max({std::fabs(target.score[2]), std::fabs(other.score[2]), 1e-300});
  std::vector<Ring> kept;
  kept.reserve(w.visible.size() + 2);
  Ring in, out;
  for (Ring &p : w.visible) {
    if (!overlap(bound(p), cb) || separated(p, cand) || separated(cand, p)) {
      kept.push_back(std::move(p));
      continue;
    }
    bool pos = false, neg = false;
    for (int i = 0; i < p.size(); ++i) {
      double z = depth_at(d, p[i]);
      pos |= z > band;
      neg |= z < -band;
    }
    if (!pos && !neg) {
      ++w.tally.ties;
      pos = target.id > other.id;
    }
    if (!pos) {
      kept.push_back(std::move(p));
      continue;
    }
    Ring inside = p;
    std::vector<Ring> outside;
    if (neg) {
      split(inside, [&](GP v) { return depth_at(d, v); }, in, out, w.tally);
      inside = std::move(in);
      if (!out.empty()) outside.push_back(std::move(out));
    }
    for (int j = 0; j < cand.size() && !inside.empty(); ++j) {
      GP a = cand[j], b = cand[(j + 1) % cand.size()];
      split(inside, [&](GP v) { return double(orient(a, b, v)); }, in, out, w.tally);
      inside = std::move(in);
      if (!out.empty()) outside.push_back(std::move(out));
    }
    if (inside.empty()) {
      kept.push_back(std::move(p));
    } else {
      ++w.tally.cuts;
      w.hidden.push_back(std::move(inside));
      for (Ring &o : outside) kept.push_back(std::move(o));
    }
  }
  w.visible = std::move(kept);
}

// Median-split hierarchy over boxes (peel_common.h:20-56): leaves of at most
// four, split at the median centre along the wider axis (ties by index).
struct Hierarchy {
  struct Node {
    Box box;
    int left = -1, right = -1, begin = 0, end = 0;
  };
  std::vector<Node> nodes;
  std::vector<int> ids;
  std::vector<int> max_rank;
  const std::vector<Box> *boxes = nullptr;
  explicit Hierarchy(const std::vector<Box> &b) : boxes(&b) {
    ids.resize(b.size());
    std::iota(ids.begin(), ids.end(), 0);
    nodes.reserve(b.size() / 2 + 1);
    if (!b.empty()) build(0, (int)b.size());
  }
  int build(int begin, int end) {
    int i = (int)nodes.size();
    nodes.emplace_back();
    Box box;
    for (int j = begin; j < end; ++j) box.add((*boxes)[ids[j]]);
    nodes[i].box = box;
    if (end - begin <= 4) {
      nodes[i].begin = begin, nodes[i].end = end;
      return i;
    }
    int axis = (box.hi[1] - box.lo[1]) > (box.hi[0] - box.lo[0]) ? 1 : 0, mid = (begin + end) / 2;
    std::nth_element(ids.begin() + begin, ids.begin() + mid, ids.begin() + end, [&](int a, int b) {
      int64_t x = (*boxes)[a].centre(axis), y = (*boxes)[b].centre(axis);
      return x == y ? a < b : x < y;
    });
    int l = build(begin, mid), r = build(mid, end);
    nodes[i].left = l, nodes[i].right = r;
    return i;
  }
  // largest rank below each node; a node whose largest rank is below the
  // layer holds only peeled whole sources (peel_common.h:59-70)
  void summarize(const std::vector<int> &ranks) {
    max_rank.assign(nodes.size(), -1);
    for (int i = (int)nodes.size() - 1; i >= 0; --i) {
      const Node &n = nodes[i];
      if (n.left < 0)
        for (int j = n.begin; j < n.end; ++j) max_rank[i] = std::max(max_rank[i], ranks[ids[j]]);
      else
        max_rank[i] = std::max(max_rank[n.left], max_rank[n.right]);
    }
  }
  template <class F>
  void visit(const Box &b, F &&f, int i = 0) const {
    if (nodes.empty() || !overlap(nodes[i].box, b)) return;
    const Node &n = nodes[i];
    if (n.left < 0) {
      for (int j = n.begin; j < n.end; ++j)
        if (overlap((*boxes)[ids[j]], b)) f(ids[j]);
    } else {
      visit(b, f, n.left);
      visit(b, f, n.right);
    }
  }
};

struct Scene {
  int bits = 30;
  std::vector<Source> sources;
  std::vector<std::pair<GP, GP>> contours;
  int collapsed = 0;
  // mesh topology (peel_common.h:104-124): boundary, silhouette, non-manifold
  // or inconsistent, ignored internal, and window/near clipping edges
  long boundary = 0, silhouette = 0, nonmanifold = 0, ignored = 0, clipping = 0;
};

// Projection and topology (fast_peel.cpp:159-318): untimed setup.
Scene project(const double *V, const int *F, int nf, const double *M, const double *S, int bits) {
  Scene s;
  s.bits = bits;
  const bool persp = S[4] != 0.0;
  const double near = S[5], grid = double(int64_t(1) << bits);
  struct Incidence {
    int direction, facing;
    GP a, b;
  };
  std::map<std::array<double, 3>, int> ids;
  std::map<std::pair<int, int>, std::vector<Incidence>> edges;
  using P3 = std::array<double, 3>;
  using P2 = std::array<double, 2>;
  for (int f = 0; f < nf; ++f) {
    std::vector<P3> cam;
    int vid[3];
    bool clipped = false;
    for (int j = 0; j < 3; ++j) {
      const double *w = V + 3 * F[3 * f + j];
      P3 world{w[0], w[1], w[2]};
      vid[j] = ids.emplace(world, (int)ids.size()).first->second;
      P3 c;
      for (int i = 0; i < 3; ++i) c[i] = M[4 * i] * w[0] + M[4 * i + 1] * w[1] + M[4 * i + 2] * w[2] + M[4 * i + 3];
      cam.push_back(c);
      clipped |= persp && c[2] < near;
    }
    if (persp) {  // keep z >= near
      std::vector<P3> kept;
      for (size_t i = 0; i < cam.size(); ++i) {
        P3 a = cam[i], b = cam[(i + 1) % cam.size()];
        double da = a[2] - near, db = b[2] - near;
        if (da >= 0) kept.push_back(a);
        if ((da < 0 && db > 0) || (da > 0 && db < 0)) {
          double t = da / (da - db);
          kept.push_back({a[0] + t * (b[0] - a[0]), a[1] + t * (b[1] - a[1]), a[2] + t * (b[2] - a[2])});
        }
      }
      cam = kept;
    }
    if (cam.size() < 3) {
      ++s.collapsed;
      continue;
    }
    std::vector<P2> p;
    std::vector<double> depth;
    for (const P3 &c : cam) {
      double q = persp ? 1 / c[2] : 1;
      p.push_back({(c[0] * q - S[0]) * grid / (S[2] - S[0]), (c[1] * q - S[1]) * grid / (S[3] - S[1])});
      depth.push_back(persp ? -q : c[2]);
    }
    size_t k = 2;
    double det = 0;
    for (; k < p.size(); ++k) {
      det = (p[1][0] - p[0][0]) * (p[k][1] - p[0][1]) - (p[1][1] - p[0][1]) * (p[k][0] - p[0][0]);
      if (det != 0) break;
    }
    if (k == p.size()) {
      ++s.collapsed;
      continue;
    }
    double x1 = p[1][0] - p[0][0], y1 = p[1][1] - p[0][1], x2 = p[k][0] - p[0][0], y2 = p[k][1] - p[0][1];
    Source src;
    src.id = f;
    src.score[0] = ((depth[1] - depth[0]) * y2 - (depth[k] - depth[0]) * y1) / det;
    src.score[1] = (x1 * (depth[k] - depth[0]) - x2 * (depth[1] - depth[0])) / det;
    src.score[2] = depth[0] - src.score[0] * p[0][0] - src.score[1] * p[0][1];
    for (const P2 &v : p) clipped |= v[0] < 0 || v[0] > grid || v[1] < 0 || v[1] > grid;
    GP rounded[3];
    if (!clipped)
      for (int i = 0; i < 3; ++i) rounded[i] = {int32_t(std::llround(p[i][0])), int32_t(std::llround(p[i][1]))};
    for (int axis = 0; axis < 2; ++axis)
      for (int side = 0; side < 2; ++side) {
        std::vector<P2> kept;
        auto value = [&](const P2 &v) { return side ? grid - v[axis] : v[axis]; };
        for (size_t i = 0; i < p.size(); ++i) {
          P2 a = p[i], b = p[(i + 1) % p.size()];
          double da = value(a), db = value(b);
          if (da >= 0) kept.push_back(a);
          if ((da < 0 && db > 0) || (da > 0 && db < 0)) {
            double t = da / (da - db);
            kept.push_back({a[0] + t * (b[0] - a[0]), a[1] + t * (b[1] - a[1])});
          }
        }
        p = kept;
      }
    for (const P2 &v : p)
      src.ring.push({int32_t(std::llround(std::clamp(v[0], 0., grid))), int32_t(std::llround(std::clamp(v[1], 0., grid)))});
    if (src.ring.size() > 2 && twice_area(src.ring) < 0) src.ring.reverse();
    Tally t;
    clean(src.ring, t);
    if (src.ring.empty()) {
      ++s.collapsed;
      continue;
    }
    if (clipped) {
      for (int i = 0; i < src.ring.size(); ++i) s.contours.push_back({src.ring[i], src.ring[(i + 1) % src.ring.size()]});
      s.clipping += src.ring.size();
    } else {
      for (int i = 0; i < 3; ++i) {
        int a = vid[i], b = vid[(i + 1) % 3];
        edges[{std::min(a, b), std::max(a, b)}].push_back({a < b ? 1 : -1, det > 0 ? 1 : -1, rounded[i], rounded[(i + 1) % 3]});
      }
    }
    s.sources.push_back(std::move(src));
  }
  // internal iff exactly two opposite incidences of equal facing (peel_common.h:104-107)
  for (auto &e : edges) {
    const auto &g = e.second;
    if (g.size() == 2 && g[0].direction != g[1].direction && g[0].facing == g[1].facing) {
      ++s.ignored;
      continue;
    }
    if (g.size() == 1)
      ++s.boundary;
    else if (g.size() != 2 || g[0].direction == g[1].direction)
      ++s.nonmanifold;
    else
      ++s.silhouette;
    s.contours.push_back({g[0].a, g[0].b});
  }
  return s;
}

struct Layer {
  std::vector<Patch> patches;
  size_t faces = 0;
  std::vector<double> fan_xy;  // [faces][3][2] grid coordinates, when kept
  std::vector<int> fan_src;
};

// Conforming boundaries and centroid fans of one layer (peel_common.h:203-258).
void mesh(Layer &layer, int threads, bool keep) {
  std::vector<GP> pts;
  for (const Patch &p : layer.patches)
    for (int i = 0; i < p.ring.size(); ++i) pts.push_back(p.ring[i]);
  std::sort(pts.begin(), pts.end());
  pts.erase(std::unique(pts.begin(), pts.end()), pts.end());
  std::vector<Box> boxes(pts.size());
  for (size_t i = 0; i < pts.size(); ++i) boxes[i].add(pts[i]);
  Hierarchy tree(boxes);
  std::vector<std::vector<double>> out(layer.patches.size());
  std::vector<size_t> count(layer.patches.size(), 0);
#pragma omp parallel for schedule(dynamic, 32) num_threads(threads)
  for (long k = 0; k < (long)layer.patches.size(); ++k) {
    const Ring &r = layer.patches[k].ring;
    std::vector<GP> ring;
    for (int i = 0; i < r.size(); ++i) {
      GP a = r[i], b = r[(i + 1) % r.size()];
      Box box;
      box.add(a), box.add(b);
      std::vector<GP> along;
      tree.visit(box, [&](int j) {
        GP v = pts[j];
        if (v != b && orient(a, b, v) == 0) along.push_back(v);
      });
      std::sort(along.begin(), along.end());
      if (b < a) std::reverse(along.begin(), along.end());
      ring.insert(ring.end(), along.begin(), along.end());
    }
    size_t n = ring.size();
    std::vector<double> &f = out[k];
    auto emit = [&](double ax, double ay, GP b, GP c) {
      if (keep) f.insert(f.end(), {ax, ay, double(b.x), double(b.y), double(c.x), double(c.y)});
      ++count[k];
    };
    if (n == 3) {
      emit(ring[0].x, ring[0].y, ring[1], ring[2]);
    } else if (n > 3) {
      double cx = 0, cy = 0;
      for (GP v : ring) cx += v.x, cy += v.y;
      cx /= n, cy /= n;
      for (size_t i = 0; i < n; ++i) emit(cx, cy, ring[i], ring[(i + 1) % n]);
    }
  }
  for (size_t k = 0; k < out.size(); ++k) {
    layer.faces += count[k];
    if (keep) {
      layer.fan_xy.insert(layer.fan_xy.end(), out[k].begin(), out[k].end());
      layer.fan_src.insert(layer.fan_src.end(), count[k], layer.patches[k].source);
    }
  }
}

struct Extraction {
  std::vector<Layer> layers;
  long constant_rank = 0, residual_sources = 0, leftover = 0, ray_candidates = 0;
  Tally tally;
  int error = 0;  // 1: no progress, 2: layer limit
};

std::vector<Box> patch_boxes(const Scene &s, const std::vector<Patch> &patches) {
  std::vector<Box> boxes(patches.size());
  for (size_t i = 0; i < patches.size(); ++i) {
    Box &b = boxes[i];
    b = bound(patches[i].ring);
    const double *sc = s.sources[patches[i].source].score;
    const Ring &r = patches[i].ring;
    for (int k = 0; k < r.size(); ++k) {
      double z = depth_at(sc, r[k]);
      double err = 64 * EPS64 * (std::fabs(sc[0] * r[k].x) + std::fabs(sc[1] * r[k].y) + std::fabs(sc[2]));
      b.near = std::min(b.near, z - err);
      b.far = std::max(b.far, z + err);
    }
  }
  return boxes;
}

// the segment may meet the source, with a four-cell guard (fast_peel.cpp:530-550)
bool touches_contour(const Ring &src, GP e0, GP e1) {
  for (int k = 0; k < src.size(); ++k) {
    GP a = src[k], b = src[(k + 1) % src.size()];
    int64_t margin = 4 * (int64_t(std::abs(b.x - a.x)) + std::abs(b.y - a.y));
    if (orient(a, b, e0) < -margin && orient(a, b, e1) < -margin) return false;
  }
  bool positive = false, negative = false;
  int64_t margin = 4 * (int64_t(std::abs(e1.x - e0.x)) + std::abs(e1.y - e0.y));
  for (int k = 0; k < src.size(); ++k) {
    int64_t v = orient(e0, e1, src[k]);
    positive |= v >= -margin;
    negative |= v <= margin;
  }
  return positive && negative;
}

// Does source j cover the sample (3x the centroid, exact) nearer than the
// target depth? 1 yes, 0 no, -1 ambiguous (edge or tie) (fast_peel.cpp:564-590).
int ray_order(const Source &target, const Source &other, int64_t px, int64_t py) {
  double zt = target.score[0] * (px / 3.) + target.score[1] * (py / 3.) + target.score[2];
  bool edge = false;
  for (int k = 0; k < other.ring.size(); ++k) {
    GP a = other.ring[k], b = other.ring[(k + 1) % other.ring.size()];
    int64_t c = int64_t(b.x - a.x) * (py - 3 * int64_t(a.y)) - int64_t(b.y - a.y) * (px - 3 * int64_t(a.x));
    if (c < 0) return 0;
    edge |= c == 0;
  }
  if (edge) return -1;
  double z = other.score[0] * (px / 3.) + other.score[1] * (py / 3.) + other.score[2];
  double band = 64 * EPS64 * std::max({std::fabs(z), std::fabs(zt), 1e-300});
  if (std::fabs(z - zt) <= band) return -1;
  return z < zt ? 1 : 0;
}

// Near-first query with retired subtrees (peel_common.h:172-200).
void query(const Scene &s, const std::vector<Patch> &input, const Hierarchy &tree, const Patch &target,
           const Box &box, Work &w, const std::vector<int> *ranks, int layer, int i = 0) {
  if (w.visible.empty() || tree.nodes.empty()) return;
  const Hierarchy::Node &node = tree.nodes[i];
  if (ranks && !tree.max_rank.empty() && tree.max_rank[i] < layer) return;
  if (!overlap(node.box, box) || node.box.near > box.far) return;
  if (node.left < 0) {
    for (int j = node.begin; j < node.end && !w.visible.empty(); ++j) {
      int c = tree.ids[j];
      const Box &cb = (*tree.boxes)[c];
      if (input[c].source == target.source || (ranks && (*ranks)[c] < layer) || !overlap(cb, box) ||
          cb.near > box.far)
        continue;
      ++w.tally.candidates;
      cut(w, input[c].ring, s.sources[target.source], s.sources[input[c].source]);
    }
  } else {
    int a = node.left, b = node.right;
    if (tree.nodes[a].box.near > tree.nodes[b].box.near) std::swap(a, b);
    query(s, input, tree, target, box, w, ranks, layer, a);
    query(s, input, tree, target, box, w, ranks, layer, b);
  }
}

// Classification, then one clipping pass per layer (peel_common.h:135-170, 344-460).
Extraction extract(const Scene &s, int layers, int threads, bool keep_faces) {
  Extraction ex;
  const int n = (int)s.sources.size();
  std::vector<Patch> originals(n);
  for (int i = 0; i < n; ++i) originals[i] = Patch{i, s.sources[i].ring};
  std::vector<Box> original_boxes = patch_boxes(s, originals);
  Hierarchy original_tree(original_boxes);
  // ---- classification: one ray after the source is clear of contours ----
  std::vector<int> ranks(n, -1);
  long ray_candidates = 0;
  {
    std::vector<Box> edges;
    edges.reserve(s.contours.size());
    for (const auto &e : s.contours) {
      Box b;
      b.add(e.first), b.add(e.second);
      for (int d = 0; d < 2; ++d) b.lo[d] -= 4, b.hi[d] += 4;
      edges.push_back(b);
    }
    Hierarchy contour_tree(edges);
#pragma omp parallel for schedule(dynamic, 64) num_threads(threads) reduction(+ : ray_candidates)
    for (int i = 0; i < n; ++i) {
      const Source &src = s.sources[i];
      bool boundary = false;
      contour_tree.visit(original_boxes[i], [&](int j) {
        if (!boundary) boundary = touches_contour(src.ring, s.contours[j].first, s.contours[j].second);
      });
      if (boundary || src.ring.size() != 3) continue;
      int64_t px = 0, py = 0;
      for (int k = 0; k < 3; ++k) px += src.ring[k].x, py += src.ring[k].y;
      Box point;
      point.add(GP{int32_t(px / 3), int32_t(py / 3)});
      point.add(GP{int32_t((px + 2) / 3), int32_t((py + 2) / 3)});
      int rank = 0;
      bool ambiguous = false;
      original_tree.visit(point, [&](int j) {
        if (j == i) return;
        ++ray_candidates;
        int hit = ray_order(src, s.sources[j], px, py);
        ambiguous |= hit < 0;
        rank += hit > 0;
      });
      if (!ambiguous) ranks[i] = rank;
    }
  }
  ex.ray_candidates = ray_candidates;
  original_tree.summarize(ranks);
  std::vector<Patch> residual;
  std::vector<std::vector<int>> direct;
  for (int i = 0; i < n; ++i) {
    if (ranks[i] < 0) {
      residual.push_back(originals[i]);
    } else {
      if ((int)direct.size() <= ranks[i]) direct.resize(ranks[i] + 1);
      direct[ranks[i]].push_back(i);
      ++ex.constant_rank;
    }
  }
  ex.residual_sources = (long)residual.size();
  const int limit = layers ? layers : n;
  for (int k = 0; k < limit && (!residual.empty() || k < (int)direct.size()); ++k) {
    std::vector<Box> boxes = patch_boxes(s, residual);
    Hierarchy tree(boxes);
    std::vector<Work> work(residual.size());
#pragma omp parallel for schedule(dynamic, 32) num_threads(threads)
    for (long i = 0; i < (long)residual.size(); ++i) {
      Work &w = work[i];
      const Patch &target = residual[i];
      w.visible.push_back(target.ring);
      query(s, originals, original_tree, target, boxes[i], w, &ranks, k);
      query(s, residual, tree, target, boxes[i], w, nullptr, 0);
      if (w.visible.empty()) {
        w.hidden.clear();
        w.hidden.push_back(target.ring);
        ++w.tally.occluded;
      }
    }
    Layer layer;
    if (k < (int)direct.size())
      for (int i : direct[k]) layer.patches.push_back(originals[i]);
    std::vector<Patch> next;
    for (size_t i = 0; i < work.size(); ++i) {
      for (Ring &r : work[i].visible) layer.patches.push_back(Patch{residual[i].source, std::move(r)});
      for (Ring &r : work[i].hidden) next.push_back(Patch{residual[i].source, std::move(r)});
      const Tally &t = work[i].tally;
      ex.tally.candidates += t.candidates, ex.tally.cuts += t.cuts, ex.tally.repairs += t.repairs;
      ex.tally.collapsed += t.collapsed, ex.tally.ties += t.ties, ex.tally.occluded += t.occluded;
    }
    if (layer.patches.empty()) {
      ex.error |= 1;
      break;
    }
    mesh(layer, threads, keep_faces);
    ex.layers.push_back(std::move(layer));
    residual = std::move(next);
  }
  if (!layers && !residual.empty()) ex.error |= 2;
  ex.leftover = (long)residual.size();
  return ex;
}

}  // namespace

extern "C" {

// counts[0..18]: layers, patches, faces, error, constant-rank sources,
// residual sources, candidate tests, polygon differences, convexity repairs,
// collapsed polygons, depth ties, active sources, contour segments, boundary,
// silhouette, non-manifold, ignored internal and clipping edges, ray
// candidates of the classification.
// seconds[0]: setup (projection + topology); seconds[1 + r]: extraction r.
// Patch rings are written (grid coordinates / 2^bits, i.e. the normalised
// window) when poly_xy is not null: up to max_vertices per ring, poly_n the
// full count.
int oracle_fast_peel(const double *V, int nv, const int *F, int nf, const double *M, const double *S, int layers,
                     int threads, int repeat, int bits, double *poly_xy, int *poly_n, int *poly_src, int *poly_layer,
                     int cap_poly, int max_vertices, double *tri_uv, int *tri_src, int *tri_layer, int cap_tri,
                     long *counts, double *seconds) {
  (void)nv;
  for (int k = 0; k < 19; ++k) counts[k] = 0;
  if (bits < 12 || bits > 30 || repeat < 1) return 1;
  if (threads < 1) threads = 1;
  auto t0 = Clock::now();
  Scene s = project(V, F, nf, M, S, bits);
  seconds[0] = since(t0);
  const bool keep = tri_uv != nullptr;
  Extraction ex;
  for (int r = 0; r < repeat; ++r) {
    auto t = Clock::now();
    ex = extract(s, layers, threads, keep);
    seconds[1 + r] = since(t);
  }
  const double g = double(int64_t(1) << bits);
  long patches = 0, faces = 0;
  for (size_t k = 0; k < ex.layers.size(); ++k) {
    const Layer &l = ex.layers[k];
    for (const Patch &p : l.patches) {
      if (poly_xy && patches < cap_poly) {
        int m = std::min(p.ring.size(), max_vertices);
        for (int j = 0; j < m; ++j) {
          poly_xy[(patches * max_vertices + j) * 2] = p.ring[j].x / g;
          poly_xy[(patches * max_vertices + j) * 2 + 1] = p.ring[j].y / g;
        }
        poly_n[patches] = p.ring.size();
        poly_src[patches] = s.sources[p.source].id;
        poly_layer[patches] = (int)k;
      }
      ++patches;
    }
    if (keep)
      for (size_t f = 0; f < l.faces; ++f, ++faces) {
        if (faces >= cap_tri) continue;
        for (int c = 0; c < 6; ++c) tri_uv[faces * 6 + c] = l.fan_xy[f * 6 + c] / g;
        tri_src[faces] = s.sources[l.fan_src[f]].id;
        tri_layer[faces] = (int)k;
      }
    else
      faces += (long)l.faces;
  }
  counts[0] = (long)ex.layers.size();
  counts[1] = patches;
  counts[2] = faces;
  counts[3] = ex.error | (patches > cap_poly && poly_xy ? 4 : 0) | (keep && faces > cap_tri ? 8 : 0);
  counts[4] = ex.constant_rank;
  counts[5] = ex.residual_sources;
  counts[6] = (long)ex.tally.candidates;
  counts[7] = (long)ex.tally.cuts;
  counts[8] = (long)ex.tally.repairs;
  counts[9] = (long)ex.tally.collapsed;
  counts[10] = (long)ex.tally.ties;
  counts[11] = (long)s.sources.size();
  counts[12] = (long)s.contours.size();
  counts[13] = s.boundary;
  counts[14] = s.silhouette;
  counts[15] = s.nonmanifold;
  counts[16] = s.ignored;
  counts[17] = s.clipping;
  counts[18] = ex.ray_candidates;
  return 0;
}

}  // extern "C"
