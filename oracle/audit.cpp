// ORACLE — test infrastructure only (see expansion.hpp).
//
// Region audit of an extraction result against exact per-point ranks, in the
// manner of the reference's rational ray audit (scripts/audit_gpu.py:22-137):
// at each sample point whose exact rank list is unambiguous, every covering
// source must own exactly one output fragment that contains the point, with
// that fragment's layer equal to the exact rank; no fragment of a source that
// does not cover the point may contain it. Containment in a fragment (a convex
// counter-clockwise polygon with binary64 vertices, normalised coordinates) is
// decided exactly; a point closer than 1e-11 to a fragment's boundary makes
// that fragment's ownership ambiguous, as the reference audit skips such
// points (audit_gpu.py:78-80): fragment vertices are binary64 constructions,
// so a point within rounding distance of a source edge may fall on either
// side of the fragment's rounded edge.
#include "expansion.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

using namespace oracle;

namespace {

// Sign of orient2d(a, b, c) for binary64 points: a floating-point filter with a
// conservative error bound (Shewchuk's orient2d bound, loosened), exact
// expansions otherwise.
int orient_sign(double ax, double ay, double bx, double by, double cx, double cy) {
  double l = (bx - ax) * (cy - ay), r = (by - ay) * (cx - ax);
  double det = l - r;
  double bound = 1e-15 * (std::fabs(l) + std::fabs(r));
  if (det > bound) return 1;
  if (-det > bound) return -1;
  Exp bxax = two(bx, -ax), cyay = two(cy, -ay), byay = two(by, -ay), cxax = two(cx, -ax);
  return sub(mul(bxax, cyay), mul(byay, cxax)).sign();
}

// Distance from p to the segment ab (binary64; only compared with 1e-11).
double segment_distance(double px, double py, double ax, double ay, double bx, double by) {
  double dx = bx - ax, dy = by - ay, len2 = dx * dx + dy * dy;
  double t = len2 > 0 ? ((px - ax) * dx + (py - ay) * dy) / len2 : 0.0;
  t = std::min(1.0, std::max(0.0, t));
  return std::hypot(px - (ax + t * dx), py - (ay + t * dy));
}

constexpr double BOUNDARY_DISTANCE = 1e-11;

}  // namespace

extern "C" {

// fragments: frag_offsets[nfrag + 1] into frag_xy (2 doubles per vertex).
// ranks CSR as produced by oracle_exact_ranks (offsets, sources, ranks, point_flags).
// counters[0..7]: checked points, skipped points, checked (point, source) pairs,
// missing, duplicate, wrong layer, spurious, boundary-skipped pairs.
// examples: up to max_examples rows of (point, source, expected, got, kind).
void oracle_audit_fragments(const long *frag_offsets, const double *frag_xy, const int *frag_sources,
                            const int *frag_layers, int nfrag, const double *points, int nq,
                            const long *offsets, const int *sources, const int *ranks,
                            const int *point_flags, long *counters, int *examples,
                            int max_examples) {
  for (int k = 0; k < 8; ++k) counters[k] = 0;
  int nex = 0;
  // Grid over fragment bounds.
  int g = 1;
  while (g * g < nfrag / 2 && g < 1024) g *= 2;
  std::vector<std::vector<int>> grid(g * g);
  for (int f = 0; f < nfrag; ++f) {
    double lo[2] = {1e300, 1e300}, hi[2] = {-1e300, -1e300};
    for (long v = frag_offsets[f]; v < frag_offsets[f + 1]; ++v)
      for (int k = 0; k < 2; ++k) {
        lo[k] = std::min(lo[k], frag_xy[2 * v + k]);
        hi[k] = std::max(hi[k], frag_xy[2 * v + k]);
      }
    if (frag_offsets[f + 1] - frag_offsets[f] < 3) continue;
    int x0 = std::max(0, std::min(g - 1, (int)std::floor(lo[0] * g)));
    int x1 = std::max(0, std::min(g - 1, (int)std::floor(hi[0] * g)));
    int y0 = std::max(0, std::min(g - 1, (int)std::floor(lo[1] * g)));
    int y1 = std::max(0, std::min(g - 1, (int)std::floor(hi[1] * g)));
    for (int y = y0; y <= y1; ++y)
      for (int x = x0; x <= x1; ++x) grid[y * g + x].push_back(f);
  }
  auto example = [&](int i, int s, int want, int got, int kind) {
    if (nex < max_examples) {
      int *e = examples + 5 * nex++;
      e[0] = i, e[1] = s, e[2] = want, e[3] = got, e[4] = kind;
    }
  };
  std::vector<int> inside, boundary;
  for (int i = 0; i < nq; ++i) {
    double qx = points[2 * i], qy = points[2 * i + 1];
    if (point_flags[i] != 0 || !(qx > 0 && qx < 1 && qy > 0 && qy < 1)) {
      counters[1]++;
      continue;
    }
    int gx = std::min(g - 1, (int)(qx * g)), gy = std::min(g - 1, (int)(qy * g));
    inside.clear();
    boundary.clear();
    for (int f : grid[gy * g + gx]) {
      int state = 1;
      long b = frag_offsets[f], n = frag_offsets[f + 1] - b;
      for (long v = 0; v < n && state == 1; ++v) {
        const double *p = frag_xy + 2 * (b + v), *q = frag_xy + 2 * (b + (v + 1) % n);
        if (segment_distance(qx, qy, p[0], p[1], q[0], q[1]) < BOUNDARY_DISTANCE) state = 0;
      }
      for (long v = 0; v < n && state == 1; ++v) {
        const double *p = frag_xy + 2 * (b + v), *q = frag_xy + 2 * (b + (v + 1) % n);
        int s = orient_sign(p[0], p[1], q[0], q[1], qx, qy);
        if (s < 0) state = -1;
        else if (s == 0) state = 0;
      }
      if (state == 1) inside.push_back(f);
      if (state == 0) boundary.push_back(f);
    }
    // A fragment boundary through the point makes ownership ambiguous for that
    // source only; other sources are still checked.
    counters[0]++;
    for (long k = offsets[i]; k < offsets[i + 1]; ++k) {
      int s = sources[k];
      bool on_edge = false;
      for (int f : boundary) on_edge |= frag_sources[f] == s;
      int count = 0, layer = -1;
      for (int f : inside)
        if (frag_sources[f] == s) {
          ++count;
          layer = frag_layers[f];
        }
      if (on_edge && count == 0) {
        counters[7]++;
        continue;
      }
      counters[2]++;
      if (count == 0) {
        counters[3]++;
        example(i, s, ranks[k], -1, 0);
      } else if (count > 1) {
        counters[4]++;
        example(i, s, ranks[k], count, 1);
      } else if (layer != ranks[k]) {
        counters[5]++;
        example(i, s, ranks[k], layer, 2);
      }
    }
    for (int f : inside) {
      bool covering = false;
      for (long k = offsets[i]; k < offsets[i + 1]; ++k) covering |= sources[k] == frag_sources[f];
      if (!covering) {
        counters[6]++;
        example(i, frag_sources[f], -1, frag_layers[f], 3);
      }
    }
  }
}

}  // extern "C"
