// ORACLE — test infrastructure only (see expansion.hpp). Built into
// oracle/liboracle.so; used by tests/ as the partition reference of the
// layered route. Never linked by the product.
//
// CPU FP64 restatement of the reference's DevicePeeler route, sequential,
// following warp_peel/kernels.py:539-938 and solver.py:198-380:
//  contour_bounds, touches, classify   kernels.py:539-683
//  separated, split_at, split_edge     kernels.py:179-232
//  cut                                 kernels.py:735-812
//  peel (one pass per layer), advance  kernels.py:815-938
// in the device's binary64 operation order (-ffp-contract=off), with the
// device's candidate order: for a target piece, the leaves of the static
// source hierarchy in depth-first order, and at each leaf the whole source
// (if not yet peeled) followed by that source's live residual pieces in slot
// order. Residual slots are ordered by (source, parent slot, piece), so the
// whole extraction is a function of the input. `order_seed` != 0 instead
// shuffles each pass's residual slots (the order a device would produce with
// unordered appends) to study order sensitivity.
#include <cstdio>
#include <random>

#include "cpu_prepare.hpp"

namespace {

constexpr int L_POLY = 1, L_PIECES = 2, L_RESIDUAL = 4, L_OUTPUT = 8, L_LAYERS = 64;

inline V2 sub(V2 a, V2 b) { return {a.x - b.x, a.y - b.y}; }

Box ring_box(const Ring &p, const double *s) {
  double lx = 1e30, ly = 1e30, lz = 1e30, hx = -1e30, hy = -1e30, hz = -1e30;
  for (int j = 0; j < p.n; ++j) {
    V2 x = p.v[j];
    double z = value(s, x);
    lx = std::fmin(lx, x.x), ly = std::fmin(ly, x.y), lz = std::fmin(lz, z);
    hx = std::fmax(hx, x.x), hy = std::fmax(hy, x.y), hz = std::fmax(hz, z);
  }
  return pad_box(lx, ly, lz, hx, hy, hz);
}

bool overlaps(const Box &b, const float *lo, const float *hi) {
  return b.lo[0] <= hi[0] && b.lo[1] <= hi[1] && b.lo[2] <= hi[2] && b.hi[0] >= lo[0] && b.hi[1] >= lo[1] &&
         b.hi[2] >= lo[2];
}

// touches (kernels.py:560-574)
bool touches(const Ring &p, V2 a, V2 b) {
  bool hit = true;
  for (int j = 0; j < p.n; ++j) {
    V2 u = p.v[j], v = p.v[(j + 1) % p.n];
    if (cross(u, v, a) < -EPS && cross(u, v, b) < -EPS) hit = false;
  }
  bool positive = false, negative = false;
  for (int j = 0; j < p.n; ++j) {
    double s = cross(a, b, p.v[j]);
    positive = positive || s >= -EPS;
    negative = negative || s <= EPS;
  }
  return hit && positive && negative;
}

// separated (kernels.py:221-232)
bool separated(const Ring &a, const Ring &b) {
  for (int j = 0; j < a.n; ++j) {
    V2 u = a.v[j], v = a.v[(j + 1) % a.n];
    bool positive = false;
    for (int k = 0; k < b.n && !positive; ++k) positive = cross(u, v, b.v[k]) > 0.0;
    if (!positive) return true;
  }
  return false;
}

// Strict convexity of a constructed ring: a vertex whose turn is not left by
// more than TURN_TOL times the L1 lengths of its two chords is dropped, until
// every turn is. The bound exceeds the rounding error of the turn, so every
// kept turn is left in exact arithmetic and each edge's supporting line has
// the whole ring on its left; a dropped vertex lies within about TURN_TOL of
// the chord of its neighbours. (Clipping by the edge lines of a ring with a
// slightly reflex vertex, from two constructed points ~1e-12 apart, clips by
// a rotated line and misplaces a strip up to ~1e-9 wide.) Not in the
// reference; the device's layered route applies the same rule.
constexpr double TURN_TOL = 1.0e-12;
bool g_convexify = true;  // off only to show what the rule prevents
void convexify(Ring &p) {
  if (p.n > MAXV || !g_convexify) return;
  bool changed = true;
  while (changed && p.n >= 3) {
    changed = false;
    for (int j = 0; j < p.n && p.n >= 3;) {
      V2 a = p.v[(j + p.n - 1) % p.n], b = p.v[j], c = p.v[(j + 1) % p.n];
      double ux = b.x - a.x, uy = b.y - a.y, wx = c.x - a.x, wy = c.y - a.y;
      double turn = ux * wy - uy * wx;
      double scale = std::fabs(ux) + std::fabs(uy) + std::fabs(wx) + std::fabs(wy);
      if (turn <= TURN_TOL * scale) {
        for (int k = j; k + 1 < p.n; ++k) p.v[k] = p.v[k + 1];
        p.n -= 1;
        changed = true;
      } else {
        ++j;
      }
    }
  }
  if (p.n < 3) p.n = 0;
}

// split_at (kernels.py:179-203), outputs made strictly convex
void split_at(const Ring &p, double sx, double sy, double sz, V2 anchor, Ring &in, Ring &out) {
  in = Ring();
  out = Ring();
  for (int j = 0; j < p.n; ++j) {
    V2 a = p.v[j], b = p.v[(j + 1) % p.n];
    double da = sx * (a.x - anchor.x) + sy * (a.y - anchor.y) + sz;
    double db = sx * (b.x - anchor.x) + sy * (b.y - anchor.y) + sz;
    if (da >= 0.0) in.append(a);
    if (da <= 0.0) out.append(a);
    if ((da > 0.0 && db < 0.0) || (da < 0.0 && db > 0.0)) {
      double t = da / (da - db);
      V2 x{a.x + t * (b.x - a.x), a.y + t * (b.y - a.y)};
      in.append(x);
      out.append(x);
    }
  }
  clean(in);
  clean(out);
  convexify(in);
  convexify(out);
}

struct Cutter {
  int cap;  // pieces per worker pool
  int error = 0;
  long ties = 0;
  // cut (kernels.py:735-812): subtract `cand`, nearer where delta < 0 on the
  // target, from the visible pieces; returns the new visible count or -1.
  int cut(std::vector<Ring> &vis, std::vector<Ring> &hidden, const Ring &cand, const double *delta, int target,
          int other) {
    std::vector<Ring> tmp;
    int count = 0;
    auto put = [&](const Ring &r) {
      if (count < cap) tmp.push_back(r);
      count += 1;
    };
    for (size_t v = 0; v < vis.size(); ++v) {
      const Ring &p = vis[v];
      if (separated(p, cand) || separated(cand, p)) {
        put(p);
      } else {
        bool positive = false, negative = false;
        for (int e = 0; e < p.n; ++e) {
          double d = delta[0] * p.v[e].x + delta[1] * p.v[e].y + delta[2];
          positive = positive || d > 0.0;
          negative = negative || d < 0.0;
        }
        if (!positive && !negative) {
          positive = target > other;
          ties += 1;
        }
        if (!positive) {
          put(p);
        } else {
          Ring inside = p, next, out;
          int base = count;
          if (negative) {
            split_at(inside, delta[0], delta[1], delta[2], V2{0.0, 0.0}, next, out);
            if (next.n > MAXV || out.n > MAXV) return error |= L_POLY, -1;
            inside = next;
            if (out.n > 0) put(out);
          }
          for (int e = 0; e < cand.n && inside.n > 0; ++e) {
            V2 a = cand.v[e], b = cand.v[(e + 1) % cand.n];
            V2 d = sub(b, a);
            split_at(inside, -d.y, d.x, 0.0, a, next, out);
            if (next.n > MAXV || out.n > MAXV) return error |= L_POLY, -1;
            inside = next;
            if (out.n > 0) put(out);
          }
          if (inside.n == 0) {
            // a collapsed inside restores the piece unchanged
            tmp.resize(std::min(base, cap));
            count = base;
            put(p);
          } else {
            hidden.push_back(inside);
          }
        }
      }
      if (count > cap || (int)hidden.size() > cap) return error |= L_PIECES, -1;
    }
    vis.swap(tmp);
    return count;
  }
};

struct Piece {
  Ring ring;
  int source;
  int parent;  // slot of the target it was hidden under, -1 at classification
  int index;   // position among that target's hidden pieces
  Box box;
};

}  // namespace

extern "C" {

void oracle_cpu_layered_options(int convexify_rings) { g_convexify = convexify_rings != 0; }

// Convexify one ring of n <= 32 vertices (xy interleaved) in place; returns
// the new count.
int oracle_convexify(double *xy, int n) {
  Ring r;
  for (int j = 0; j < n && j < MAXV; ++j) r.v[j] = V2{xy[2 * j], xy[2 * j + 1]};
  r.n = n;
  bool saved = g_convexify;
  g_convexify = true;
  convexify(r);
  g_convexify = saved;
  for (int j = 0; j < r.n; ++j) xy[2 * j] = r.v[j].x, xy[2 * j + 1] = r.v[j].y;
  return r.n;
}

// counts[0..9]: polygons, layers, error bits, depth ties, passes, peak
// residual pieces, max hidden pieces of one target, whole sources emitted by
// classification, residual sources, candidate tests.
// stats (optional, [stats_rows][3]): per layer targets, candidate tests,
// fully hidden targets.
int oracle_cpu_layered_peel(const double *V, int nv, const int *F, int nf, const double *M, const double *S,
                            int layers_limit, int layer_capacity, int pieces_cap, int residual_cap,
                            unsigned order_seed, double *poly_xy, int *poly_n, int *poly_src, int *poly_layer,
                            int cap_poly, long *counts, long *stats, int stats_rows) {
  for (int k = 0; k < 10; ++k) counts[k] = 0;
  Prep prep;
  if (int bits = prepare(V, nv, F, nf, M, S, prep)) {
    counts[2] = bits;
    return 0;
  }
  const std::vector<Ring> &src = prep.src;
  const double *score = prep.score.data();
  const int limit = layers_limit ? layers_limit : layer_capacity;
  const bool all_layers = layers_limit == 0;
  int error = 0, layers = 0;
  long emitted = 0;
  auto save = [&](const Ring &p, int source, int layer) {
    long slot = emitted++;
    if (slot < cap_poly) {
      for (int j = 0; j < p.n && j < MAXV; ++j) {
        poly_xy[(slot * MAXV + j) * 2] = p.v[j].x;
        poly_xy[(slot * MAXV + j) * 2 + 1] = p.v[j].y;
      }
      poly_n[slot] = p.n;
      poly_src[slot] = source;
      poly_layer[slot] = layer;
      layers = std::max(layers, layer + 1);
    } else {
      error |= L_OUTPUT;
    }
  };
  // ---- contour_bounds + contour hierarchy (kernels.py:539-557) -----------
  std::vector<Box> cbox(nf);
  for (int i = 0; i < nf; ++i) {
    const Ring &p = src[i];
    double lx = 1e30, ly = 1e30, lz = 1e30, hx = -1e30, hy = -1e30, hz = -1e30;
    for (int j = 0; j < p.n; ++j) {
      if (!prep.contour_edge(i, j)) continue;
      V2 a = p.v[j], b = p.v[(j + 1) % p.n];
      lx = std::fmin(lx, std::fmin(a.x, b.x)), ly = std::fmin(ly, std::fmin(a.y, b.y)), lz = std::fmin(lz, 0.0);
      hx = std::fmax(hx, std::fmax(a.x, b.x)), hy = std::fmax(hy, std::fmax(a.y, b.y)), hz = std::fmax(hz, 0.0);
    }
    Box &b = cbox[i];
    b.lo[0] = (float)(lx - 2.0e-7), b.lo[1] = (float)(ly - 2.0e-7), b.lo[2] = (float)(lz - 2.0e-7);
    b.hi[0] = (float)(hx + 2.0e-7), b.hi[1] = (float)(hy + 2.0e-7), b.hi[2] = (float)(hz + 2.0e-7);
  }
  Tree contour_tree, source_tree;
  contour_tree.build(cbox, false);
  source_tree.build(prep.src_box, false);
  // ---- classify (kernels.py:597-683) --------------------------------------
  std::vector<int> ranks(nf, -1);
  std::vector<Piece> cur;
  for (int i = 0; i < nf; ++i) {
    const Ring &p = src[i];
    int rank = -1;
    if (p.n > 0) {
      double lx = 1e30, ly = 1e30, hx = -1e30, hy = -1e30;
      for (int j = 0; j < p.n; ++j)
        lx = std::fmin(lx, p.v[j].x), ly = std::fmin(ly, p.v[j].y), hx = std::fmax(hx, p.v[j].x),
        hy = std::fmax(hy, p.v[j].y);
      float lo[3] = {(float)(lx - 2.0e-7), (float)(ly - 2.0e-7), -1.0e-6f};
      float hi[3] = {(float)(hx + 2.0e-7), (float)(hy + 2.0e-7), 1.0e-6f};
      int node = 1;
      bool boundary = false;
      while (node > 0 && !boundary) {
        int leaf = contour_tree.next_leaf(lo, hi, node);
        if (leaf < 0) break;
        int j = contour_tree.ids[leaf];
        const Ring &q = src[j];
        for (int e = 0; e < q.n && !boundary; ++e)
          boundary = prep.contour_edge(j, e) && touches(p, q.v[e], q.v[(e + 1) % q.n]);
      }
      if (!boundary) {
        V2 sample{0.0, 0.0};
        for (int e = 0; e < p.n; ++e) {
          sample.x += p.v[e].x / (double)p.n;
          sample.y += p.v[e].y / (double)p.n;
        }
        double z = value(score + 4 * i, sample);
        float qlo[3] = {(float)sample.x - 2.0e-7f, (float)sample.y - 2.0e-7f, -1.0e30f};
        float qhi[3] = {(float)sample.x + 2.0e-7f, (float)sample.y + 2.0e-7f,
                        (float)(z + 2.0e-7 * std::fmax(1.0, std::fabs(z)))};
        int node2 = 1;
        rank = 0;
        bool ambiguous = false;
        while (node2 > 0) {
          int leaf = source_tree.next_leaf(qlo, qhi, node2);
          if (leaf < 0) break;
          int j = source_tree.ids[leaf];
          if (j == i) continue;
          const Ring &q = src[j];
          bool inside = true, edge = false;
          for (int e = 0; e < q.n; ++e) {
            double s = cross(q.v[e], q.v[(e + 1) % q.n], sample);
            inside = inside && s >= 0.0;
            edge = edge || std::fabs(s) <= EPS;
          }
          if (inside) {
            double delta = z - value(score + 4 * j, sample);
            if (edge || std::fabs(delta) <= 1.0e-12 * std::fmax(1.0, std::fabs(z)))
              ambiguous = true;
            else if (delta > 0.0)
              rank += 1;
          }
        }
        if (ambiguous) rank = -1;
      }
    }
    if (rank < 0) {
      Ring q = p;
      convexify(q);
      if ((int)cur.size() < residual_cap)
        cur.push_back({q, i, -1, 0, ring_box(q, score + 4 * i)});
      else
        error |= L_RESIDUAL;
      counts[8] += 1;
    } else if (rank < limit) {
      save(p, i, rank);
      counts[7] += 1;
    } else if (all_layers) {
      error |= L_LAYERS;
    }
    ranks[i] = rank;
  }
  // ---- passes (kernels.py:815-938) ----------------------------------------
  std::mt19937 rng(order_seed);
  Cutter cutter{pieces_cap};
  long peak = (long)cur.size(), max_hidden = 0, tests_all = 0;
  int pass = 0;
  std::vector<int> first(nf, 0), number(nf, 0);
  while (!cur.empty() && error == 0 && pass < limit) {
    if (order_seed) std::shuffle(cur.begin(), cur.end(), rng);
    // residual pieces of each source: one contiguous run in slot order
    std::fill(number.begin(), number.end(), 0);
    if (!order_seed) {
      for (int k = (int)cur.size() - 1; k >= 0; --k) first[cur[k].source] = k, number[cur[k].source] += 1;
    }
    std::vector<std::vector<int>> runs;
    if (order_seed) {
      runs.assign(nf, {});
      for (int k = 0; k < (int)cur.size(); ++k) runs[cur[k].source].push_back(k);
    }
    std::vector<Piece> next;
    long targets = 0, tested_all = 0, hidden_all = 0;
    for (int t = 0; t < (int)cur.size() && cutter.error == 0; ++t) {
      const Piece &target = cur[t];
      const int source = target.source;
      const double *sc = score + 4 * source;
      std::vector<Ring> vis{target.ring}, hidden;
      int nv_ = 1;
      float lo[3] = {target.box.lo[0], target.box.lo[1], -1.0e30f};
      const float *hi = target.box.hi;
      long tested = 0;
      int node = 1;
      auto try_cut = [&](const Ring &cand, int other) {
        double delta[3] = {sc[0] - score[4 * other], sc[1] - score[4 * other + 1], sc[2] - score[4 * other + 2]};
        tested += 1;
        nv_ = cutter.cut(vis, hidden, cand, delta, source, other);
      };
      while (node > 0 && nv_ > 0) {
        int leaf = source_tree.next_leaf(lo, hi, node);
        if (leaf < 0) break;
        int c = source_tree.ids[leaf];
        if (c == source) continue;
        if (ranks[c] >= pass) try_cut(src[c], c);
        if (order_seed) {
          for (size_t r = 0; r < runs[c].size() && nv_ > 0; ++r) {
            const Piece &k = cur[runs[c][r]];
            if (overlaps(k.box, lo, hi)) try_cut(k.ring, c);
          }
        } else {
          for (int k = first[c]; k < first[c] + number[c] && nv_ > 0; ++k)
            if (overlaps(cur[k].box, lo, hi)) try_cut(cur[k].ring, c);
        }
      }
      targets += 1;
      tested_all += tested;
      max_hidden = std::max<long>(max_hidden, (long)hidden.size());
      if (nv_ == 0) {
        hidden.assign(1, target.ring);
        hidden_all += 1;
      }
      if (nv_ < 0) break;
      for (const Ring &r : vis) save(r, source, pass);
      for (int j = 0; j < (int)hidden.size(); ++j) {
        if ((long)next.size() >= residual_cap) {
          error |= L_RESIDUAL;
          break;
        }
        next.push_back({hidden[j], source, t, j, ring_box(hidden[j], sc)});
      }
    }
    error |= cutter.error;
    if (stats && pass < stats_rows) {
      stats[3 * pass] = targets;
      stats[3 * pass + 1] = tested_all;
      stats[3 * pass + 2] = hidden_all;
    }
    tests_all += tested_all;
    // next residual in (source, parent slot, piece) order
    std::stable_sort(next.begin(), next.end(), [](const Piece &a, const Piece &b) {
      if (a.source != b.source) return a.source < b.source;
      if (a.parent != b.parent) return a.parent < b.parent;
      return a.index < b.index;
    });
    cur.swap(next);
    peak = std::max<long>(peak, (long)cur.size());
    pass += 1;
    if (!cur.empty() && pass >= limit && all_layers) error |= L_LAYERS;
  }
  counts[0] = emitted;
  counts[1] = layers;
  counts[2] = error;
  counts[3] = cutter.ties;
  counts[4] = pass;
  counts[5] = peak;
  counts[6] = max_hidden;
  counts[9] = tests_all;
  return 0;
}

}  // extern "C"
