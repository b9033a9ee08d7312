// Layered route: the DevicePeeler's whole-face ranks and per-layer clipping.
//
// Reference semantics (file:line in alecjacobson/exact-depth-peeling):
//  contour_bounds      warp_peel/kernels.py:539-557
//  touches             warp_peel/kernels.py:560-574
//  save_polygon        warp_peel/kernels.py:577-594
//  classify            warp_peel/kernels.py:597-683
//  start_loop          warp_peel/kernels.py:686-689
//  separated/split     warp_peel/kernels.py:184-232
//  cut                 warp_peel/kernels.py:735-812
//  peel                warp_peel/kernels.py:815-908
//  advance             warp_peel/kernels.py:924-938
//  orchestration       warp_peel/solver.py:198-380 (_prepare, _layer)
//
// A source whose own or a neighbour's contour edge may touch it, or whose
// centroid ray is ambiguous (edge or depth tie), goes to the residual pool;
// every other source gets its rank from one centroid ray and is emitted whole.
// Each pass then clips every residual piece against the sources and pieces
// that may cover it: the visible parts are the current layer, the hidden
// parts (or the whole piece, if nothing stays visible) the next residual.
//
// MI355X layout instead of the reference's per-layer hierarchies: one static
// source hierarchy answers every candidate query. After classification and
// after every pass the residual pieces are ordered by (source, parent slot)
// with a stable radix sort, so each source's live pieces are one run of
// consecutive slots and a query that reaches a source's leaf walks that run;
// the order is a function of the input, so every run cuts in the same order.
// The per-worker piece pools are vertex-major with the worker index fastest,
// so the lanes of a wave touch adjacent addresses when they walk their pieces
// in step.
//
// Every ring this route constructs is made strictly convex (convexify_ring):
// `separated` and the edge-by-edge clipping of `cut` rely on each edge line
// having the whole ring on its left, which a slightly reflex constructed
// vertex breaks.
#include <algorithm>

#include "workspace.hpp"

namespace peel {

// A ring in strided storage: vertex j at base[j * stride].
struct RingRef {
  double2 *base;
  long stride;
  __device__ __forceinline__ double2 operator[](int j) const { return base[(long)j * stride]; }
  __device__ __forceinline__ double2 &at(int j) const { return base[(long)j * stride]; }
};

__device__ __forceinline__ RingRef source_ring(const Ws &w, int i) {
  return RingRef{w.src_xy + i, (long)w.nf};
}
__device__ __forceinline__ RingRef residual_ring(const Ws &w, int par, int k) {
  return RingRef{w.res_xy[par] + k, (long)w.cap.frontier};
}

// Sort key of an appended residual piece: its source, then the slot of the
// target it was hidden under (pieces of one target are appended in order).
__device__ __forceinline__ uint64_t residual_key(const Ws &w, int source, int parent) {
  return ((uint64_t)(uint32_t)source << w.res_parent_bits) | (uint64_t)(uint32_t)parent;
}

// Strict convexity of a constructed ring (not in the reference): a vertex
// whose turn is not left by more than TURN_TOL times the L1 lengths of its two
// chords is dropped, until every turn is. The bound exceeds the rounding error
// of the turn, so every kept turn is left in exact arithmetic; a dropped
// vertex lies within about TURN_TOL of its neighbours' chord. Two constructed
// points ~1e-12 apart can otherwise form a slightly reflex vertex whose short
// edge has a rotated supporting line, and clipping by that line misplaces a
// strip up to ~1e-9 wide on the wrong layer.
constexpr double TURN_TOL = 1.0e-12;
__device__ int convexify_ring(RingRef p, int n) {
  if (n > MAXV) return n;
  bool changed = true;
  while (changed && n >= 3) {
    changed = false;
    for (int j = 0; j < n && n >= 3;) {
      double2 a = p[(j + n - 1) % n], b = p[j], c = p[(j + 1) % n];
      double ux = b.x - a.x, uy = b.y - a.y, wx = c.x - a.x, wy = c.y - a.y;
      double turn = ux * wy - uy * wx;
      double scale = fabs(ux) + fabs(uy) + fabs(wx) + fabs(wy);
      if (turn <= TURN_TOL * scale) {
        for (int k = j; k + 1 < n; ++k) p.at(k) = p[k + 1];
        n -= 1;
        changed = true;
      } else {
        ++j;
      }
    }
  }
  return n < 3 ? 0 : n;
}

// Worker pools: q = 0, 1 alternate as visible/temporary, q = 2 holds hidden
// pieces; slots pieces, pieces + 1 of q = 2 are the inside work rings and
// slot pieces + 2 receives outside rings past the capacity.
__device__ __forceinline__ long pool_slot(const Ws &w, int q, int slot) {
  return (long)q * (w.pieces + 3) + slot;
}
__device__ __forceinline__ RingRef pool_ring(const Ws &w, int q, int slot, int worker) {
  return RingRef{w.pool_xy + pool_slot(w, q, slot) * MAXV * w.workers + worker, (long)w.workers};
}
__device__ __forceinline__ int &pool_n(const Ws &w, int q, int slot, int worker) {
  return w.pool_n[pool_slot(w, q, slot) * w.workers + worker];
}

__device__ __forceinline__ void copy_ring(RingRef from, int n, RingRef to) {
  for (int j = 0; j < n && j < MAXV; ++j) to.at(j) = from[j];
}

__device__ __forceinline__ Box ring_box(RingRef p, int n, double4 score) {
  double lx = 1e30, ly = 1e30, lz = 1e30, hx = -1e30, hy = -1e30, hz = -1e30;
  for (int j = 0; j < n; ++j) {
    double2 x = p[j];
    double z = value(score, x);
    lx = fmin(lx, x.x), ly = fmin(ly, x.y), lz = fmin(lz, z);
    hx = fmax(hx, x.x), hy = fmax(hy, x.y), hz = fmax(hz, z);
  }
  Box b;
  pad_box(lx, ly, lz, hx, hy, hz, b.lo, b.hi);
  return b;
}

// contour_edge (kernels.py:523-536): ring edge j of source i is a contour
__device__ __forceinline__ bool contour_edge(const Ws &w, int i, int j) {
  if (w.clipped[i] != 0) return true;
  int e = w.facing[i] < 0 ? (4 - j) % 3 : j;
  return w.masks[3 * i + e] != 0;
}

// ---------------------------------------------------------------------------
__global__ void k_layered_reset(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  for (int j = i; j < w.nf; j += gridDim.x * blockDim.x) {
    w.run_first[j] = -1;
    w.run_last[j] = -1;
  }
  for (int j = i; j < 3 * w.cap.layers; j += gridDim.x * blockDim.x) w.stats[j] = 0;
  if (i == 0) w.control[C_LOOP] = 0;
}

// contour_bounds (kernels.py:539-557): box of a source's contour edges, z = 0
__global__ void k_contour_bounds(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.nf) return;
  RingRef p = source_ring(w, i);
  int n = w.src_n[i];
  double lx = 1e30, ly = 1e30, lz = 1e30, hx = -1e30, hy = -1e30, hz = -1e30;
  for (int j = 0; j < n; ++j) {
    if (!contour_edge(w, i, j)) continue;
    double2 a = p[j], b = p[(j + 1) % n];
    lx = fmin(lx, fmin(a.x, b.x)), ly = fmin(ly, fmin(a.y, b.y)), lz = fmin(lz, 0.0);
    hx = fmax(hx, fmax(a.x, b.x)), hy = fmax(hy, fmax(a.y, b.y)), hz = fmax(hz, 0.0);
  }
  Box b;
  b.lo = make_float4((float)(lx - 2.0e-7), (float)(ly - 2.0e-7), (float)(lz - 2.0e-7), 0.f);
  b.hi = make_float4((float)(hx + 2.0e-7), (float)(hy + 2.0e-7), (float)(hz + 2.0e-7), 0.f);
  w.contour_boxes[i] = b;
}

// touches (kernels.py:560-574): the segment (a, b) may meet the ring
__device__ bool touches(RingRef p, int n, double2 a, double2 b) {
  bool hit = true;
  for (int j = 0; j < n; ++j) {
    double2 u = p[j], v = p[(j + 1) % n];
    if (cross3(u, v, a) < -EPS && cross3(u, v, b) < -EPS) hit = false;
  }
  bool positive = false, negative = false;
  for (int j = 0; j < n; ++j) {
    double s = cross3(a, b, p[j]);
    positive = positive || s >= -EPS;
    negative = negative || s <= EPS;
  }
  return hit && positive && negative;
}

// save_polygon (kernels.py:577-594)
__device__ void save_polygon(const Ws &w, RingRef p, int n, int source, int layer) {
  int slot = atomicAdd(w.status + ST_POLYS, 1);
  if (slot < w.cap.regions) {
    long R = w.cap.regions;
    for (int j = 0; j < n && j < MAXV; ++j) w.out_xy[(long)j * R + slot] = p[j];
    w.out_n[slot] = n;
    w.out_src[slot] = source;
    w.out_layer[slot] = layer;
    atomicMax(w.status + ST_LAYERS, layer + 1);
  } else {
    set_error(w.status, ERR_OUTPUT);
  }
}

// classify (kernels.py:597-683)
__global__ __launch_bounds__(256) void k_classify(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= w.nf) return;
  RingRef p = source_ring(w, i);
  int n = w.src_n[i];
  int rank = -1;
  if (n > 0) {
    double lx = 1e30, ly = 1e30, hx = -1e30, hy = -1e30;
    for (int j = 0; j < n; ++j) {
      double2 x = p[j];
      lx = fmin(lx, x.x), ly = fmin(ly, x.y), hx = fmax(hx, x.x), hy = fmax(hy, x.y);
    }
    float4 lo = make_float4((float)(lx - 2.0e-7), (float)(ly - 2.0e-7), -1.0e-6f, 0.f);
    float4 hi = make_float4((float)(hx + 2.0e-7), (float)(hy + 2.0e-7), 1.0e-6f, 0.f);
    Tree ct = w.contour_tree.tree();
    int node = 1;
    bool boundary = false;
    while (node > 0 && !boundary) {
      int leaf = query_leaf(ct, lo, hi, node);
      if (leaf < 0) break;
      int j = ct.ids[leaf];
      RingRef q = source_ring(w, j);
      int qn = w.src_n[j];
      for (int e = 0; e < qn; ++e) {
        if (contour_edge(w, j, e) && touches(p, n, q[e], q[(e + 1) % qn])) {
          boundary = true;
          break;
        }
      }
    }
    if (!boundary) {
      double2 sample = make_double2(0.0, 0.0);
      for (int e = 0; e < n; ++e) {
        double2 x = p[e];
        sample.x += x.x / (double)n;
        sample.y += x.y / (double)n;
      }
      double z = value(w.scores[i], sample);
      float4 qlo = make_float4((float)sample.x - 2.0e-7f, (float)sample.y - 2.0e-7f, -1.0e30f, 0.f);
      float4 qhi = make_float4((float)sample.x + 2.0e-7f, (float)sample.y + 2.0e-7f,
                               (float)(z + 2.0e-7 * fmax(1.0, fabs(z))), 0.f);
      Tree st = w.source_tree.tree();
      int node2 = 1;
      rank = 0;
      bool ambiguous = false;
      while (node2 > 0) {
        int leaf = query_leaf(st, qlo, qhi, node2);
        if (leaf < 0) break;
        int j = st.ids[leaf];
        if (j == i) continue;
        RingRef q = source_ring(w, j);
        int qn = w.src_n[j];
        bool inside = true, edge = false;
        for (int e = 0; e < qn; ++e) {
          double s = cross3(q[e], q[(e + 1) % qn], sample);
          inside = inside && s >= 0.0;
          edge = edge || fabs(s) <= EPS;
        }
        if (inside) {
          double delta = z - value(w.scores[j], sample);
          if (edge || fabs(delta) <= 1.0e-12 * fmax(1.0, fabs(z)))
            ambiguous = true;
          else if (delta > 0.0)
            rank += 1;
        }
      }
      if (ambiguous) rank = -1;
    }
  }
  if (rank < 0) {
    int slot = atomicAdd(w.status + ST_CUR, 1);
    if (slot < w.cap.frontier) {
      RingRef q = residual_ring(w, 1, slot);
      copy_ring(p, n, q);
      int qn = convexify_ring(q, n);
      w.res_n[1][slot] = qn;
      w.res_src[1][slot] = i;
      w.res_box[1][slot] = ring_box(q, qn, w.scores[i]);
      w.res_keys[slot] = residual_key(w, i, 0);
      w.res_ids[slot] = slot;
    } else {
      set_error(w.status, ERR_RESIDUAL);
    }
  } else if (rank < w.limit) {
    save_polygon(w, p, n, i, rank);
  } else if (w.all_layers) {
    set_error(w.status, ERR_LAYERS);
  }
  w.ranks[i] = rank;
}

// start_loop (kernels.py:686-689)
__global__ void k_start_loop(Ws w) {
  w.status[ST_PEAK] = w.status[ST_CUR];
  int loop = (w.status[ST_CUR] > 0 && w.status[ST_ERR] == 0) ? 1 : 0;
  w.control[C_LOOP] = loop;
  w.control[C_SORTN] = loop ? w.status[ST_CUR] : 0;
}

// ---------------------------------------------------------------------------
// separated (kernels.py:221-232): some edge of a has every vertex of b on or
// right of it
__device__ bool separated(RingRef a, int an, RingRef b, int bn) {
  for (int j = 0; j < an; ++j) {
    double2 u = a[j], v = a[(j + 1) % an];
    bool positive = false;
    for (int k = 0; k < bn && !positive; ++k) positive = cross3(u, v, b[k]) > 0.0;
    if (!positive) return true;
  }
  return false;
}

// split_at (kernels.py:179-203): keep value(s, x - anchor) >= 0 in `in` and
// <= 0 in `out`, both cleaned and made strictly convex; counts past MAXV are
// reported, not stored.
__device__ void split_at(RingRef p, int n, double3 s, double2 anchor, RingRef in, int &in_n, RingRef out,
                         int &out_n) {
  in_n = 0;
  out_n = 0;
  for (int j = 0; j < n; ++j) {
    double2 a = p[j], b = p[(j + 1) % n];
    double da = s.x * (a.x - anchor.x) + s.y * (a.y - anchor.y) + s.z;
    double db = s.x * (b.x - anchor.x) + s.y * (b.y - anchor.y) + s.z;
    if (da >= 0.0) {
      if (in_n < MAXV) in.at(in_n) = a;
      in_n += 1;
    }
    if (da <= 0.0) {
      if (out_n < MAXV) out.at(out_n) = a;
      out_n += 1;
    }
    if ((da > 0.0 && db < 0.0) || (da < 0.0 && db > 0.0)) {
      double t = da / (da - db);
      double2 x = make_double2(a.x + t * (b.x - a.x), a.y + t * (b.y - a.y));
      if (in_n < MAXV) in.at(in_n) = x;
      in_n += 1;
      if (out_n < MAXV) out.at(out_n) = x;
      out_n += 1;
    }
  }
  in_n = convexify_ring(in, clean_slot(in.base, (int)in.stride, in_n));
  out_n = convexify_ring(out, clean_slot(out.base, (int)out.stride, out_n));
}

// cut (kernels.py:735-812): subtract `cand` (nearer where delta < 0 on the
// target) from the visible pieces; returns the new visible count or -1.
__device__ int cut(const Ws &w, int worker, int &vis, int nv, int &nh, RingRef cand, int cn, double3 delta,
                   int target_source, int other_source, long long &ties) {
  const int cap = w.pieces;
  const int tmp = vis ^ 1;
  const RingRef work[2] = {pool_ring(w, 2, cap, worker), pool_ring(w, 2, cap + 1, worker)};
  const RingRef spill = pool_ring(w, 2, cap + 2, worker);
  int count = 0;
  for (int v = 0; v < nv; ++v) {
    RingRef p = pool_ring(w, vis, v, worker);
    int pn = pool_n(w, vis, v, worker);
    if (separated(p, pn, cand, cn) || separated(cand, cn, p, pn)) {
      if (count < cap) {
        copy_ring(p, pn, pool_ring(w, tmp, count, worker));
        pool_n(w, tmp, count, worker) = pn;
      }
      count += 1;
    } else {
      bool positive = false, negative = false;
      for (int e = 0; e < pn; ++e) {
        double2 x = p[e];
        double d = delta.x * x.x + delta.y * x.y + delta.z;
        positive = positive || d > 0.0;
        negative = negative || d < 0.0;
      }
      if (!positive && !negative) {
        positive = target_source > other_source;
        ties += 1;
      }
      if (!positive) {
        if (count < cap) {
          copy_ring(p, pn, pool_ring(w, tmp, count, worker));
          pool_n(w, tmp, count, worker) = pn;
        }
        count += 1;
      } else {
        RingRef inside = p;
        int in_n = pn, which = 0;
        int base = count;
        if (negative) {
          RingRef out = count < cap ? pool_ring(w, tmp, count, worker) : spill;
          int next_n, out_n;
          split_at(inside, in_n, delta, make_double2(0.0, 0.0), work[which], next_n, out, out_n);
          if (next_n > MAXV || out_n > MAXV) {
            set_error(w.status, ERR_POLY);
            return -1;
          }
          inside = work[which];
          in_n = next_n;
          which ^= 1;
          if (out_n > 0) {
            if (count < cap) pool_n(w, tmp, count, worker) = out_n;
            count += 1;
          }
        }
        for (int e = 0; e < cn && in_n > 0; ++e) {
          // split_edge (kernels.py:206-212): the edge's own frame, anchored at a
          double2 a = cand[e], b = cand[(e + 1) % cn];
          double2 d = make_double2(b.x - a.x, b.y - a.y);
          RingRef out = count < cap ? pool_ring(w, tmp, count, worker) : spill;
          int next_n, out_n;
          split_at(inside, in_n, make_double3(-d.y, d.x, 0.0), a, work[which], next_n, out, out_n);
          if (next_n > MAXV || out_n > MAXV) {
            set_error(w.status, ERR_POLY);
            return -1;
          }
          inside = work[which];
          in_n = next_n;
          which ^= 1;
          if (out_n > 0) {
            if (count < cap) pool_n(w, tmp, count, worker) = out_n;
            count += 1;
          }
        }
        if (in_n == 0) {
          // a collapsed inside restores the piece unchanged
          count = base;
          if (count < cap) {
            copy_ring(p, pn, pool_ring(w, tmp, count, worker));
            pool_n(w, tmp, count, worker) = pn;
          }
          count += 1;
        } else {
          if (nh < cap) {
            copy_ring(inside, in_n, pool_ring(w, 2, nh, worker));
            pool_n(w, 2, nh, worker) = in_n;
          }
          nh += 1;
        }
      }
    }
    if (count > cap || nh > cap) {
      set_error(w.status, ERR_PIECES);
      return -1;
    }
  }
  vis = tmp;
  return count;
}

// peel (kernels.py:815-908): one pass over the ordered pieces of pool 0; the
// hidden pieces are appended to pool 1 with their sort keys.
__global__ __launch_bounds__(256) void k_peel(Ws w) {
  if (load_i(w.control + C_LOOP) == 0) return;
  int worker = blockIdx.x * blockDim.x + threadIdx.x;
  const int layer = w.status[ST_PASS];
  const long R = w.cap.frontier;
  long long targets = 0, tested_all = 0, hidden_all = 0, ties = 0;
  int max_hidden = 0;
  const Tree tree = w.source_tree.tree();
  for (int target = worker; worker < w.workers && target < load_i(w.status + ST_CUR) &&
                            load_i(w.status + ST_ERR) == 0;
       target += w.workers) {
    RingRef p = residual_ring(w, 0, target);
    int pn = w.res_n[0][target];
    int source = w.res_src[0][target];
    int vis = 0, nv = 1, nh = 0;
    copy_ring(p, pn, pool_ring(w, 0, 0, worker));
    pool_n(w, 0, 0, worker) = pn;
    double4 score = w.scores[source];
    Box tb = w.res_box[0][target];
    float4 lo = tb.lo, hi = tb.hi;
    lo.z = -1.0e30f;
    long long tested = 0;
    int node = 1;
    while (node > 0 && nv > 0) {
      int leaf = query_leaf(tree, lo, hi, node);
      if (leaf < 0) break;
      int c = tree.ids[leaf];
      if (c == source) continue;
      double4 sc = w.scores[c];
      double3 delta = make_double3(score.x - sc.x, score.y - sc.y, score.z - sc.z);
      // a whole source not yet peeled (retire_bounds, kernels.py:717-732)
      if (w.ranks[c] >= layer) {
        tested += 1;
        nv = cut(w, worker, vis, nv, nh, source_ring(w, c), w.src_n[c], delta, source, c, ties);
      }
      // this source's live residual pieces (active_bounds, kernels.py:700-714)
      int first = w.run_first[c];
      if (first < 0) continue;
      int last = w.run_last[c];
      for (int k = first; k <= last && nv > 0; ++k) {
        if (!overlaps(w.res_box[0][k], lo, hi)) continue;
        tested += 1;
        nv = cut(w, worker, vis, nv, nh, residual_ring(w, 0, k), w.res_n[0][k], delta, source, c, ties);
      }
    }
    targets += 1;
    tested_all += tested;
    max_hidden = max(max_hidden, nh);
    if (nv == 0) {
      copy_ring(p, pn, pool_ring(w, 2, 0, worker));
      pool_n(w, 2, 0, worker) = pn;
      nh = 1;
      hidden_all += 1;
    }
    if (nv >= 0 && load_i(w.status + ST_ERR) == 0) {
      for (int j = 0; j < nv; ++j)
        save_polygon(w, pool_ring(w, vis, j, worker), pool_n(w, vis, j, worker), source, layer);
      int offset = atomicAdd(w.status + ST_NEXT, nh);
      if ((long)offset + nh > R) {
        set_error(w.status, ERR_RESIDUAL);
      } else {
        for (int j = 0; j < nh; ++j) {
          RingRef h = pool_ring(w, 2, j, worker);
          int hn = pool_n(w, 2, j, worker);
          int slot = offset + j;
          copy_ring(h, hn, residual_ring(w, 1, slot));
          w.res_n[1][slot] = hn;
          w.res_src[1][slot] = source;
          w.res_box[1][slot] = ring_box(h, hn, score);
          w.res_keys[slot] = residual_key(w, source, target);
          w.res_ids[slot] = slot;
        }
      }
    }
  }
  if (worker < w.workers && targets) {
    atomicAdd((unsigned long long *)&w.stats[3 * layer], (unsigned long long)targets);
    atomicAdd((unsigned long long *)&w.stats[3 * layer + 1], (unsigned long long)tested_all);
    atomicAdd((unsigned long long *)&w.stats[3 * layer + 2], (unsigned long long)hidden_all);
    atomicMax(w.status + ST_MAXH, max_hidden);
  }
  if (ties) atomicAdd(w.status + ST_TIES, (int)ties);
}

// advance (kernels.py:924-938)
__global__ void k_layer_advance(Ws w) {
  if (w.control[C_LOOP] == 0) return;
  int *s = w.status;
  s[ST_CUR] = s[ST_NEXT];
  s[ST_PEAK] = max(s[ST_PEAK], s[ST_CUR]);
  s[ST_NEXT] = 0;
  s[ST_PASS] += 1;
  if (s[ST_CUR] > 0 && s[ST_PASS] >= w.limit && w.all_layers) s[ST_ERR] |= ERR_LAYERS;
  int loop = (s[ST_CUR] > 0 && s[ST_PASS] < w.limit && s[ST_ERR] == 0) ? 1 : 0;
  w.control[C_LOOP] = loop;
  w.control[C_SORTN] = loop ? s[ST_CUR] : 0;
}

// Move the appended pieces of pool 1 to pool 0 in sorted order.
__global__ void k_residual_gather(Ws w) {
  if (load_i(w.control + C_LOOP) == 0) return;
  const int n = w.status[ST_CUR];
  const long R = w.cap.frontier;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int from = w.res_ids[i];
    int pn = w.res_n[1][from];
    const double2 *a = w.res_xy[1] + from;
    double2 *b = w.res_xy[0] + i;
    for (int j = 0; j < pn && j < MAXV; ++j) b[j * R] = a[j * R];
    w.res_n[0][i] = pn;
    w.res_src[0][i] = w.res_src[1][from];
    w.res_box[0][i] = w.res_box[1][from];
  }
}

// Clear every source's run of pool 0.
__global__ void k_runs_reset(Ws w) {
  if (load_i(w.control + C_LOOP) == 0) return;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < w.nf; j += gridDim.x * blockDim.x) {
    w.run_first[j] = -1;
    w.run_last[j] = -1;
  }
}

// Each source's pieces are consecutive in pool 0: record where they start and end.
__global__ void k_runs(Ws w) {
  if (load_i(w.control + C_LOOP) == 0) return;
  const int n = w.status[ST_CUR];
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    int s = w.res_src[0][i];
    if (i == 0 || w.res_src[0][i - 1] != s) w.run_first[s] = i;
    if (i == n - 1 || w.res_src[0][i + 1] != s) w.run_last[s] = i;
  }
}

// One ring operation on one lane, for unit tests of the device functions:
// op 0 clean (both routes), op 1 split_edge by the edge param[0..3] = (a, b)
// (the layered route's split_at in the edge frame, strictly convex outputs),
// op 2 convexify_ring. Rings are contiguous [MAXV] double2; out holds two.
__global__ void k_ring_test(int op, const double2 *xy, int n, const double *param, double2 *out, int *out_n) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  RingRef a{out, 1}, b{out + MAXV, 1};
  if (op == 1) {
    RingRef src{const_cast<double2 *>(xy), 1};
    double2 e0 = make_double2(param[0], param[1]), e1 = make_double2(param[2], param[3]);
    double2 d = make_double2(e1.x - e0.x, e1.y - e0.y);
    int in_n, o_n;
    split_at(src, n, make_double3(-d.y, d.x, 0.0), e0, a, in_n, b, o_n);
    out_n[0] = in_n;
    out_n[1] = o_n;
    return;
  }
  for (int j = 0; j < n && j < MAXV; ++j) a.at(j) = xy[j];
  out_n[0] = op == 0 ? clean_slot(out, 1, n) : convexify_ring(a, n);
  out_n[1] = 0;
}

static inline dim3 grid_for(long n, int block = 256) {
  long g = (n + block - 1) / block;
  return dim3((unsigned)(g < 1 ? 1 : g));
}

static inline dim3 capped_grid(long n, long cap) {
  long g = (n + 255) / 256;
  return dim3((unsigned)(g > cap ? cap : (g < 1 ? 1 : g)));
}

// Sort the appended pieces by (source, parent slot), move them to pool 0 and
// record each source's run (launches that return at once after the last pass).
static void order_residual(const Ws &w, hipStream_t s) {
  radix_sort(s, w.res_keys, w.res_ids, w.res_keys_alt, w.res_ids_alt, w.control + C_SORTN, 0, w.cap.frontier, 0,
             w.res_key_bits, w.sort);
  hipLaunchKernelGGL(k_residual_gather, capped_grid(w.cap.frontier, 8192), dim3(256), 0, s, w);
  hipLaunchKernelGGL(k_runs_reset, capped_grid(w.nf, 256), dim3(256), 0, s, w);
  hipLaunchKernelGGL(k_runs, capped_grid(w.cap.frontier, 8192), dim3(256), 0, s, w);
}

// Every pass is in the graph (no conditional nodes): a pass after the last
// one returns at its first instruction.
void enqueue_layered(const Ws &w, hipStream_t s, Timeline *t) {
  long reset_n = std::max<long>(w.nf, 3L * w.cap.layers);
  long g = (reset_n + 255) / 256;
  hipLaunchKernelGGL(k_layered_reset, dim3((unsigned)(g > 1024 ? 1024 : (g < 1 ? 1 : g))), dim3(256), 0, s, w);
  if (w.nf > 0) {
    hipLaunchKernelGGL(k_contour_bounds, grid_for(w.nf), dim3(256), 0, s, w);
    Hierarchy ct = w.contour_tree;
    build_hierarchy(s, w.contour_boxes, nullptr, w.nf, w.nf, false, ct, w.sort);
    mark(t, s, "contour_hierarchy");
    hipLaunchKernelGGL(k_classify, grid_for(w.nf), dim3(256), 0, s, w);
  }
  hipLaunchKernelGGL(k_start_loop, dim3(1), dim3(1), 0, s, w);
  order_residual(w, s);
  mark(t, s, "k_classify");
  for (int layer = 0; layer < w.limit; ++layer) {
    hipLaunchKernelGGL(k_peel, grid_for(w.workers), dim3(256), 0, s, w);
    hipLaunchKernelGGL(k_layer_advance, dim3(1), dim3(1), 0, s, w);
    order_residual(w, s);
  }
  mark(t, s, "k_peel");
}

}  // namespace peel

extern "C" int peel_ring_test(int32_t op, const double *xy, int32_t n, const double *param, double *out_xy,
                              int32_t *out_n, void *stream) {
  if (op < 0 || op > 2 || n < 0 || n > peel::MAXV || !xy || !out_xy || !out_n || (op == 1 && !param)) return 1;
  hipLaunchKernelGGL(peel::k_ring_test, dim3(1), dim3(64), 0, static_cast<hipStream_t>(stream), op,
                     reinterpret_cast<const double2 *>(xy), n, param, reinterpret_cast<double2 *>(out_xy), out_n);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}
