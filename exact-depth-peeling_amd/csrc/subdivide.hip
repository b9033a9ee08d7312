// Stage 2: contour-guided supporting-line subdivision as one persistent
// work-queue kernel.
//
// Reference semantics: warp_peel/frontier.py:15-245 (seed, clean,
// front_segment, subdivide, copy_frontier) driven by a CUDA conditional while
// node (warp_peel/events.py:222-264, 804). HIP graphs have no conditional
// nodes, so the rounds become a device work queue instead:
//
//  * Seeds (the projected sources, read in place from their rings) are
//    handed out by one global ticket counter, one wave-aggregated claim at a
//    time. Every cut stores its two children in two consecutive frontier
//    slots k, k + 1.
//  * The lane that cuts goes on with the left child itself (depth first, no
//    hand-off); the right child goes to its block's queue in LDS and is
//    taken by an idle lane of the same block. Children never leave their
//    block, so handing one over needs no agent-scope fence: the child's data
//    is stored by the same CU that reads it (workgroup scope), and agent-scope
//    fences would slow every wave's loads (each acquire invalidates caches,
//    each release writes the L2 back).
//  * Per visit a lane inspects at most 32 contour candidates from the task's
//    hierarchy cursor, exactly like one reference round; a visit without a
//    cut keeps the task and continues from the cursor.
//  * Each block counts its live tasks in LDS: a seed taken adds one, a queued
//    right child adds one, a finished task subtracts one. A lane leaves once
//    the seeds are all handed out and its block has no live task. An error
//    bit also exits all lanes, and a wall-clock guard turns a stall into
//    error bit 2048.
//
// Each task's sequence of cuts depends only on its own ring and cursor, so the
// set of final patches is the same as the reference's round-synchronous
// frontier for the same hierarchy (only the output order differs). A cut
// whose child collapses carries the parent on (frontier.py:200-206): here the
// lane simply goes on with the same ring from the new cursor.
#include "workspace.hpp"

namespace peel {

// The seed counter and the frontier slot counter (C_HEAD, C_ALLOC) are
// cleared by the pipeline's first kernel, k_reset.

struct Slot {
  double2 *xy;  // vertex j at xy[j * stride]
  int stride;
};

__device__ __forceinline__ Slot slot_of(const Ws &w, int task) {
  if (task < w.nf) return Slot{w.src_xy + task, w.nf};
  return Slot{w.fr_xy + (task - w.nf), w.cap.frontier};
}

__device__ __forceinline__ double2 vtx(const Slot &s, int j) { return s.xy[(long)j * s.stride]; }


// Each lane has three rings in LDS: its current ring and the two children of
// a cut, which are built and cleaned there (a cut and its cleanup make no
// global round trips; only the right child, which another lane takes, is
// copied to its slot). The first SUB_LDS vertices of a ring live in LDS,
// vertex-major so the 64 lanes of a wave read vertex j of their rings from
// consecutive banks; the rare longer ring keeps the rest in its slot (slots
// are never reused within a run). 288 B of LDS per lane (six vertices a
// ring: measured faster than twelve on both the torus knot and the shells,
// and faster than four or three, which fit three or four blocks a CU instead
// of two but send more ring vertices to their slots).
constexpr int SUB_THREADS = 256;
constexpr int SUB_LDS = 6;
constexpr int SUB_RING = SUB_LDS * SUB_THREADS;  // LDS words of one ring for the block
__shared__ double2 sub_ring_lds[3 * SUB_RING];
// The block's queue of right children (frontier slots; -1: empty), a ring of
// BQ entries with monotonic head and tail, and the block's live tasks.
constexpr int BQ = 1024;
constexpr int BQ_MARGIN = SUB_THREADS;  // room for every lane of the block to push at once
__shared__ int sub_queue[BQ];
__shared__ int sub_head, sub_tail, sub_live;
// a lane keeps up to DEFER right children itself when its block's queue is full
constexpr int DEFER = 4;
// kept out of line so the compiler does not fold the LDS and global accesses
// into generic ones
__device__ __noinline__ double2 slot_vertex(const double2 *slot, long stride, int j) {
  return slot[(long)j * stride];
}
__device__ __noinline__ void slot_store(double2 *slot, long stride, int j, double2 v) {
  slot[(long)j * stride] = v;
}
struct LdsRing {
  int lane;       // this lane's column
  int base;       // which of the lane's three rings: 0, SUB_RING or 2 * SUB_RING
  double2 *slot;  // the ring's slot: vertex j at slot[j * stride]
  long stride;
  __device__ double2 operator[](int j) const {
    if (j < SUB_LDS) return sub_ring_lds[base + j * SUB_THREADS + lane];
    return slot_vertex(slot, stride, j);
  }
  __device__ void put(int j, double2 v) const {
    if (j < SUB_LDS)
      sub_ring_lds[base + j * SUB_THREADS + lane] = v;
    else
      slot_store(slot, stride, j, v);
  }
};

// frontier.py:81-103
__device__ bool front_segment(const LdsRing &s, int n, double2 a, double2 b, double za, double zb) {
  double low = 0.0, high = 1.0;
  for (int j = 0; j <= n; ++j) {
    double da = za, db = zb;
    if (j < n) {
      double2 u = s[j], v = s[(j + 1) % n];
      da = cross3(u, v, a);
      db = cross3(u, v, b);
    }
    if (da <= 0.0 && db <= 0.0) return false;
    if (da < 0.0) low = fmax(low, da / (da - db));
    if (db < 0.0) high = fmin(high, da / (da - db));
  }
  return high > low;
}

// front_segment with a binary32 filter on its ring-edge tests: an edge whose
// two binary32 orientations are both beyond `tol` on the negative side
// rejects (binary64 would reject there or earlier), both beyond it on the
// positive side leaves the interval as binary64 would; only the other edges
// are evaluated in binary64, where the interval needs their values. The result
// therefore equals front_segment's.
__device__ bool front_segment32(const LdsRing &s, int n, double2 a, double2 b, double za, double zb, float tol,
                                int &filtered, int &fallbacks) {
  const float2 a32 = to32(a), b32 = to32(b);
  double low = 0.0, high = 1.0;
  double2 u = s[0];
  float2 u32 = to32(u);
  for (int j = 0; j < n; ++j) {
    double2 v = s[(j + 1) % n];
    float2 v32 = to32(v);
    float da32 = cross32(u32, v32, a32), db32 = cross32(u32, v32, b32);
    if (da32 < -tol && db32 < -tol) {
      filtered += 1;
      return false;
    }
    if (da32 > tol && db32 > tol) {
      filtered += 1;
    } else {
      fallbacks += 1;
      double da = cross3(u, v, a), db = cross3(u, v, b);
      if (da <= 0.0 && db <= 0.0) return false;
      if (da < 0.0) low = fmax(low, da / (da - db));
      if (db < 0.0) high = fmin(high, da / (da - db));
    }
    u = v;
    u32 = v32;
  }
  if (za <= 0.0 && zb <= 0.0) return false;
  if (za < 0.0) low = fmax(low, za / (za - zb));
  if (zb < 0.0) high = fmin(high, za / (za - zb));
  return high > low;
}

__device__ __forceinline__ int append(const LdsRing &r, int n, double2 x) {
  if (n < MAXV) r.put(n, x);
  return n + 1;
}

// clean (frontier.py:45-71 = kernels.py:146-176), in place on a lane's ring;
// the same operations as clean_slot, in the same order.
__device__ int clean_ring(const LdsRing &r, int n) {
  if (n > MAXV) return n;
  int count = 0;
  double2 last = make_double2(0, 0);
  for (int j = 0; j < n; ++j) {
    double2 x = r[j];
    if (count == 0 || dlength(make_double2(x.x - last.x, x.y - last.y)) > EPS * 0.01) {
      r.put(count, x);
      last = x;
      count += 1;
    }
  }
  if (count > 1) {
    double2 first = r[0];
    if (dlength(make_double2(first.x - last.x, first.y - last.y)) <= EPS * 0.01) count -= 1;
  }
  if (count < 3) return 0;
  double2 origin = r[0];
  double area2 = 0.0, perimeter = 0.0;
  for (int j = 0; j < count; ++j) {
    double2 a = r[j], b = r[(j + 1) % count];
    area2 += cross3(origin, a, b);
    perimeter += fabs(b.x - a.x) + fabs(b.y - a.y);
  }
  if (area2 <= EPS * perimeter) return 0;
  return count;
}

// save_polygon (kernels.py:579-596) with layer 0; ranks are assigned later.
// The output cursor is claimed once per wave.
__device__ void save_output(const Ws &w, const LdsRing &s, int n, int source, int o) {
  if (o < w.cap.regions) {
    for (int j = 0; j < n; ++j) w.out_xy[(long)j * w.cap.regions + o] = s[j];
    w.out_n[o] = n;
    w.out_src[o] = source;
    w.out_layer[o] = 0;
  } else {
    set_error(w.status, ERR_OUTPUT);
  }
}

__global__ __launch_bounds__(SUB_THREADS) void k_subdivide(Ws w) {
  LdsRing ring{(int)threadIdx.x, 0, nullptr, 0};
  TrailWalk walk;
  bool fresh = false;  // the task's walk starts at its first visit, from its binary cursor
  // A child that this lane goes on with may keep its parent's walk: the
  // parent's box (plo, phi) bounds the masks the walk holds.
  bool inherit = false;
  float4 plo = make_float4(0.f, 0.f, 0.f, 0.f), phi = plo;
  const int nf = w.nf;
  const Tree tree = w.contour_tree.tree();
  int lane = threadIdx.x & 63;
  uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  int task = -1;
  int n = 0, source = 0, node = 1, gen = 0;
  long long visits = 0, tests = 0, splits = 0;
  int f_straddle = 0, b_straddle = 0, f_front = 0, b_front = 0;
  int max_gen = 0, peak = 0;
  int deferred[DEFER], ndefer = 0;  // right children this lane keeps (queue full)
  bool seeds_left = true;           // wave-uniform
  uint64_t idle_since = wall_clock64();
  for (int i = threadIdx.x; i < BQ; i += SUB_THREADS) sub_queue[i] = -1;
  if (threadIdx.x == 0) sub_head = sub_tail = sub_live = 0;
  __syncthreads();
  for (int iter = 0;; ++iter) {
    if ((iter & 7) == 0 && load_i(w.status + ST_ERR) != 0) break;
    // ---- a lane without a task takes one of its own deferred children, else
    // one from its block's queue, else a seed
    int got = -1;  // frontier slot of a child taken
    bool took = false;
    if (task < 0 && ndefer > 0) {
      ndefer -= 1;
#pragma unroll
      for (int d = 0; d < DEFER; ++d)
        if (d == ndefer) got = deferred[d];
    }
    {
      const bool need = task < 0 && got < 0;
      const uint64_t mask = __ballot(need);
      if (mask) {
        const int count = __popcll((unsigned long long)mask), rank = __popcll(mask & lt);
        const int leader = __ffsll((unsigned long long)mask) - 1;
        int first = 0, take = 0;
        if (lane == leader) {
          int h = __hip_atomic_load(&sub_head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          while (true) {
            int t = __hip_atomic_load(&sub_tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            take = min(count, t - h);
            if (take <= 0) {
              take = 0;
              break;
            }
            int old = atomicCAS(&sub_head, h, h + take);
            if (old == h) break;
            h = old;
          }
          first = h;
        }
        first = __shfl(first, leader);
        take = __shfl(take, leader);
        if (need && rank < take) {
          // the producer stores the entry right after reserving it
          const int pos = (first + rank) & (BQ - 1);
          int slot;
          while ((slot = __hip_atomic_load(&sub_queue[pos], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP)) < 0) {
          }
          __hip_atomic_store(&sub_queue[pos], -1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
          got = slot;
        }
        const bool want = need && rank >= take && seeds_left;
        const uint64_t smask = __ballot(want);
        if (smask) {
          const int sl = __ffsll((unsigned long long)smask) - 1, scount = __popcll((unsigned long long)smask);
          int base = 0;
          if (lane == sl) base = atomicAdd(w.control + C_HEAD, scount);
          base = __shfl(base, sl);
          if (base + scount >= nf) seeds_left = false;
          const int ticket = base + __popcll(smask & lt);
          if (want && ticket < nf) {
            n = w.src_n[ticket];
            if (n > 0) {
              task = ticket;
              source = ticket;
              node = 1;
              gen = 0;
              took = true;
              atomicAdd(&sub_live, 1);
            }
          }
        }
      }
    }
    if (got >= 0) {
      // a right child of this block: its slot was stored by this CU
      task = nf + got;
      n = w.fr_n[got];
      source = w.fr_src[got];
      node = w.fr_node[got];
      gen = w.fr_gen[got];
      took = true;
    }
    if (took) {
      Slot src = slot_of(w, task);
      for (int j = 0; j < n && j < SUB_LDS; ++j) ring.put(j, vtx(src, j));
      ring.slot = src.xy;
      ring.stride = src.stride;
      fresh = true;
    }
    // (a ballot inside the branch below would only see the idle lanes)
    const bool wave_has_task = __ballot(task >= 0) != 0;
    if (task < 0) {
      if (!seeds_left && ndefer == 0 &&
          __hip_atomic_load(&sub_live, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP) == 0)
        break;
      if (!wave_has_task) {
        __builtin_amdgcn_s_sleep(8);
        if (wall_clock64() - idle_since > 1000000000ull) {
          set_error(w.status, ERR_ROUNDS);
          break;
        }
      }
      continue;
    }
    idle_since = wall_clock64();
    // ---- one visit: at most 32 contour candidates (frontier.py:106-223)
    visits += 1;
    if (gen > max_gen) max_gen = gen;
    if (gen >= w.cap.rounds) {
      set_error(w.status, ERR_ROUNDS);
      break;
    }
    const LdsRing &s = ring;
    double4 score = w.scores[source];
    double lx = 1e30, ly = 1e30, lz = 1e30, hx = -1e30, hy = -1e30, hz = -1e30;
    for (int j = 0; j < n; ++j) {
      double2 x = s[j];
      double z = value(score, x);
      lx = fmin(lx, x.x), ly = fmin(ly, x.y), lz = fmin(lz, z);
      hx = fmax(hx, x.x), hy = fmax(hy, x.y), hz = fmax(hz, z);
    }
    float4 lo, hi;
    pad_box(lx, ly, lz, hx, hy, hz, lo, hi);
    lo.z = -1e30f;
    if (fresh) {
      // the leaves still to visit are those from the leftmost leaf under the
      // cursor on (a task continues after the candidate that cut its parent)
      fresh = false;
      if (inherit && lo.x >= plo.x && lo.y >= plo.y && hi.x <= phi.x && hi.y <= phi.y && hi.z <= phi.z) {
        // Inside the parent's box: every node the child's walk would visit
        // overlapped the parent's box, so it is pending in the parent's walk.
        // Pending nodes are tested against the child's box when expanded;
        // only the current bucket's pending leaves need testing again.
        if (walk.bucket) walk.bucket &= walk.bucket_mask(tree, walk.cur, lo, hi);
      } else {
        walk.trail = 0;
        walk.bucket = 0;
        if (node == 1) {
          walk.start(tree, 1, w.bucket_log, lo, hi);  // a seed: from the root
        } else if (node > 0) {
          int first = node << (__clz(node) - __clz(tree.leaves));
          walk.start_at(tree, 1, first, w.bucket_log, lo, hi);
        }
      }
      inherit = false;
    }
    bool done = false, cut = false;
    for (int step = 0; step < 32; ++step) {
      int leaf = walk.next_node(tree, lo, hi);
      if (leaf < 0) {
        node = 0;
        done = true;
        break;
      }
      node = successor(leaf);
      tests += 1;
      // the candidate's record in leaf order: endpoints and minimum depths
      const int c = leaf - tree.leaves;
      double2 a = w.cl_a[c], b = w.cl_b[c], near = w.cl_near[c];
      bool positive = false, negative = false;
      float tol = 0.f;
      if (w.filter32) {
        // binary32 straddle signs beyond the candidate's static bound, the
        // rest in binary64 (same signs, so the same decision)
        tol = candidate_tol32(lx, ly, hx, hy, a, b);
        const float2 a32 = to32(a), b32 = to32(b);
        for (int j = 0; j < n; ++j) {
          double2 p = s[j];
          float side32 = cross32(a32, b32, to32(p));
          if (side32 > tol) {
            positive = true;
            f_straddle += 1;
          } else if (side32 < -tol) {
            negative = true;
            f_straddle += 1;
          } else {
            b_straddle += 1;
            double side = cross3(a, b, p);
            positive |= side > 0.0;
            negative |= side < 0.0;
          }
        }
      } else {
        for (int j = 0; j < n; ++j) {
          double side = cross3(a, b, s[j]);
          positive |= side > 0.0;
          negative |= side < 0.0;
        }
      }
      if (!(positive && negative)) continue;
      double za = value(score, a) - near.x;
      double zb = value(score, b) - near.y;
      if (w.filter32 ? !front_segment32(s, n, a, b, za, zb, tol, f_front, b_front)
                     : !front_segment(s, n, a, b, za, zb))
        continue;
      // ---- cut by the supporting line of (a, b), relative to a
      int k = atomicAdd(w.control + C_ALLOC, 2);
      if (k + 2 > w.cap.frontier) {
        set_error(w.status, ERR_RESIDUAL);
        cut = true;
        break;
      }
      const int fcap = w.cap.frontier;
      // the children in the lane's two other LDS rings (slots k and k + 1
      // hold vertices past SUB_LDS)
      const LdsRing L{ring.lane, (ring.base + SUB_RING) % (3 * SUB_RING), w.fr_xy + k, fcap};
      const LdsRing R{ring.lane, (ring.base + 2 * SUB_RING) % (3 * SUB_RING), w.fr_xy + k + 1, fcap};
      int left = 0, right = 0;
      double2 d = make_double2(b.x - a.x, b.y - a.y);
      for (int j = 0; j < n; ++j) {
        double2 u = s[j], v = s[(j + 1) % n];
        double du = -d.y * (u.x - a.x) + d.x * (u.y - a.y) + 0.0;
        double dv = -d.y * (v.x - a.x) + d.x * (v.y - a.y) + 0.0;
        if (du >= 0.0) left = append(L, left, u);
        if (du <= 0.0) right = append(R, right, u);
        if ((du > 0.0 && dv < 0.0) || (du < 0.0 && dv > 0.0)) {
          double t = du / (du - dv);
          double2 x = make_double2(u.x + t * (v.x - u.x), u.y + t * (v.y - u.y));
          left = append(L, left, x);
          right = append(R, right, x);
        }
      }
      left = clean_ring(L, left);
      right = clean_ring(R, right);
      if (left > MAXV || right > MAXV) {
        set_error(w.status, ERR_POLY);
        cut = true;
        break;
      }
      if (left > 0 && right > 0) {
        splits += 1;
        // the right child goes to the queue, whole in its slot
        for (int j = 0; j < right && j < SUB_LDS; ++j) R.slot[(long)j * fcap] = R[j];
        w.fr_n[k + 1] = right;
        w.fr_src[k + 1] = source;
        w.fr_node[k + 1] = node;
        w.fr_gen[k + 1] = gen + 1;
        // counted live before it can be taken; the slot's stores complete
        // before the queue entry that names it
        int live = atomicAdd(&sub_live, 1) + 1;
        peak = max(peak, live);
        __builtin_amdgcn_s_waitcnt(0);
        int tail = __hip_atomic_load(&sub_tail, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        int head = __hip_atomic_load(&sub_head, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        if (tail - head < BQ - BQ_MARGIN) {
          int pos = atomicAdd(&sub_tail, 1) & (BQ - 1);
          __hip_atomic_store(&sub_queue[pos], k + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
        } else if (ndefer < DEFER) {
#pragma unroll
          for (int d = 0; d < DEFER; ++d)
            if (d == ndefer) deferred[d] = k + 1;
          ndefer += 1;
        } else {
          set_error(w.status, ERR_RESIDUAL);
        }
        // and this lane goes on with the left child
        ring = L;
        n = left;
      }
      // a collapsed child leaves the parent's ring to go on from the cursor
      inherit = true;
      plo = lo;
      phi = hi;
      cut = true;
      break;
    }
    bool finish = !cut && done;
    int o = wave_claim32(w.status + ST_POLYS, finish);
    if (finish) {
      save_output(w, s, n, source, o);
      atomicSub(&sub_live, 1);
      task = -1;
    } else {
      // a child (or a task whose visit ran out of candidates) continues from
      // the cursor in the next visit, like a re-enqueued patch
      gen += 1;
      fresh = fresh || cut;
    }
  }
  // Per-wave partial counters, summed by the next kernel: same-address
  // atomics from every wave at once would serialise at the queue's end.
  long long part[SUB_COUNTERS] = {visits, tests, splits, f_straddle, b_straddle, f_front, b_front, 0, 0};
  for (int k = 0; k < SUB_SUMS; ++k) part[k] = wave_sum(part[k]);
  part[SUB_SUMS] = wave_max(max_gen) + 1;
  part[SUB_SUMS + 1] = wave_max(peak);
  const int wave_id = blockIdx.x * (SUB_THREADS / 64) + (threadIdx.x >> 6);
  if ((threadIdx.x & 63) == 0) {
    if (wave_id < SUB_PARTIAL_WAVES) {
      for (int k = 0; k < SUB_COUNTERS; ++k) w.sub_partials[(long)wave_id * SUB_COUNTERS + k] = part[k];
    } else {
      const int slot[SUB_SUMS] = {M_VISITS, M_CONTOUR_TESTS, M_SPLITS, M_FILTER_STRADDLE,
                                  M_FALLBACK_STRADDLE, M_FILTER_FRONT, M_FALLBACK_FRONT};
      for (int k = 0; k < SUB_SUMS; ++k)
        if (part[k]) atomicAdd((unsigned long long *)&w.metrics[slot[k]], (unsigned long long)part[k]);
      atomicMax((unsigned long long *)&w.metrics[M_ROUNDS], (unsigned long long)part[SUB_SUMS]);
      atomicMax(w.status + ST_PEAK, (int)part[SUB_SUMS + 1]);
    }
  }
}

void enqueue_subdivide(const Ws &w, hipStream_t s, Timeline *t) {
  hipLaunchKernelGGL(k_subdivide, dim3(w.persistent_blocks), dim3(SUB_THREADS), 0, s, w);
  mark(t, s, "k_subdivide");
}

}  // namespace peel

namespace peel {
int subdivide_blocks_per_cu() {
  int blocks = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k_subdivide, SUB_THREADS, 0) != hipSuccess)
    return 2;
  return blocks;
}
}  // namespace peel
