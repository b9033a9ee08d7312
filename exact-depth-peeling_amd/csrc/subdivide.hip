// Stage 2: contour-guided supporting-line subdivision as one persistent
// work-queue kernel.
//
// Reference semantics: warp_peel/frontier.py:15-245 (seed, clean,
// front_segment, subdivide, copy_frontier) driven by a CUDA conditional while
// node (warp_peel/events.py:222-264, 804). HIP graphs have no conditional
// nodes, so the rounds become a device work queue instead:
//
//  * Tickets [0, nf) are the projected sources themselves (seeds, read in
//    place from the source rings); ticket nf + q is queue entry q. Every cut
//    stores its two children in two consecutive frontier slots k, k + 1.
//  * The lane that cuts goes on with the left child itself (depth first, no
//    hand-off) and publishes the right child as the next queue entry: the
//    entry holds 1 + the child's slot, stored after a release fence.
//  * A lane claims the next ticket (wave-aggregated atomicAdd) and processes
//    it once its entry is published (acquire at agent scope). Per visit it
//    inspects at most 32 contour candidates from the task's hierarchy cursor,
//    exactly like one reference round; a visit without a cut keeps the task
//    and continues from the cursor.
//  * C_PENDING counts live tasks: seeds start it at nf, a split adds one (the
//    right child; the left child takes the parent's place), a finished task
//    subtracts one. It only reaches zero when no task is live and no cut is
//    in flight, so a lane that holds a ticket that is not published and sees
//    zero can exit: the entry will never be published. Every wave reaches that
//    exit; an error bit also exits all lanes, and a wall-clock guard turns a
//    stall into error bit 2048.
//
// Each task's sequence of cuts depends only on its own ring and cursor, so the
// set of final patches is the same as the reference's round-synchronous
// frontier for the same hierarchy (only the output order differs). A cut
// whose child collapses carries the parent on (frontier.py:200-206): here the
// lane simply goes on with the same ring from the new cursor.
#include "workspace.hpp"

namespace peel {

// The queue's ready flags and counters (C_ALLOC, C_HEAD, C_PENDING = nf) are
// set by the pipeline's first kernel, k_reset.

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
// are never reused within a run). 576 B of LDS per lane: one block per CU.
constexpr int SUB_THREADS = 256;
constexpr int SUB_LDS = 12;
constexpr int SUB_RING = SUB_LDS * SUB_THREADS;  // LDS words of one ring for the block
__shared__ double2 sub_ring_lds[3 * SUB_RING];
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
  int task = -1, ticket = -1;
  int n = 0, source = 0, node = 1, gen = 0;
  long long visits = 0, tests = 0, splits = 0;
  int f_straddle = 0, b_straddle = 0, f_front = 0, b_front = 0;
  int max_gen = 0;
  uint64_t idle_since = wall_clock64();
  while (true) {
    if (load_i(w.status + ST_ERR) != 0) break;
    if (task < 0) {
      bool need = ticket < 0;
      uint64_t mask = __ballot(need);
      if (mask) {
        int leader = __ffsll((unsigned long long)mask) - 1;
        int base = 0;
        if (lane == leader) base = atomicAdd(w.control + C_HEAD, __popcll(mask));
        base = __shfl(base, leader);
        if (need) ticket = base + __popcll(mask & lt);
      }
      // A ticket past the frontier capacity can never be published: it only
      // waits for the pending count to reach zero. Polling is relaxed; the
      // acquire fence runs once, when a published entry is taken.
      int entry = ticket >= nf && ticket - nf < w.cap.frontier ? load_i(w.fr_ready + (ticket - nf)) : 0;
      bool ready = ticket < nf || entry != 0;
      if (__ballot(ready && ticket >= nf)) acquire_fence();
      if (ready) {
        task = ticket < nf ? ticket : nf + entry - 1;
        ticket = -1;
        if (task < nf) {
          n = w.src_n[task];
          source = task;
          node = 1;
          gen = 0;
        } else {
          int q = task - nf;
          n = w.fr_n[q];
          source = w.fr_src[q];
          node = w.fr_node[q];
          gen = w.fr_gen[q];
        }
        if (n == 0) {
          atomicSub(w.control + C_PENDING, 1);
          task = -1;
        } else {
          Slot src = slot_of(w, task);
          for (int j = 0; j < n && j < SUB_LDS; ++j) ring.put(j, vtx(src, j));
          ring.slot = src.xy;
          ring.stride = src.stride;
          fresh = true;
        }
      }
    }
    // (a ballot inside the branch below would only see the idle lanes)
    const bool wave_has_task = __ballot(task >= 0) != 0;
    if (task < 0) {
      if (load_i(w.control + C_PENDING) == 0) break;
      if (!wave_has_task) {
        __builtin_amdgcn_s_sleep(32);
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
        int q = atomicAdd(w.control + C_QUEUE, 1);
        release_fence();
        __hip_atomic_store(w.fr_ready + q, k + 2, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        int live = atomicAdd(w.control + C_PENDING, 1) + 1;
        atomicMax(w.status + ST_PEAK, live);
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
      atomicSub(w.control + C_PENDING, 1);
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
  long long part[SUB_COUNTERS] = {visits, tests, splits, f_straddle, b_straddle, f_front, b_front, 0};
  for (int k = 0; k < SUB_COUNTERS - 1; ++k) part[k] = wave_sum(part[k]);
  part[SUB_COUNTERS - 1] = wave_max(max_gen) + 1;
  const int wave_id = blockIdx.x * (SUB_THREADS / 64) + (threadIdx.x >> 6);
  if ((threadIdx.x & 63) == 0) {
    if (wave_id < SUB_PARTIAL_WAVES) {
      for (int k = 0; k < SUB_COUNTERS; ++k) w.sub_partials[(long)wave_id * SUB_COUNTERS + k] = part[k];
    } else {
      const int slot[SUB_COUNTERS - 1] = {M_VISITS, M_CONTOUR_TESTS, M_SPLITS, M_FILTER_STRADDLE,
                                          M_FALLBACK_STRADDLE, M_FILTER_FRONT, M_FALLBACK_FRONT};
      for (int k = 0; k < SUB_COUNTERS - 1; ++k)
        if (part[k]) atomicAdd((unsigned long long *)&w.metrics[slot[k]], (unsigned long long)part[k]);
      atomicMax((unsigned long long *)&w.metrics[M_ROUNDS], (unsigned long long)part[SUB_COUNTERS - 1]);
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
