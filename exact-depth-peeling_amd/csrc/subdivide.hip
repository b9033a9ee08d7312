// Stage 2: contour-guided supporting-line subdivision as one persistent
// work-queue kernel.
//
// Reference semantics: warp_peel/frontier.py:15-245 (seed, clean,
// front_segment, subdivide, copy_frontier) driven by a CUDA conditional while
// node (warp_peel/events.py:222-264, 804). HIP graphs have no conditional
// nodes, so the rounds become a device work queue instead:
//
//  * Tasks are slots. Slots [0, nf) are the projected sources themselves
//    (seeds, read in place from the source rings); every cut allocates two
//    consecutive frontier slots [nf + k, nf + k + 2).
//  * A lane claims the next ticket (= slot index, wave-aggregated atomicAdd)
//    and processes it once the slot's ready flag is published (release/acquire
//    at agent scope). Per visit it inspects at most 32 contour candidates from
//    the slot's hierarchy cursor, exactly like one reference round; a visit
//    without a cut keeps the task and continues from the cursor.
//  * C_PENDING counts live tasks: seeds start it at nf, a cut adds two and
//    retires one (a single +1), a finished task subtracts one. It only reaches
//    zero when no task is live and no cut is in flight, so a lane that holds a
//    ticket that is not ready and sees zero can exit: the ticket will never be
//    allocated. Every wave reaches that exit; an error bit also exits all
//    lanes, and a wall-clock guard turns a stall into error bit 2048.
//
// Each task's sequence of cuts depends only on its own ring and cursor, so the
// set of final patches is the same as the reference's round-synchronous
// frontier for the same hierarchy (only the output order differs).
#include "workspace.hpp"

namespace peel {

// The queue's ready flags and counters (C_ALLOC, C_HEAD, C_PENDING = nf) are
// set by the pipeline's first kernel, k_reset.

struct Slot {
  const double2 *xy;  // vertex j at xy[j * stride]
  int stride;
};

__device__ __forceinline__ Slot slot_of(const Ws &w, int task) {
  if (task < w.nf) return Slot{w.src_xy + task, w.nf};
  return Slot{w.fr_xy + (task - w.nf), w.cap.frontier};
}

__device__ __forceinline__ double2 vtx(const Slot &s, int j) { return s.xy[(long)j * s.stride]; }


// The first SUB_LDS vertices of a lane's current ring live in LDS,
// vertex-major so the 64 lanes of a wave read vertex j of their rings from
// consecutive banks; the rare longer ring reads the rest from its slot (slots
// are never reused within a run). 192 B of LDS per lane lets several blocks
// share a CU.
constexpr int SUB_THREADS = 256;
constexpr int SUB_LDS = 12;
__shared__ double2 sub_ring_lds[SUB_LDS * SUB_THREADS];
// kept out of line so the compiler does not fold the LDS and global loads
// into one generic load
__device__ __noinline__ double2 slot_vertex(const double2 *slot, long stride, int j) {
  return slot[(long)j * stride];
}
struct LdsRing {
  int lane;             // this lane's column of sub_ring_lds
  const double2 *slot;  // the task's slot: vertex j at slot[j * stride]
  long stride;
  __device__ double2 operator[](int j) const {
    if (j < SUB_LDS) return sub_ring_lds[j * SUB_THREADS + lane];
    return slot_vertex(slot, stride, j);
  }
  __device__ void put(int j, double2 v) const { sub_ring_lds[j * SUB_THREADS + lane] = v; }
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

__device__ __forceinline__ int append_slot(double2 *xy, int stride, int n, double2 x) {
  if (n < MAXV) xy[(long)n * stride] = x;
  return n + 1;
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
  LdsRing ring{(int)threadIdx.x, nullptr, 0};
  TrailWalk walk;
  bool fresh = false;  // the task's walk starts at its first visit, from its binary cursor
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
      // A ticket past the frontier capacity can never be allocated: it only
      // waits for the pending count to reach zero. Polling is relaxed; the
      // acquire fence runs once, when a published slot is taken.
      bool ready = ticket < nf || (ticket - nf < w.cap.frontier &&
                                   load_i(w.fr_ready + (ticket - nf)) != 0);
      if (__ballot(ready && ticket >= nf)) acquire_fence();
      if (ready) {
        task = ticket;
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
    if (task < 0) {
      if (load_i(w.control + C_PENDING) == 0) break;
      if (__ballot(task >= 0) == 0) {
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
      walk.trail = 0;
      walk.bucket = 0;
      if (node > 0) {
        int first = node << (__clz(node) - __clz(tree.leaves));
        walk.start_at(tree, 1, first, w.bucket_log, lo, hi);
      }
    }
    bool done = false, cut = false;
    for (int step = 0; step < 32; ++step) {
      int leaf = 0;
      int c = walk.next(tree, lo, hi, leaf);
      if (c < 0) {
        node = 0;
        done = true;
        break;
      }
      node = successor(leaf);
      tests += 1;
      double2 a = w.cseg_a[c], b = w.cseg_b[c];
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
      double za = value(score, a) - w.cnear[c];
      double zb = value(score, b) - w.cnear[w.nseg + c];
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
      int fcap = w.cap.frontier;
      double2 *L = w.fr_xy + k, *R = w.fr_xy + k + 1;
      int left = 0, right = 0;
      double2 d = make_double2(b.x - a.x, b.y - a.y);
      for (int j = 0; j < n; ++j) {
        double2 u = s[j], v = s[(j + 1) % n];
        double du = -d.y * (u.x - a.x) + d.x * (u.y - a.y) + 0.0;
        double dv = -d.y * (v.x - a.x) + d.x * (v.y - a.y) + 0.0;
        if (du >= 0.0) left = append_slot(L, fcap, left, u);
        if (du <= 0.0) right = append_slot(R, fcap, right, u);
        if ((du > 0.0 && dv < 0.0) || (du < 0.0 && dv > 0.0)) {
          double t = du / (du - dv);
          double2 x = make_double2(u.x + t * (v.x - u.x), u.y + t * (v.y - u.y));
          left = append_slot(L, fcap, left, x);
          right = append_slot(R, fcap, right, x);
        }
      }
      left = clean_slot(L, fcap, left);
      right = clean_slot(R, fcap, right);
      if (left > MAXV || right > MAXV) {
        set_error(w.status, ERR_POLY);
        cut = true;
        break;
      }
      if (left > 0 && right > 0) {
        splits += 1;
      } else {
        // A collapsed child cancels the cut: carry the parent (frontier.py:200-206).
        for (int j = 0; j < n; ++j) L[(long)j * fcap] = s[j];
        left = n;
        right = 0;
      }
      w.fr_n[k] = left;
      w.fr_n[k + 1] = right;
      w.fr_src[k] = source;
      w.fr_src[k + 1] = source;
      w.fr_node[k] = node;
      w.fr_node[k + 1] = node;
      w.fr_gen[k] = gen + 1;
      w.fr_gen[k + 1] = gen + 1;
      // one release for both children, then the two ready flags
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
      __hip_atomic_store(w.fr_ready + k, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __hip_atomic_store(w.fr_ready + k + 1, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      int live = atomicAdd(w.control + C_PENDING, 1) + 1;
      atomicMax(w.status + ST_PEAK, live);
      cut = true;
      break;
    }
    bool finish = !cut && done;
    int o = wave_claim32(w.status + ST_POLYS, finish);
    if (cut) {
      task = -1;
    } else if (finish) {
      save_output(w, s, n, source, o);
      atomicSub(w.control + C_PENDING, 1);
      task = -1;
    } else {
      gen += 1;  // cursor continues in the next visit, like a re-enqueued patch
    }
  }
  // one atomic per wave for the counters
  wave_add_metric(&w.metrics[M_VISITS], visits);
  wave_add_metric(&w.metrics[M_CONTOUR_TESTS], tests);
  wave_add_metric(&w.metrics[M_SPLITS], splits);
  if (w.filter32) {
    wave_add_metric(&w.metrics[M_FILTER_STRADDLE], f_straddle);
    wave_add_metric(&w.metrics[M_FALLBACK_STRADDLE], b_straddle);
    wave_add_metric(&w.metrics[M_FILTER_FRONT], f_front);
    wave_add_metric(&w.metrics[M_FALLBACK_FRONT], b_front);
  }
  max_gen = wave_max(max_gen);
  if ((threadIdx.x & 63) == 0)
    atomicMax((unsigned long long *)&w.metrics[M_ROUNDS], (unsigned long long)(max_gen + 1));
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
