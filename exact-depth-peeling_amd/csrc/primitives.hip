// LSD radix sort, exclusive scan and the Morton hierarchy for gfx950.
//
// Radix sort: 8-bit digits, 256-thread blocks (four wave64s) over 4096-key
// tiles, one kernel per pass with decoupled look-back (Merrill and Garland,
// "Single-pass Parallel Prefix Scan with Decoupled Look-back", 2016; used per
// radix pass as in Adinets and Merrill, "Onesweep", 2022):
//   rs_count  one read of the keys counts the digits of every pass;
//   rs_sweep  a block takes the next tile ticket, ranks its keys by digit with
//             wave64 ballots (eight ballots give each lane the mask of lanes
//             holding the same digit) and per-wave LDS counters, publishes its
//             per-digit counts, adds the counts of the tiles before it by
//             walking back to the nearest published prefix, publishes its own
//             prefix and scatters stably;
//   rs_done   copies back after an odd number of moving passes and clears the
//             counters, tickets and look-back words for the next sort (the
//             hierarchy builds do this in h_leaves instead: rs_finish).
// Tickets are taken in execution order, so a tile waits only on tiles whose
// blocks are already running and publish their counts before they wait
// themselves; every wait is bounded (ERR_WAIT). A pass is trivial when one
// digit holds every key: it moves nothing and the ping-pong parity follows
// the non-trivial passes on the device, so leading or trailing constant
// digits cost no data movement and no host read. Look-back words pack a
// 2-bit state (0 none, 1 tile count, 2 inclusive prefix) over a 30-bit count.
#include "primitives.hpp"

namespace peel {

int pow2_at_least(int n) {
  int s = 1;
  while (s < n) s <<= 1;
  return s;
}

__device__ __forceinline__ int live_count(const int *count, int count_static, int capacity) {
  int n = count ? load_i(count) : count_static;
  return n < 0 ? 0 : (n > capacity ? capacity : n);
}

__device__ __forceinline__ int block_exclusive(int v, int *lds, int &total) {
  // 256-thread exclusive scan through wave64 shuffles + LDS for the 4 wave sums
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  int x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    int y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) lds[wave] = x;
  __syncthreads();
  int prefix = 0;
  for (int w = 0; w < wave; ++w) prefix += lds[w];
  total = lds[0] + lds[1] + lds[2] + lds[3];
  __syncthreads();
  return prefix + x - v;
}

constexpr unsigned LB_COUNT = 1u << 30, LB_PREFIX = 2u << 30, LB_MASK = (1u << 30) - 1;
constexpr int LB_SPIN_LIMIT = 1 << 20;
constexpr int LB_BATCH = 8;   // predecessor words loaded together in the look-back
constexpr int RC_TILE = 1024;  // keys per block of the digit count (independent of the sweep's tiles)

__global__ __launch_bounds__(RS_THREADS) void rs_count(const uint64_t *keys, const int *count, int count_static,
                                                        int capacity, int first_shift, int passes, int *counts) {
  __shared__ int bins[RS_MAX_SWEEPS][256];
  int n = live_count(count, count_static, capacity);
  // chunks of RC_TILE keys, grid-strided over at most one block per CU
  if ((long)blockIdx.x * RC_TILE >= n) return;
  for (int i = threadIdx.x; i < RS_MAX_SWEEPS * 256; i += RS_THREADS) (&bins[0][0])[i] = 0;
  __syncthreads();
  for (long begin = (long)blockIdx.x * RC_TILE; begin < n; begin += (long)gridDim.x * RC_TILE) {
    long end = min((long)n, begin + RC_TILE);
    for (long i = begin + threadIdx.x; i < end; i += RS_THREADS) {
      uint64_t k = keys[i];
      for (int p = 0; p < passes; ++p) atomicAdd(&bins[p][(int)((k >> (first_shift + 8 * p)) & 255)], 1);
    }
  }
  __syncthreads();
  for (int p = 0; p < passes; ++p)
    if (bins[p][threadIdx.x]) atomicAdd(&counts[p * 256 + threadIdx.x], bins[p][threadIdx.x]);
}

__global__ __launch_bounds__(RS_THREADS) void rs_sweep(const uint64_t *k0, const int *v0, uint64_t *k1, int *v1,
                                                        const int *count, int count_static, int capacity,
                                                        int shift, int pass, const int *counts, int *tickets,
                                                        unsigned *tiles, int max_blocks, int *flags,
                                                        int *status) {
  __shared__ int wave_count[4][256];
  __shared__ int digit_base[256];
  __shared__ int lds[4];
  __shared__ int trivial, tile_shared;
  int n = live_count(count, count_static, capacity);
  // this digit's count in every pass so far, loaded together
  int digit_counts[RS_MAX_SWEEPS];
#pragma unroll
  for (int q = 0; q < RS_MAX_SWEEPS; ++q) digit_counts[q] = q <= pass ? counts[q * 256 + threadIdx.x] : 0;
  // parity of the moving passes before this one, and whether this one moves
  int sel = 0, moves = 0;
#pragma unroll
  for (int q = 0; q < RS_MAX_SWEEPS; ++q) {
    if (q > pass) break;
    if (threadIdx.x == 0) trivial = 0;
    __syncthreads();
    if (digit_counts[q] == n) trivial = 1;
    __syncthreads();
    int m = n > 0 && !trivial;
    __syncthreads();
    if (q < pass) sel ^= m;
    else moves = m;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) flags[pass] = moves;
  if (!moves) return;
  // only as many blocks as there are tiles take a ticket (blocks 0 .. tiles-1
  // take tickets 0 .. tiles-1, in execution order); the others leave at once
  if ((long)blockIdx.x * RS_TILE >= n) return;
  if (threadIdx.x == 0) tile_shared = atomicAdd(tickets + pass, 1);
  __syncthreads();
  const int tile = tile_shared;
  const int begin = tile * RS_TILE;
  if (begin >= n) return;
  const uint64_t *kin = sel ? k1 : k0;
  const int *vin = sel ? v1 : v0;
  uint64_t *kout = sel ? (uint64_t *)k0 : k1;
  int *vout = sel ? (int *)v0 : v1;
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < 4 * 256; i += RS_THREADS) (&wave_count[0][0])[i] = 0;
  __syncthreads();
  volatile int *counters = wave_count[wave];
  int base = begin + wave * 64 * RS_ITEMS;
  uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint64_t key[RS_ITEMS];
  int val[RS_ITEMS], rank[RS_ITEMS], dig[RS_ITEMS];
  // every load of the tile first, so that they are all in flight together
  // (the ranking below orders itself through the LDS counters)
#pragma unroll
  for (int j = 0; j < RS_ITEMS; ++j) {
    int idx = base + j * 64 + lane;
    bool valid = idx < n;
    key[j] = valid ? kin[idx] : 0ull;
    val[j] = valid ? vin[idx] : 0;
  }
#pragma unroll
  for (int j = 0; j < RS_ITEMS; ++j) {
    int idx = base + j * 64 + lane;
    bool valid = idx < n;
    int d = valid ? (int)((key[j] >> shift) & 255) : 0;
    dig[j] = d;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      uint64_t m = __ballot((d >> b) & 1);
      peers &= ((d >> b) & 1) ? m : ~m;
    }
    int below = __popcll(peers & lt);
    int total = __popcll(peers);
    int c = valid ? counters[d] : 0;
    __builtin_amdgcn_wave_barrier();
    if (valid && below == 0) counters[d] = c + total;
    __builtin_amdgcn_wave_barrier();
    rank[j] = c + below;
  }
  __syncthreads();
  // this tile's count of each digit; per-wave offsets inside the tile
  const int d = threadIdx.x;
  int tile_count = 0;
  for (int w = 0; w < 4; ++w) {
    int c = wave_count[w][d];
    wave_count[w][d] = tile_count;
    tile_count += c;
  }
  unsigned *word = tiles + ((long)pass * max_blocks + tile) * 256 + d;
  int before = 0;
  if (tile == 0) {
    __hip_atomic_store(word, LB_PREFIX | (unsigned)tile_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  } else {
    __hip_atomic_store(word, LB_COUNT | (unsigned)tile_count, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    // The words of up to LB_BATCH predecessors are loaded together and
    // consumed nearest first: a prefix ends the walk, a count is added and
    // the walk goes on, an unpublished word is loaded again (with the rest
    // of its batch). The sum is the one a word-by-word walk finds, with one
    // load latency per batch instead of per predecessor.
    int spins = 0;
    const unsigned *column = tiles + (long)pass * max_blocks * 256 + d;
    for (int t = tile - 1; t >= 0;) {
      unsigned got[LB_BATCH];
#pragma unroll
      for (int j = 0; j < LB_BATCH; ++j)
        got[j] = t - j >= 0 ? __hip_atomic_load(column + (long)(t - j) * 256, __ATOMIC_RELAXED,
                                                __HIP_MEMORY_SCOPE_AGENT)
                            : LB_PREFIX;
      bool done = false, waiting = false;
#pragma unroll
      for (int j = 0; j < LB_BATCH; ++j) {
        if (done || waiting) continue;
        unsigned state = got[j] & ~LB_MASK;
        if (state == 0) {
          waiting = true;
          continue;
        }
        if (t < 0) {  // past tile 0 (only reached after tile 0's prefix)
          done = true;
          continue;
        }
        before += (int)(got[j] & LB_MASK);
        if (state == LB_PREFIX) done = true;
        --t;
      }
      if (done) break;
      if (waiting) {
        if (++spins > LB_SPIN_LIMIT) {
          set_error(status, ERR_WAIT);
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    __hip_atomic_store(word, LB_PREFIX | (unsigned)(before + tile_count), __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_AGENT);
  }
  int all;
  int this_pass = 0;
#pragma unroll
  for (int q = 0; q < RS_MAX_SWEEPS; ++q)
    if (q == pass) this_pass = digit_counts[q];
  int digit_start = block_exclusive(this_pass, lds, all);
  digit_base[d] = digit_start + before;
  __syncthreads();
#pragma unroll
  for (int j = 0; j < RS_ITEMS; ++j) {
    int idx = base + j * 64 + lane;
    if (idx < n) {
      int pos = digit_base[dig[j]] + wave_count[wave][dig[j]] + rank[j];
      kout[pos] = key[j];
      vout[pos] = val[j];
    }
  }
}

// The end of a sort, for the grid thread `first` of `stride`: the ids (and,
// with `copy_keys`, the keys) of the items it owns move back after an odd
// number of moving passes, and the look-back words, digit counts and tickets
// are left zero for the next sort.
__device__ void rs_finish(uint64_t *k0, int *v0, const uint64_t *k1, const int *v1, int n, int passes,
                          const int *flags, int *counts, int *tickets, unsigned *tiles, int max_blocks,
                          bool copy_keys, long first, long stride) {
  int sel = 0;
  for (int p = 0; p < passes; ++p) sel ^= flags[p];
  if (sel)
    for (long i = first; i < n; i += stride) {
      if (copy_keys) k0[i] = k1[i];
      v0[i] = v1[i];
    }
  const long used = (long)((n + RS_TILE - 1) / RS_TILE) * 256;
  for (int p = 0; p < passes; ++p)
    for (long i = first; i < used; i += stride) tiles[(long)p * max_blocks * 256 + i] = 0u;
  for (long i = first; i < (long)passes * 256; i += stride) counts[i] = 0;
  if (first < passes) tickets[first] = 0;
}

__global__ void rs_done(uint64_t *k0, int *v0, const uint64_t *k1, const int *v1, const int *count,
                        int count_static, int capacity, int passes, const int *flags, int *counts, int *tickets,
                        unsigned *tiles, int max_blocks) {
  rs_finish(k0, v0, k1, v1, live_count(count, count_static, capacity), passes, flags, counts, tickets, tiles,
            max_blocks, true, (long)blockIdx.x * blockDim.x + threadIdx.x, (long)gridDim.x * blockDim.x);
}

void radix_sort(hipStream_t s, uint64_t *keys, int *ids, uint64_t *alt_keys, int *alt_ids,
                const int *count, int count_static, int capacity, int bit_begin, int bit_end,
                const SortScratch &scratch, bool finish) {
  int blocks = (capacity + RS_TILE - 1) / RS_TILE;
  if (blocks < 1 || bit_end <= bit_begin) return;
  int first = bit_begin / 8, last = (bit_end + 7) / 8;
  int passes = last - first;  // at most 8 for 64-bit keys
  int count_blocks = (capacity + RC_TILE - 1) / RC_TILE;
  if (count_blocks > 256) count_blocks = 256;
  hipLaunchKernelGGL(rs_count, dim3(count_blocks), dim3(RS_THREADS), 0, s, keys, count, count_static, capacity,
                     first * 8, passes, scratch.counts);
  for (int pass = 0; pass < passes; ++pass)
    hipLaunchKernelGGL(rs_sweep, dim3(blocks), dim3(RS_THREADS), 0, s, keys, ids, alt_keys, alt_ids, count,
                       count_static, capacity, (first + pass) * 8, pass, scratch.counts, scratch.tickets,
                       scratch.tiles, scratch.max_blocks, scratch.flags, scratch.status);
  if (!finish) return;  // the caller's next kernel runs rs_finish
  int grid = blocks * 4 < 4096 ? blocks * 4 : 4096;
  hipLaunchKernelGGL(rs_done, dim3(grid), dim3(256), 0, s, keys, ids, alt_keys, alt_ids, count, count_static,
                     capacity, passes, scratch.flags, scratch.counts, scratch.tickets, scratch.tiles,
                     scratch.max_blocks);
}

// ---------------------------------------------------------------------------
// Exclusive scan in one launch with decoupled look-back, as the radix passes:
// only blocks with a live tile take a ticket (in execution order), each tile
// publishes its total, the first wave walks back over up to 64 predecessors'
// words at a time to the nearest published prefix, and the tile publishes its
// own prefix before it writes. The last tile to finish clears the words and
// counters for the next scan. Words pack a 2-bit state (0 none, 1 tile total,
// 2 inclusive prefix) over a 62-bit value.
constexpr int SC_THREADS = 256, SC_ITEMS = 16, SC_TILE = SC_THREADS * SC_ITEMS;
constexpr unsigned long long SC_TOTAL = 1ull << 62, SC_PREFIX = 2ull << 62, SC_VALUE = (1ull << 62) - 1;

__global__ __launch_bounds__(SC_THREADS) void sc_onepass(const int *in, int *out, const int *count,
                                                          int count_static, int capacity, int *total_out,
                                                          unsigned long long *words, int *counters,
                                                          int *status) {
  __shared__ int lds[4];
  __shared__ int tile_shared, last_shared;
  __shared__ long long before_shared;
  int n = live_count(count, count_static, capacity);
  if (n == 0) {
    if (blockIdx.x == 0 && threadIdx.x == 0 && total_out) *total_out = 0;
    return;
  }
  const int tiles = (n + SC_TILE - 1) / SC_TILE;
  if ((int)blockIdx.x >= tiles) return;
  if (threadIdx.x == 0) tile_shared = atomicAdd(counters, 1);
  __syncthreads();
  const int tile = tile_shared;
  const int first = tile * SC_TILE + threadIdx.x * SC_ITEMS;
  int vals[SC_ITEMS];
  int local = 0;
#pragma unroll
  for (int j = 0; j < SC_ITEMS; ++j) {
    int i = first + j;
    vals[j] = i < n ? in[i] : 0;
    local += vals[j];
  }
  int total;
  int running = block_exclusive(local, lds, total);
  if (threadIdx.x < 64) {
    const int lane = threadIdx.x;
    long long before = 0;
    if (tile == 0) {
      if (lane == 0)
        __hip_atomic_store(words, SC_PREFIX | (unsigned long long)total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      if (lane == 0)
        __hip_atomic_store(words + tile, SC_TOTAL | (unsigned long long)total, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
      int t = tile - 1, spins = 0;  // nearest predecessor not yet added
      while (true) {
        int j = t - lane;
        unsigned long long word =
            j >= 0 ? __hip_atomic_load(words + j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : SC_PREFIX;
        uint64_t pending = __ballot((word >> 62) == 0);
        uint64_t prefix = __ballot((word >> 62) == 2);
        int stop = prefix ? __ffsll((unsigned long long)prefix) - 1 : 63;      // last lane this window needs
        int gap = pending ? __ffsll((unsigned long long)pending) - 1 : 64;     // first unpublished lane
        int take = gap <= stop ? gap : stop + 1;                                // lanes [0, take) are added
        before += wave_sum(lane < take ? (long long)(word & SC_VALUE) : 0ll);
        if (gap > stop && prefix) break;
        t -= take;
        if (gap <= stop) {
          if (++spins > LB_SPIN_LIMIT) {
            if (lane == 0) set_error(status, ERR_WAIT);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
        }
      }
      if (lane == 0)
        __hip_atomic_store(words + tile, SC_PREFIX | (unsigned long long)(before + total), __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
    }
    if (lane == 0) before_shared = before;
  }
  __syncthreads();
  const int offset = (int)before_shared;
  if (threadIdx.x == 0 && total_out && tile == tiles - 1) *total_out = offset + total;
  running += offset;
#pragma unroll
  for (int j = 0; j < SC_ITEMS; ++j) {
    int i = first + j;
    if (i < n) out[i] = running;
    running += vals[j];
  }
  // the last tile to finish leaves the words and counters zero
  if (threadIdx.x == 0) last_shared = atomicAdd(counters + 1, 1) == tiles - 1;
  __syncthreads();
  if (last_shared) {
    for (int j = threadIdx.x; j < tiles; j += SC_THREADS) words[j] = 0ull;
    if (threadIdx.x == 0) {
      counters[0] = 0;
      counters[1] = 0;
    }
  }
}

void scan_exclusive(hipStream_t s, const int *in, int *out, const int *count, int count_static,
                    int capacity, int *total, const ScanScratch &scratch) {
  int blocks = (capacity + SC_TILE - 1) / SC_TILE;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(sc_onepass, dim3(blocks), dim3(SC_THREADS), 0, s, in, out, count, count_static, capacity,
                     total, scratch.words, scratch.counters, scratch.status);
}

// ---------------------------------------------------------------------------
// Morton hierarchy (warp_peel/kernels.py:67-116, solver.py:34-89).
// Morton keys of the live boxes; thread 0 also sizes the tree: the power of two
// at least the live count. (Counting the sort's digits here, on at most 256
// blocks as rs_count does, saved its launch on the 65,536-face torus knot but
// cost the large hierarchies more: the keys' producer then walks its boxes in
// long per-block loops instead of one box per thread.)
__global__ void h_keys(const Box *boxes, const int *count, int count_static, int capacity,
                       int layered, uint64_t *keys, int *ids, int *leaves) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int n = live_count(count, count_static, capacity);
  if (i == 0) {
    int s = 1;
    while (s < n) s <<= 1;
    *leaves = s;
  }
  if (i >= n) return;
  ids[i] = i;
  uint64_t key = ~0ull;
  Box b = boxes[i];
  if (b.lo.x <= b.hi.x && b.lo.y <= b.hi.y) {
    float cx = (b.lo.x + b.hi.x) * 0.5f, cy = (b.lo.y + b.hi.y) * 0.5f;
    cx = fminf(fmaxf(cx, 0.0f), 1.0f);
    cy = fminf(fmaxf(cy, 0.0f), 1.0f);
    uint32_t x = (uint32_t)(int)(cx * 32767.0f), y = (uint32_t)(int)(cy * 32767.0f);
    key = 0;
    for (int bit = 0; bit < 15; ++bit) {
      key |= (uint64_t)((x >> bit) & 1u) << (2 * bit);
      key |= (uint64_t)((y >> bit) & 1u) << (2 * bit + 1);
    }
    if (layered) key |= (uint64_t)(uint32_t)(int)b.lo.z << 32;
  }
  keys[i] = key;
}

// The leaves in Morton order. Each thread first finishes the sort for the
// item it owns (the grid covers the tree's leaves, at least the live count)
// and then reads that item's id.
__global__ void h_leaves(const Box *boxes, const int *count, int count_static, int capacity,
                         int *ids, const int *alt_ids, int passes, SortScratch sort, const int *leaves,
                         Box *nodes, LeafGather g) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int n = live_count(count, count_static, capacity);
  // the items, look-back words and counts end below max(n, passes * 256)
  if (i < n || i < passes * 256)
    rs_finish(nullptr, ids, nullptr, alt_ids, n, passes, sort.flags, sort.counts, sort.tickets, sort.tiles,
              sort.max_blocks, false, i, (long)gridDim.x * blockDim.x);
  int size = *leaves;
  if (i >= size) return;
  Box b;
  b.lo = make_float4(1e30f, 1e30f, 1e30f, 0.f);
  b.hi = make_float4(-1e30f, -1e30f, -1e30f, 0.f);
  if (i < n) {
    int id = ids[i];
    b = boxes[id];
    if (g.out_a) {
      g.out_a[i] = g.a[id];
      g.out_b[i] = g.b[id];
      g.out_near[i] = make_double2(g.near[id], g.near[g.near_stride + id]);
    }
  }
  nodes[size + i] = b;
}

__device__ __forceinline__ Box unite(const Box &a, const Box &b) {
  Box r;
  r.lo = make_float4(fminf(a.lo.x, b.lo.x), fminf(a.lo.y, b.lo.y), fminf(a.lo.z, b.lo.z), 0.f);
  r.hi = make_float4(fmaxf(a.hi.x, b.hi.x), fmaxf(a.hi.y, b.hi.y), fmaxf(a.hi.z, b.hi.z), 0.f);
  return r;
}

// The internal levels, nine per launch: each block reduces a group of up to
// 512 nodes of the stage's bottom level in LDS, nine levels at a time, and the
// next stage starts from the nodes this one wrote. Stages are separate
// launches (in-kernel hand-offs between blocks would need agent-scope fences,
// which slow every wave's loads); a stage above the root returns at once.
__global__ __launch_bounds__(256) void h_levels(Box *nodes, const int *leaves, int stage) {
  __shared__ Box lds[HIER_GROUP];
  const int bottom = *leaves >> (HIER_STAGE_LEVELS * stage);
  const int group = blockIdx.x;
  const int width = bottom < HIER_GROUP ? bottom : HIER_GROUP;
  if (bottom <= 1 || group * width >= bottom) return;
  int levels = 0;
  while ((1 << (levels + 1)) <= width) ++levels;
  const int first = bottom + group * width;
  for (int i = threadIdx.x; i < width; i += 256) lds[i] = nodes[first + i];
  __syncthreads();
  int count = width, start = first;
  for (int l = 0; l < levels; ++l) {
    int half = count >> 1;
    Box parent;
    if ((int)threadIdx.x < half) parent = unite(lds[2 * threadIdx.x], lds[2 * threadIdx.x + 1]);
    __syncthreads();
    start >>= 1;
    if ((int)threadIdx.x < half) {
      lds[threadIdx.x] = parent;
      nodes[start + threadIdx.x] = parent;
    }
    __syncthreads();
    count = half;
  }
}

void build_hierarchy(hipStream_t s, const Box *boxes, const int *count, int count_static,
                     int capacity, bool layered, Hierarchy &h, const SortScratch &scratch,
                     const LeafGather &gather) {
  int cap = capacity < h.size ? capacity : h.size;
  const int bits = layered ? 64 : 32;
  int grid = (cap + 255) / 256;
  hipLaunchKernelGGL(h_keys, dim3(grid < 1 ? 1 : grid), dim3(256), 0, s, boxes, count, count_static,
                     cap, (int)layered, h.keys, h.ids, h.leaves);
  radix_sort(s, h.keys, h.ids, h.alt_keys, h.alt_ids, count, count_static, cap, 0, bits, scratch, false);
  hipLaunchKernelGGL(h_leaves, dim3((h.size + 255) / 256), dim3(256), 0, s, boxes, count, count_static, cap,
                     h.ids, h.alt_ids, (bits + 7) / 8, scratch, h.leaves, h.nodes, gather);
  for (int stage = 0; (1L << (HIER_STAGE_LEVELS * stage)) < h.size; ++stage) {
    long blocks = (h.size >> (HIER_STAGE_LEVELS * stage)) / HIER_GROUP;
    hipLaunchKernelGGL(h_levels, dim3((unsigned)(blocks < 1 ? 1 : blocks)), dim3(256), 0, s, h.nodes, h.leaves,
                       stage);
  }
}

}  // namespace peel
