// LSD radix sort, exclusive scan and the Morton hierarchy for gfx950.
//
// Radix sort: 8-bit digits, 256-thread blocks (four wave64s) over 4096-key
// tiles. A pass is three kernels: per-tile digit histograms in LDS; a scan
// with one block per digit that turns its row of tile counts into offsets
// within the digit and records the digit's total; and a stable scatter that
// adds the exclusive scan of the 256 digit totals (computed in LDS by every
// block) and ranks keys inside the tile with wave64 ballots (eight ballots
// give each lane the mask of lanes holding the same digit) plus per-wave LDS
// digit counters. A pass is trivial when one digit holds every key; the
// scatter then does nothing and the ping-pong parity is tracked on the device,
// so leading or trailing constant digits cost one histogram pass and no data
// movement, without any host read.
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

// parity of the non-trivial passes before `pass`
__device__ __forceinline__ int current_buffer(const int *flags, int pass) {
  int sel = 0;
  for (int p = 0; p < pass; ++p) sel ^= flags[p];
  return sel;
}

__global__ __launch_bounds__(RS_THREADS) void rs_hist(const uint64_t *k0, const uint64_t *k1,
                                                       const int *count, int count_static,
                                                       int capacity, int shift, int pass,
                                                       int *flags, int *hist, int max_blocks) {
  __shared__ int bins[256];
  int n = live_count(count, count_static, capacity);
  int begin = blockIdx.x * RS_TILE;
  // non-trivial until the scan finds one digit holding every key
  if (blockIdx.x == 0 && threadIdx.x == 0) flags[pass] = n > 0 ? 1 : 0;
  if (begin >= n) return;
  const uint64_t *keys = current_buffer(flags, pass) ? k1 : k0;
  bins[threadIdx.x] = 0;
  __syncthreads();
  int end = min(n, begin + RS_TILE);
  for (int i = begin + threadIdx.x; i < end; i += RS_THREADS)
    atomicAdd(&bins[(int)((keys[i] >> shift) & 255)], 1);
  __syncthreads();
  hist[threadIdx.x * max_blocks + blockIdx.x] = bins[threadIdx.x];
}

// One block per digit: exclusive scan of the digit's tile counts (coalesced
// chunks of 256) into offsets within the digit, and the digit's total.
__global__ __launch_bounds__(RS_THREADS) void rs_scan(const int *count, int count_static, int capacity,
                                                       int pass, int *flags, int *hist, int *totals,
                                                       int max_blocks) {
  __shared__ int lds[4];
  int n = live_count(count, count_static, capacity);
  int nb = (n + RS_TILE - 1) / RS_TILE;
  int digit = blockIdx.x;
  int *row = hist + (long)digit * max_blocks;
  int running = 0;
  for (int c = 0; c < nb; c += RS_THREADS) {
    int b = c + threadIdx.x;
    int v = b < nb ? row[b] : 0;
    int total;
    int excl = block_exclusive(v, lds, total);
    if (b < nb) row[b] = running + excl;
    running += total;
  }
  if (threadIdx.x == 0) {
    totals[digit] = running;
    if (n > 0 && running == n) flags[pass] = 0;
  }
}

__global__ __launch_bounds__(RS_THREADS) void rs_scatter(const uint64_t *k0, const int *v0,
                                                          uint64_t *k1, int *v1,
                                                          const int *count, int count_static,
                                                          int capacity, int shift, int pass,
                                                          const int *flags, const int *hist,
                                                          const int *totals, int max_blocks) {
  __shared__ int wave_count[4][256];
  __shared__ int digit_base[256];
  __shared__ int lds[4];
  int n = live_count(count, count_static, capacity);
  int tile = blockIdx.x * RS_TILE;
  if (tile >= n || flags[pass] == 0) return;
  int sel = current_buffer(flags, pass);
  const uint64_t *kin = sel ? k1 : k0;
  const int *vin = sel ? v1 : v0;
  uint64_t *kout = sel ? (uint64_t *)k0 : k1;
  int *vout = sel ? (int *)v0 : v1;
  int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < 4 * 256; i += RS_THREADS) (&wave_count[0][0])[i] = 0;
  int all;
  int digit_offset = block_exclusive(totals[threadIdx.x], lds, all);
  digit_base[threadIdx.x] = digit_offset + hist[threadIdx.x * max_blocks + blockIdx.x];
  __syncthreads();
  volatile int *counters = wave_count[wave];
  int base = tile + wave * 64 * RS_ITEMS;
  uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  uint64_t key[RS_ITEMS];
  int val[RS_ITEMS], rank[RS_ITEMS], dig[RS_ITEMS];
#pragma unroll
  for (int j = 0; j < RS_ITEMS; ++j) {
    int idx = base + j * 64 + lane;
    bool valid = idx < n;
    key[j] = valid ? kin[idx] : 0ull;
    val[j] = valid ? vin[idx] : 0;
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
  {
    int d = threadIdx.x, running = 0;
    for (int w = 0; w < 4; ++w) {
      int c = wave_count[w][d];
      wave_count[w][d] = running;
      running += c;
    }
  }
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

__global__ void rs_finish(uint64_t *k0, int *v0, const uint64_t *k1, const int *v1,
                          const int *count, int count_static, int capacity, int passes,
                          const int *flags) {
  int n = live_count(count, count_static, capacity);
  if (!current_buffer(flags, passes)) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    k0[i] = k1[i];
    v0[i] = v1[i];
  }
}

void radix_sort(hipStream_t s, uint64_t *keys, int *ids, uint64_t *alt_keys, int *alt_ids,
                const int *count, int count_static, int capacity, int bit_begin, int bit_end,
                const SortScratch &scratch) {
  int blocks = (capacity + RS_TILE - 1) / RS_TILE;
  if (blocks < 1 || bit_end <= bit_begin) return;
  int first = bit_begin / 8, last = (bit_end + 7) / 8;
  int passes = 0;
  for (int digit = first; digit < last; ++digit, ++passes) {
    int shift = digit * 8;
    hipLaunchKernelGGL(rs_hist, dim3(blocks), dim3(RS_THREADS), 0, s, keys, alt_keys, count,
                       count_static, capacity, shift, passes, scratch.flags, scratch.hist,
                       scratch.max_blocks);
    hipLaunchKernelGGL(rs_scan, dim3(256), dim3(RS_THREADS), 0, s, count, count_static, capacity, passes,
                       scratch.flags, scratch.hist, scratch.totals, scratch.max_blocks);
    hipLaunchKernelGGL(rs_scatter, dim3(blocks), dim3(RS_THREADS), 0, s, keys, ids, alt_keys,
                       alt_ids, count, count_static, capacity, shift, passes, scratch.flags,
                       scratch.hist, scratch.totals, scratch.max_blocks);
  }
  int grid = blocks * 4 < 4096 ? blocks * 4 : 4096;
  hipLaunchKernelGGL(rs_finish, dim3(grid), dim3(256), 0, s, keys, ids, alt_keys, alt_ids, count,
                     count_static, capacity, passes, scratch.flags);
}

// ---------------------------------------------------------------------------
// Exclusive scan: tile sums, one-block scan of tile sums, tile scans.
constexpr int SC_THREADS = 256, SC_ITEMS = 16, SC_TILE = SC_THREADS * SC_ITEMS;


__global__ __launch_bounds__(SC_THREADS) void sc_reduce(const int *in, const int *count,
                                                         int count_static, int capacity,
                                                         int *block_sums) {
  __shared__ int lds[4];
  int n = live_count(count, count_static, capacity);
  int begin = blockIdx.x * SC_TILE;
  if (begin >= n) return;
  int local = 0;
  for (int i = begin + threadIdx.x; i < min(n, begin + SC_TILE); i += SC_THREADS) local += in[i];
  int total;
  block_exclusive(local, lds, total);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

__global__ __launch_bounds__(1024) void sc_top(const int *count, int count_static, int capacity,
                                                int *block_sums, int *total_out) {
  __shared__ int sums[1024];
  int n = live_count(count, count_static, capacity);
  int nb = (n + SC_TILE - 1) / SC_TILE;
  int t = threadIdx.x, per = (nb + 1023) / 1024;
  int b0 = min(nb, t * per), b1 = min(nb, b0 + per);
  int local = 0;
  for (int b = b0; b < b1; ++b) local += block_sums[b];
  sums[t] = local;
  for (int off = 1; off < 1024; off <<= 1) {
    __syncthreads();
    int v = t >= off ? sums[t - off] : 0;
    __syncthreads();
    sums[t] += v;
  }
  __syncthreads();
  int running = sums[t] - local;
  for (int b = b0; b < b1; ++b) {
    int c = block_sums[b];
    block_sums[b] = running;
    running += c;
  }
  if (t == 1023 && total_out) *total_out = sums[1023];
}

__global__ __launch_bounds__(SC_THREADS) void sc_apply(const int *in, int *out, const int *count,
                                                        int count_static, int capacity,
                                                        const int *block_sums) {
  __shared__ int lds[4];
  int n = live_count(count, count_static, capacity);
  int begin = blockIdx.x * SC_TILE;
  if (begin >= n) return;
  // each thread scans SC_ITEMS consecutive elements
  int first = begin + threadIdx.x * SC_ITEMS;
  int vals[SC_ITEMS];
  int local = 0;
#pragma unroll
  for (int j = 0; j < SC_ITEMS; ++j) {
    int i = first + j;
    vals[j] = i < n ? in[i] : 0;
    local += vals[j];
  }
  int total;
  int running = block_exclusive(local, lds, total) + block_sums[blockIdx.x];
#pragma unroll
  for (int j = 0; j < SC_ITEMS; ++j) {
    int i = first + j;
    if (i < n) out[i] = running;
    running += vals[j];
  }
}

void scan_exclusive(hipStream_t s, const int *in, int *out, const int *count, int count_static,
                    int capacity, int *total, const ScanScratch &scratch) {
  int blocks = (capacity + SC_TILE - 1) / SC_TILE;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(sc_reduce, dim3(blocks), dim3(SC_THREADS), 0, s, in, count, count_static,
                     capacity, scratch.block_sums);
  hipLaunchKernelGGL(sc_top, dim3(1), dim3(1024), 0, s, count, count_static, capacity,
                     scratch.block_sums, total);
  hipLaunchKernelGGL(sc_apply, dim3(blocks), dim3(SC_THREADS), 0, s, in, out, count,
                     count_static, capacity, scratch.block_sums);
}

// ---------------------------------------------------------------------------
// Morton hierarchy (warp_peel/kernels.py:67-116, solver.py:34-89).
__global__ void h_size(const int *count, int count_static, int capacity, int *leaves) {
  int n = live_count(count, count_static, capacity);
  int s = 1;
  while (s < n) s <<= 1;
  *leaves = s;
}

__global__ void h_keys(const Box *boxes, const int *count, int count_static, int capacity,
                       int layered, uint64_t *keys, int *ids) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int n = live_count(count, count_static, capacity);
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

__global__ void h_leaves(const Box *boxes, const int *count, int count_static, int capacity,
                         const int *ids, const int *leaves, Box *nodes) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  int size = *leaves;
  if (i >= size) return;
  int n = live_count(count, count_static, capacity);
  Box b;
  b.lo = make_float4(1e30f, 1e30f, 1e30f, 0.f);
  b.hi = make_float4(-1e30f, -1e30f, -1e30f, 0.f);
  if (i < n) b = boxes[ids[i]];
  nodes[size + i] = b;
}

__device__ __forceinline__ Box unite(const Box &a, const Box &b) {
  Box r;
  r.lo = make_float4(fminf(a.lo.x, b.lo.x), fminf(a.lo.y, b.lo.y), fminf(a.lo.z, b.lo.z), 0.f);
  r.hi = make_float4(fmaxf(a.hi.x, b.hi.x), fmaxf(a.hi.y, b.hi.y), fmaxf(a.hi.z, b.hi.z), 0.f);
  return r;
}

// Launch `stage` builds levels [9 stage, 9 stage + 9) above the leaves: block b
// reduces up to 512 nodes of that bottom level inside LDS.
__global__ __launch_bounds__(256) void h_levels(Box *nodes, const int *leaves, int stage) {
  __shared__ Box lds[512];
  int bottom = *leaves >> (9 * stage);
  if (bottom <= 1) return;
  int levels = 0;
  while ((1 << (levels + 1)) <= bottom && levels < 9) ++levels;
  int width = 1 << levels;
  int first = bottom + blockIdx.x * width;
  if (blockIdx.x * width >= bottom) return;
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
                     int capacity, bool layered, Hierarchy &h, const SortScratch &scratch) {
  int cap = capacity < h.size ? capacity : h.size;
  hipLaunchKernelGGL(h_size, dim3(1), dim3(1), 0, s, count, count_static, cap, h.leaves);
  int grid = (cap + 255) / 256;
  hipLaunchKernelGGL(h_keys, dim3(grid < 1 ? 1 : grid), dim3(256), 0, s, boxes, count, count_static,
                     cap, (int)layered, h.keys, h.ids);
  radix_sort(s, h.keys, h.ids, h.alt_keys, h.alt_ids, count, count_static, cap, 0, layered ? 64 : 32,
             scratch);
  hipLaunchKernelGGL(h_leaves, dim3((h.size + 255) / 256), dim3(256), 0, s, boxes, count,
                     count_static, cap, h.ids, h.leaves, h.nodes);
  int stage = 0;
  for (long bottom = h.size; bottom > 1; bottom >>= 9, ++stage) {
    long blocks = bottom / 512;
    hipLaunchKernelGGL(h_levels, dim3((unsigned)(blocks < 1 ? 1 : blocks)), dim3(256), 0, s, h.nodes,
                       h.leaves, stage);
  }
}

}  // namespace peel
