// Device-wide primitives: LSD radix sort, exclusive scan, Morton hierarchy.
// All of them take their element count from device memory, so they can be
// recorded once into a graph and replayed with changing counts.
#pragma once
#include "peel_common.hpp"

namespace peel {

constexpr int RS_THREADS = 256;
constexpr int RS_ITEMS = 8;
constexpr int RS_TILE = RS_THREADS * RS_ITEMS;  // 2048 keys per block tile
constexpr int RS_MAX_SWEEPS = 8;  // 8-bit digits of a 64-bit key

// Zero between sorts: allocation clears it and every sort's last kernel clears
// what the sort used.
struct SortScratch {
  int *counts = nullptr;      // [RS_MAX_SWEEPS][256] keys per digit of each pass
  int *tickets = nullptr;     // [RS_MAX_SWEEPS] next tile of each pass
  unsigned *tiles = nullptr;  // [RS_MAX_SWEEPS][max_blocks][256] look-back words
  int *flags = nullptr;       // [RS_MAX_SWEEPS] pass moved keys (written by every sort)
  int *status = nullptr;      // workspace status: a look-back that waits too long sets ERR_WAIT
  int max_blocks = 0;
};

// Sort (keys, ids) in place by the key bits [bit_begin, bit_end), stably.
// count: device pointer (clamped to capacity) or nullptr to use count_static.
// alt_keys / alt_ids are ping-pong buffers of the same capacity.
// With finish == false the caller's next kernel must run rs_finish over the
// items (the ids after an odd number of moving passes, and the scratch).
void radix_sort(hipStream_t s, uint64_t *keys, int *ids, uint64_t *alt_keys, int *alt_ids,
                const int *count, int count_static, int capacity, int bit_begin, int bit_end,
                const SortScratch &scratch, bool finish = true);

// Zero between scans: allocation clears it and every scan's last tile clears
// what the scan used.
struct ScanScratch {
  unsigned long long *words = nullptr;  // [max_blocks] look-back words
  int *counters = nullptr;              // [2] next tile ticket, finished tiles
  int *status = nullptr;                // workspace status: a look-back that waits too long sets ERR_WAIT
  int max_blocks = 0;
};

// Exclusive prefix sum of in[0, n) into out[0, n); total written to *total.
void scan_exclusive(hipStream_t s, const int *in, int *out, const int *count, int count_static,
                    int capacity, int *total, const ScanScratch &scratch);

// Implicit complete tree whose leaf count is decided on the device: leaves =
// the power of two at or above the live count (stored in *leaves), so a
// captured graph builds and traverses only the live part of a pool whose
// capacity is `size`.
constexpr int HIER_GROUP = 512;                   // nodes reduced by one block per stage
constexpr int HIER_STAGE_LEVELS = 9;              // levels one stage reduces (log2 HIER_GROUP)

struct Hierarchy {
  uint64_t *keys = nullptr, *alt_keys = nullptr;
  int *ids = nullptr, *alt_ids = nullptr;  // [size]
  Box *nodes = nullptr;                     // [2 * size]
  int *leaves = nullptr;                    // device: current leaf count
  int size = 1;                             // capacity (power of two)
  __device__ Tree tree() const { return Tree{nodes, ids, __hip_atomic_load(leaves, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)}; }
};

// Morton-ordered balanced hierarchy over leaf boxes[0, count) (warp_peel/
// solver.py:34-89 semantics: 15-bit x/y Morton code of the binary32 box
// centre, layer in the high word when layered, stable order, empty padding).
// Optional per-leaf records gathered into leaf order while the leaves are
// written: out_a[i], out_b[i], out_near[i] = (near[id], near[near_stride + id])
// for the leaf's id = ids[i]. A walk that reaches leaf i then reads its record
// at i, with no dependent load of the id first.
struct LeafGather {
  const double2 *a = nullptr, *b = nullptr;
  const double *near = nullptr;
  long near_stride = 0;
  double2 *out_a = nullptr, *out_b = nullptr, *out_near = nullptr;
};

void build_hierarchy(hipStream_t s, const Box *boxes, const int *count, int count_static,
                     int capacity, bool layered, Hierarchy &h, const SortScratch &scratch,
                     const LeafGather &gather = LeafGather{});

int pow2_at_least(int n);

}  // namespace peel
