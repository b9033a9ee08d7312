// Workspace layout shared by the entry points that own their pools
// (peel_create) and those that borrow them (peel_extract_static).
#pragma once
#include <string>

#include "../../include/peel355.h"
#include "workspace.hpp"

namespace peel {

// Device memory of a workspace is one block carved into 256-byte aligned
// pools. The same sequence of take() calls measures the block (base ==
// nullptr) and then hands out its pools, so the layout is written once.
struct Carver {
  char *base = nullptr;
  size_t used = 0;
  template <class T>
  void take(T *&ptr, long count) {
    if (count < 1) count = 1;
    used = (used + 255) & ~size_t(255);
    ptr = base ? reinterpret_cast<T *>(base + used) : nullptr;
    used += sizeof(T) * (size_t)count;
  }
};

// Record `message` as this thread's last error (peel_last_error) and return code.
int fail(int code, const std::string &message);
// Validate a descriptor and set every scalar of the layout; 0 or a positive error.
int plan(Ws &w, const PeelDesc *d);
// Carve every pool of a planned workspace; `outputs` false leaves the
// polygons, triangles, status and metrics to the caller.
void carve(Ws &w, Carver &c, bool outputs);
// Launch the whole pipeline on `s` (launches only).
void enqueue_all(const Ws &w, hipStream_t s, Timeline *t);

}  // namespace peel

// Return a negative code with the runtime's message when a HIP call fails.
#define HIP_TRY(expr)                                                                  \
  do {                                                                                 \
    hipError_t _e = (expr);                                                            \
    if (_e != hipSuccess)                                                              \
      return fail(-(int)_e, std::string(#expr) + ": " + hipGetErrorString(_e));        \
  } while (0)
