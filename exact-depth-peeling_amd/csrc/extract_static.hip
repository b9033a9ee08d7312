// peel_extract_static: one extraction that allocates nothing.
// Declaration and contract: include/peel355.h.
#include <algorithm>
#include <mutex>
#include <string>

#include "layout.hpp"

using namespace peel;

// One extraction without allocation: every pool except the caller's outputs
// is carved from `scratch`, a file-scope device array of fixed size that each
// GPU holds for the life of the process. Its host-side name is not a device
// address: host code uses only the per-device pointer from
// hipGetSymbolAddress, resolved once per device on that device.
static __device__ __align__(256) char scratch[PEEL_STATIC_SCRATCH_BYTES];

namespace {
constexpr int MAX_DEVICES = 64;
std::mutex g_scratch_lock;
char *g_scratch_ptr[MAX_DEVICES] = {};

// Zero `count` 16-byte words (a kernel, not hipMemsetAsync, so the call may
// also be recorded into a graph).
__global__ void k_fill_zero(uint4 *p, long count) {
  long stride = (long)gridDim.x * blockDim.x;
  for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < count; i += stride) p[i] = make_uint4(0, 0, 0, 0);
}

int scratch_for_device(int device, char **out) {
  if (device < 0 || device >= MAX_DEVICES) return fail(2, "device index out of range");
  std::lock_guard<std::mutex> guard(g_scratch_lock);
  if (!g_scratch_ptr[device]) {
    void *ptr = nullptr;
    hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr)
      return fail(e != hipSuccess ? -(int)e : -1,
                  std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
    g_scratch_ptr[device] = static_cast<char *>(ptr);
  }
  *out = g_scratch_ptr[device];
  return 0;
}
}  // namespace

extern "C" int peel_extract_static(const PeelDesc *d, const PeelOutputs *o, void *stream) {
  if (!d || !o) return fail(1, "null argument");
  if (!o->output_xy || !o->output_n || !o->output_sources || !o->output_layers || !o->triangles_uv ||
      !o->triangles_world || !o->triangle_sources || !o->triangle_layers || !o->status || !o->metrics)
    return fail(3, "null output array");
  Ws w;
  int r = plan(w, d);
  if (r) return r;
  Carver measure;
  carve(w, measure, false);
  if (measure.used > sizeof(scratch))
    return fail(2, "extraction needs " + std::to_string(measure.used) + " bytes of scratch; the static scratch holds " +
                       std::to_string(sizeof(scratch)) + " (use peel_create for larger meshes or capacities)");
  int device = 0;
  HIP_TRY(hipGetDevice(&device));
  char *base = nullptr;
  if ((r = scratch_for_device(device, &base))) return r;
  w.status = o->status;
  w.metrics = (long long *)o->metrics;
  w.out_xy = (double2 *)o->output_xy;
  w.out_n = o->output_n;
  w.out_src = o->output_sources;
  w.out_layer = o->output_layers;
  w.tri_uv = o->triangles_uv;
  w.tri_world = o->triangles_world;
  w.tri_src = o->triangle_sources;
  w.tri_layer = o->triangle_layers;
  Carver pools;
  pools.base = base;
  carve(w, pools, false);
  int cus = 0;
  HIP_TRY(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, device));
  w.persistent_blocks = std::max(1, cus);
  hipStream_t s = static_cast<hipStream_t>(stream);
  // the previous call's contents stay in the array: clear the part in use
  long words = (long)((measure.used + 15) / 16);
  hipLaunchKernelGGL(k_fill_zero, dim3((unsigned)std::min<long>(4096, (words + 255) / 256)), dim3(256), 0, s,
                     reinterpret_cast<uint4 *>(base), words);
  enqueue_all(w, s, nullptr);
  HIP_TRY(hipGetLastError());
  return 0;
}

