// Stage 3a: rank propagation across certified patch adjacencies
// (warp_peel/components.py:15-271). Placeholder launcher until the
// connectivity kernels land: every patch is its own representative.
#include "workspace.hpp"

namespace peel {

__global__ void k_identity(Ws w) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < w.cap.regions) w.reps[i] = i;
}

void enqueue_connect(const Ws &w, hipStream_t s) {
  hipLaunchKernelGGL(k_identity, dim3((w.cap.regions + 255) / 256), dim3(256), 0, s, w);
}

}  // namespace peel
