// Learned pool sizes: a small multilayer perceptron that predicts, from cheap
// mesh statistics, how much of each extraction pool a peel will use. It has no
// counterpart in the reference (whose capacities are fixed multiples of the
// face count, warp_peel/events.py:19-33); a predicted pool that overflows is
// detected by the status bits and the extraction is rerun with the
// conservative capacities.
//
// Network: inputs (<= 16) -> 64 tanh -> 64 tanh -> outputs (<= 8), binary32,
// parameters flat as [W1 (64 x I) | b1 | W2 (64 x 64) | b2 | W3 (O x 64) | b3],
// row-major (output index major).
//
// Training loss: asymmetric squared error over samples s and outputs o,
//   L = 1/(N O) * sum_{s,o} w(r) r^2,  r = y - t,  w = under if r < 0 else 1,
// so under-predictions (overflows) cost `under` times more.
//
// Kernels:
//   k_mlp_tile    one wave per tile of 64 samples: forward, loss and backward
//                 to the pre-activations in LDS (one lane per sample), then the
//                 tile's weight-gradient partial sums, each summed over the
//                 tile's samples in index order (deterministic);
//   k_mlp_reduce  one lane per parameter: sum of the tile partials in tile
//                 order (deterministic), the loss likewise;
//   k_mlp_adam    Adam step (Kingma and Ba, 2015) with host-computed bias
//                 corrections;
//   k_mlp_forward one lane per sample, hidden activations in registers.
#include <hip/hip_runtime.h>

#include <cstdint>

namespace {

constexpr int H = 64;      // hidden width
constexpr int MAX_I = 16;  // inputs
constexpr int MAX_O = 8;   // outputs
constexpr int TILE = 64;   // samples per tile (one wave, one lane each)
constexpr int PAD = H + 1; // LDS row pitch: lanes read rows, so pad the banks

struct Layout {
  int I, O;
  __host__ __device__ int w1() const { return 0; }
  __host__ __device__ int b1() const { return H * I; }
  __host__ __device__ int w2() const { return H * I + H; }
  __host__ __device__ int b2() const { return H * I + H + H * H; }
  __host__ __device__ int w3() const { return H * I + 2 * H + H * H; }
  __host__ __device__ int b3() const { return H * I + 2 * H + H * H + O * H; }
  __host__ __device__ int count() const { return H * I + 2 * H + H * H + O * H + O; }
};

__global__ void __launch_bounds__(TILE) k_mlp_tile(const float *__restrict__ p, Layout L, const float *__restrict__ x,
                                                   const float *__restrict__ t, int batch, float under,
                                                   float *__restrict__ partial) {
  __shared__ float X[TILE][MAX_I];
  __shared__ float H1[TILE][PAD], D1[TILE][PAD], H2[TILE][PAD], D2[TILE][PAD];
  __shared__ float DY[TILE][MAX_O];
  __shared__ float LOSS[TILE];
  const int s = threadIdx.x;
  const int sample = blockIdx.x * TILE + s;
  const bool valid = sample < batch;
  const int I = L.I, O = L.O;
  const float scale = 2.0f / ((float)batch * (float)O);
  for (int i = 0; i < MAX_I; ++i) X[s][i] = (valid && i < I) ? x[(long)sample * I + i] : 0.0f;
  // forward
  for (int j = 0; j < H; ++j) {
    float z = p[L.b1() + j];
    for (int i = 0; i < I; ++i) z += p[L.w1() + j * I + i] * X[s][i];
    H1[s][j] = tanhf(z);
  }
  for (int k = 0; k < H; ++k) {
    float z = p[L.b2() + k];
    for (int j = 0; j < H; ++j) z += p[L.w2() + k * H + j] * H1[s][j];
    H2[s][k] = tanhf(z);
  }
  float loss = 0.0f;
  for (int o = 0; o < MAX_O; ++o) {
    float dy = 0.0f;
    if (o < O && valid) {
      float y = p[L.b3() + o];
      for (int k = 0; k < H; ++k) y += p[L.w3() + o * H + k] * H2[s][k];
      float r = y - t[(long)sample * O + o];
      float w = r < 0.0f ? under : 1.0f;
      loss += w * r * r;
      dy = scale * w * r;
    }
    DY[s][o] = dy;
  }
  LOSS[s] = loss;
  // backward to the pre-activations
  for (int k = 0; k < H; ++k) {
    float g = 0.0f;
    for (int o = 0; o < O; ++o) g += p[L.w3() + o * H + k] * DY[s][o];
    float h = H2[s][k];
    D2[s][k] = g * (1.0f - h * h);
  }
  for (int j = 0; j < H; ++j) {
    float g = 0.0f;
    for (int k = 0; k < H; ++k) g += p[L.w2() + k * H + j] * D2[s][k];
    float h = H1[s][j];
    D1[s][j] = g * (1.0f - h * h);
  }
  __syncthreads();
  // weight-gradient partials of this tile, samples summed in index order
  const int P = L.count();
  float *out = partial + (long)blockIdx.x * (P + 1);
  for (int q = s; q < P; q += TILE) {
    float g = 0.0f;
    if (q < L.b1()) {
      int j = q / I, i = q % I;
      for (int u = 0; u < TILE; ++u) g += D1[u][j] * X[u][i];
    } else if (q < L.w2()) {
      int j = q - L.b1();
      for (int u = 0; u < TILE; ++u) g += D1[u][j];
    } else if (q < L.b2()) {
      int k = (q - L.w2()) / H, j = (q - L.w2()) % H;
      for (int u = 0; u < TILE; ++u) g += D2[u][k] * H1[u][j];
    } else if (q < L.w3()) {
      int k = q - L.b2();
      for (int u = 0; u < TILE; ++u) g += D2[u][k];
    } else if (q < L.b3()) {
      int o = (q - L.w3()) / H, k = (q - L.w3()) % H;
      for (int u = 0; u < TILE; ++u) g += DY[u][o] * H2[u][k];
    } else {
      int o = q - L.b3();
      for (int u = 0; u < TILE; ++u) g += DY[u][o];
    }
    out[q] = g;
  }
  if (s == 0) {
    float l = 0.0f;
    for (int u = 0; u < TILE; ++u) l += LOSS[u];
    out[P] = l;
  }
}

__global__ void k_mlp_reduce(const float *__restrict__ partial, int tiles, int P, float norm, float *__restrict__ grad,
                             float *__restrict__ loss) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q > P) return;
  float g = 0.0f;
  for (int b = 0; b < tiles; ++b) g += partial[(long)b * (P + 1) + q];
  if (q < P) grad[q] = g;
  else if (loss) *loss = g * norm;
}

__global__ void k_mlp_adam(float *__restrict__ p, const float *__restrict__ g, float *__restrict__ m,
                           float *__restrict__ v, int count, float lr, float beta1, float beta2, float eps, float bc1,
                           float bc2) {
  int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= count) return;
  float gq = g[q];
  float mq = beta1 * m[q] + (1.0f - beta1) * gq;
  float vq = beta2 * v[q] + (1.0f - beta2) * gq * gq;
  m[q] = mq;
  v[q] = vq;
  p[q] -= lr * (mq / bc1) / (sqrtf(vq / bc2) + eps);
}

__global__ void __launch_bounds__(256) k_mlp_forward(const float *__restrict__ p, Layout L, const float *__restrict__ x,
                                                     int batch, float *__restrict__ y) {
  int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= batch) return;
  const int I = L.I, O = L.O;
  float xi[MAX_I];
#pragma unroll
  for (int i = 0; i < MAX_I; ++i) xi[i] = i < I ? x[(long)s * I + i] : 0.0f;
  // constant trip counts keep the activations in registers
  float h1[H];
#pragma unroll
  for (int j = 0; j < H; ++j) {
    float z = p[L.b1() + j];
#pragma unroll
    for (int i = 0; i < MAX_I; ++i)
      if (i < I) z += p[L.w1() + j * I + i] * xi[i];
    h1[j] = tanhf(z);
  }
  float out[MAX_O];
#pragma unroll
  for (int o = 0; o < MAX_O; ++o) out[o] = o < O ? p[L.b3() + o] : 0.0f;
  for (int k = 0; k < H; ++k) {
    float z = p[L.b2() + k];
#pragma unroll
    for (int j = 0; j < H; ++j) z += p[L.w2() + k * H + j] * h1[j];
    float h2 = tanhf(z);
#pragma unroll
    for (int o = 0; o < MAX_O; ++o)
      if (o < O) out[o] += p[L.w3() + o * H + k] * h2;
  }
#pragma unroll
  for (int o = 0; o < MAX_O; ++o)
    if (o < O) y[(long)s * O + o] = out[o];
}

bool shape_ok(int inputs, int outputs) { return inputs >= 1 && inputs <= MAX_I && outputs >= 1 && outputs <= MAX_O; }

int launch_status() { return hipGetLastError() == hipSuccess ? 0 : -1; }

}  // namespace

extern "C" {

int32_t pool_mlp_param_count(int32_t inputs, int32_t outputs) {
  if (!shape_ok(inputs, outputs)) return -1;
  return Layout{inputs, outputs}.count();
}

int32_t pool_mlp_partial_count(int32_t inputs, int32_t outputs, int32_t batch) {
  if (!shape_ok(inputs, outputs) || batch < 1) return -1;
  return ((batch + TILE - 1) / TILE) * (Layout{inputs, outputs}.count() + 1);
}

int pool_mlp_forward(const float *params, int32_t inputs, int32_t outputs, const float *x, int32_t batch, float *y,
                     void *stream) {
  if (!shape_ok(inputs, outputs) || batch < 0 || !params || (batch > 0 && (!x || !y))) return 1;
  if (batch == 0) return 0;
  hipLaunchKernelGGL(k_mlp_forward, dim3((batch + 255) / 256), dim3(256), 0, (hipStream_t)stream, params,
                     Layout{inputs, outputs}, x, batch, y);
  return launch_status();
}

int pool_mlp_gradient(const float *params, int32_t inputs, int32_t outputs, const float *x, const float *target,
                      int32_t batch, float under_weight, float *partial, float *grad, float *loss, void *stream) {
  if (!shape_ok(inputs, outputs) || batch < 1 || !params || !x || !target || !partial || !grad) return 1;
  if (!(under_weight > 0.0f)) return 2;
  Layout L{inputs, outputs};
  int tiles = (batch + TILE - 1) / TILE;
  int P = L.count();
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(k_mlp_tile, dim3(tiles), dim3(TILE), 0, s, params, L, x, target, batch, under_weight, partial);
  hipLaunchKernelGGL(k_mlp_reduce, dim3((P + 1 + 255) / 256), dim3(256), 0, s, partial, tiles, P,
                     1.0f / ((float)batch * (float)outputs), grad, loss);
  return launch_status();
}

int pool_mlp_adam(float *params, const float *grad, float *m, float *v, int32_t count, float lr, float beta1,
                  float beta2, float eps, float bias1, float bias2, void *stream) {
  if (count < 1 || !params || !grad || !m || !v) return 1;
  if (!(bias1 > 0.0f) || !(bias2 > 0.0f)) return 2;
  hipLaunchKernelGGL(k_mlp_adam, dim3((count + 255) / 256), dim3(256), 0, (hipStream_t)stream, params, grad, m, v,
                     count, lr, beta1, beta2, eps, bias1, bias2);
  return launch_status();
}

}  // extern "C"
