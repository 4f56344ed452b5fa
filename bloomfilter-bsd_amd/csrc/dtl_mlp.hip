// dtl_mlp.hip -- the 8-64-64-1 FP32 multilayer perceptron of the learned
// configuration model: hand-written training (forward, backward, Adam) and
// inference kernels.  No reference counterpart: it stands in for the
// reference's exhaustive calibration + skyline lookup
// (model/calibration.cpp:251-311, tool/calibration/src/skyline_matrix_builder.cpp).
//
// Parameter layout (floats): W1[64][8] b1[64] W2[64][64] b2[64] W3[64] b3.
// Training: each workgroup takes 64 samples of the minibatch, keeps the
// activations in LDS, forms its share of the gradient with small LDS matrix
// products (dW2 = dH2^T H1 etc.) and adds it to the global gradient with
// float atomics; a second kernel applies Adam and clears the gradient.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "dtl_hip.h"
#include "dtl_internal.hpp"

namespace dtlh {

constexpr int kIn = 8, kH = 64;
constexpr int oW1 = 0, oB1 = oW1 + kH * kIn, oW2 = oB1 + kH, oB2 = oW2 + kH * kH, oW3 = oB2 + kH,
              oB3 = oW3 + kH, kParams = oB3 + 1;
constexpr int kRows = 64;  // samples per training workgroup

__device__ __forceinline__ float relu(float v) { return v > 0.f ? v : 0.f; }

// One sample per thread; the parameters are read from LDS.
__global__ __launch_bounds__(256) void k_mlp_forward(const float* __restrict__ params,
                                                     const float* __restrict__ x, uint32_t n,
                                                     float* __restrict__ y) {
  __shared__ float p[kParams];
  for (int i = threadIdx.x; i < kParams; i += blockDim.x) p[i] = params[i];
  __syncthreads();
  const uint32_t r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  float in[kIn];
#pragma unroll
  for (int k = 0; k < kIn; ++k) in[k] = x[static_cast<uint64_t>(r) * kIn + k];
  float h1[kH];
#pragma unroll
  for (int j = 0; j < kH; ++j) {
    float a = p[oB1 + j];
#pragma unroll
    for (int k = 0; k < kIn; ++k) a += p[oW1 + j * kIn + k] * in[k];
    h1[j] = relu(a);
  }
  float out = p[oB3];
  for (int i = 0; i < kH; ++i) {
    float a = p[oB2 + i];
#pragma unroll
    for (int j = 0; j < kH; ++j) a += p[oW2 + i * kH + j] * h1[j];
    out += p[oW3 + i] * relu(a);
  }
  y[r] = out;
}

// Gradient of 0.5 * mean(w * (f(x) - t)^2) over the minibatch rows
// idx(j) = (a * (base + j) + c) mod n, j < count (an epoch-specific
// permutation); w = wts[idx] (1 when no weights are given), plus, when
// `best` is given, dyn * exp(-max(f(x) - best[idx], 0) * inv_temp): rows the
// network currently predicts close to (or below) the best cost of their
// query weigh more, so a poor candidate predicted to be good is corrected.
// That term is a weight, held constant in the gradient.
__global__ __launch_bounds__(256) void k_mlp_grad(const float* __restrict__ params,
                                                  const float* __restrict__ x,
                                                  const float* __restrict__ t,
                                                  const float* __restrict__ wts,
                                                  const float* __restrict__ best, float dyn,
                                                  float inv_temp, uint32_t n,
                                                  uint64_t perm_a, uint64_t perm_c, uint32_t base,
                                                  uint32_t count, float* __restrict__ grad) {
  __shared__ float p[kParams];
  __shared__ float X[kRows][kIn];
  __shared__ float H1[kRows][kH + 1];
  __shared__ float H2[kRows][kH + 1];  // post-ReLU; reused for dH2 (pre-activation grad)
  __shared__ float D1[kRows][kH + 1];  // dH1 (pre-activation grad)
  __shared__ float dy[kRows];
  const int tid = threadIdx.x;
  for (int i = tid; i < kParams; i += blockDim.x) p[i] = params[i];
  const uint32_t row0 = blockIdx.x * kRows;
  const int rows = static_cast<int>(min<uint32_t>(kRows, count - row0));
  for (int i = tid; i < kRows * kIn; i += blockDim.x) {
    const int r = i / kIn, k = i % kIn;
    float v = 0.f;
    if (r < rows) {
      const uint64_t idx = (perm_a * (base + row0 + r) + perm_c) % n;
      v = x[idx * kIn + k];
    }
    X[r][k] = v;
  }
  __syncthreads();
  // H1 = relu(X W1^T + b1)
  for (int i = tid; i < kRows * kH; i += blockDim.x) {
    const int r = i / kH, j = i % kH;
    float a = p[oB1 + j];
#pragma unroll
    for (int k = 0; k < kIn; ++k) a += p[oW1 + j * kIn + k] * X[r][k];
    H1[r][j] = relu(a);
  }
  __syncthreads();
  // H2 = relu(H1 W2^T + b2)
  for (int i = tid; i < kRows * kH; i += blockDim.x) {
    const int r = i / kH, j = i % kH;
    float a = p[oB2 + j];
    for (int k = 0; k < kH; ++k) a += p[oW2 + j * kH + k] * H1[r][k];
    H2[r][j] = relu(a);
  }
  __syncthreads();
  // output, dy = (y - t) / count
  if (tid < kRows) {
    float dv = 0.f;
    if (tid < rows) {
      float yv = p[oB3];
      for (int k = 0; k < kH; ++k) yv += p[oW3 + k] * H2[tid][k];
      const uint64_t idx = (perm_a * (base + row0 + tid) + perm_c) % n;
      float wv = wts ? wts[idx] : 1.0f;
      if (best) wv += dyn * __expf(-fmaxf(yv - best[idx], 0.f) * inv_temp);
      dv = (yv - t[idx]) * wv / static_cast<float>(count);
    }
    dy[tid] = dv;
  }
  __syncthreads();
  // dW3 = sum_r dy H2 ; db3 = sum_r dy
  for (int j = tid; j < kH + 1; j += blockDim.x) {
    float g = 0.f;
    for (int r = 0; r < kRows; ++r) g += dy[r] * (j < kH ? H2[r][j] : 1.f);
    atomicAdd(&grad[oW3 + j], g);  // oB3 == oW3 + kH
  }
  __syncthreads();
  // dH2 (pre-activation) = dy W3 * [H2 > 0], stored over H2
  for (int i = tid; i < kRows * kH; i += blockDim.x) {
    const int r = i / kH, j = i % kH;
    H2[r][j] = H2[r][j] > 0.f ? dy[r] * p[oW3 + j] : 0.f;
  }
  __syncthreads();
  // dW2 = dH2^T H1 ; db2 = sum_r dH2 ; dH1 = (dH2 W2) * [H1 > 0]
  for (int i = tid; i < kH * kH; i += blockDim.x) {
    const int j = i / kH, k = i % kH;
    float g = 0.f;
    for (int r = 0; r < kRows; ++r) g += H2[r][j] * H1[r][k];
    atomicAdd(&grad[oW2 + i], g);
  }
  for (int j = tid; j < kH; j += blockDim.x) {
    float g = 0.f;
    for (int r = 0; r < kRows; ++r) g += H2[r][j];
    atomicAdd(&grad[oB2 + j], g);
  }
  for (int i = tid; i < kRows * kH; i += blockDim.x) {
    const int r = i / kH, k = i % kH;
    float a = 0.f;
    for (int j = 0; j < kH; ++j) a += H2[r][j] * p[oW2 + j * kH + k];
    D1[r][k] = H1[r][k] > 0.f ? a : 0.f;
  }
  __syncthreads();
  // dW1 = dH1^T X ; db1 = sum_r dH1
  for (int i = tid; i < kH * (kIn + 1); i += blockDim.x) {
    const int j = i / (kIn + 1), k = i % (kIn + 1);
    float g = 0.f;
    for (int r = 0; r < kRows; ++r) g += D1[r][j] * (k < kIn ? X[r][k] : 1.f);
    atomicAdd(&grad[k < kIn ? oW1 + j * kIn + k : oB1 + j], g);
  }
}

__global__ void k_mlp_adam(float* __restrict__ params, float* __restrict__ grad,
                           float* __restrict__ m, float* __restrict__ v, float lr, float b1t,
                           float b2t) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= kParams) return;
  const float g = grad[i];
  grad[i] = 0.f;
  const float mi = 0.9f * m[i] + 0.1f * g;
  const float vi = 0.999f * v[i] + 0.001f * g * g;
  m[i] = mi;
  v[i] = vi;
  const float mh = mi / (1.f - b1t), vh = vi / (1.f - b2t);
  params[i] -= lr * mh / (sqrtf(vh) + 1e-8f);
}

__global__ void k_fill_zero(float* p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0.f;
}

// Distinct keys from a seeded bijection of [0, 2^31) (same recipe as the
// benchmark's inputs); bit 31 = high_bit.
__global__ void k_gen_keys(uint32_t* __restrict__ out, uint64_t n, uint32_t seed, uint32_t high_bit) {
  const uint64_t stride = static_cast<uint64_t>(gridDim.x) * blockDim.x;
  for (uint64_t i = static_cast<uint64_t>(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += stride) {
    const uint64_t M = (1ull << 31) - 1;
    uint64_t x = (i + (static_cast<uint64_t>(seed) << 20)) & M;
    x = (x * 0x5851F42Dull) & M;
    x ^= x >> 16;
    x = (x * 0x2C1B3C6Dull) & M;
    x ^= x >> 13;
    x = (x * 0x297A2D39ull) & M;
    out[i] = static_cast<uint32_t>(x) | (high_bit << 31);
  }
}

hipError_t gen_keys(uint32_t* out, uint64_t n, uint32_t seed, uint32_t high_bit, hipStream_t st) {
  const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 65536);
  hipLaunchKernelGGL(k_gen_keys, dim3(static_cast<uint32_t>(std::max<uint64_t>(blocks, 1))), dim3(256), 0,
                     st, out, n, seed, high_bit);
  return hipGetLastError();
}

}  // namespace dtlh

using namespace dtlh;

struct dtl_mlp {
  float* d = nullptr;  // params | grad | adam m | adam v
  uint64_t step = 0;
};

extern "C" {

int dtl_mlp_input_count(void) { return kIn; }

int dtl_mlp_param_count(void) { return kParams; }

int dtl_mlp_create(uint32_t seed, dtl_mlp** out) {
  if (!out) return DTL_EINVAL;
  *out = nullptr;
  dtl_mlp* m = new dtl_mlp();
  if (hipMalloc(&m->d, sizeof(float) * kParams * 4) != hipSuccess) {
    delete m;
    return DTL_ENOMEM;
  }
  // He-uniform initialisation from a seeded xorshift stream; biases zero.
  std::vector<float> h(static_cast<size_t>(kParams) * 4, 0.f);
  uint64_t s = 0x9E3779B97F4A7C15ull ^ seed;
  auto uni = [&]() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return static_cast<float>((s >> 11) * (1.0 / 9007199254740992.0)) * 2.f - 1.f;
  };
  const float a1 = std::sqrt(6.f / kIn), a2 = std::sqrt(6.f / kH);
  for (int i = 0; i < kH * kIn; ++i) h[oW1 + i] = a1 * uni();
  for (int i = 0; i < kH * kH; ++i) h[oW2 + i] = a2 * uni();
  for (int i = 0; i < kH; ++i) h[oW3 + i] = a2 * uni() * 0.1f;
  if (hipMemcpy(m->d, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipFree(m->d);
    delete m;
    return DTL_EHIP;
  }
  *out = m;
  return DTL_OK;
}

void dtl_mlp_destroy(dtl_mlp* m) {
  if (!m) return;
  if (m->d) (void)hipFree(m->d);
  delete m;
}

int dtl_mlp_get_params(const dtl_mlp* m, float* h_params) {
  if (!m || !h_params) return DTL_EINVAL;
  return hipMemcpy(h_params, m->d, sizeof(float) * kParams, hipMemcpyDeviceToHost) == hipSuccess
             ? DTL_OK
             : DTL_EHIP;
}

int dtl_mlp_set_params(dtl_mlp* m, const float* h_params) {
  if (!m || !h_params) return DTL_EINVAL;
  return hipMemcpy(m->d, h_params, sizeof(float) * kParams, hipMemcpyHostToDevice) == hipSuccess
             ? DTL_OK
             : DTL_EHIP;
}

// Trains on n samples (d_x: n x kIn = 8 features, d_t: n targets, d_w: n sample
// weights or null; device memory) with Adam, minibatches of `batch` rows,
// `epochs` passes; each epoch visits the rows in the order (a * j + c) mod n
// for a fresh (a coprime to n, c).
int dtl_mlp_train(dtl_mlp* m, const float* d_x, const float* d_t, const float* d_w, uint32_t n,
                  uint32_t epochs, uint32_t batch, float lr, uint32_t seed, dtl_stream_t stream) {
  return dtl_mlp_train_ex(m, d_x, d_t, d_w, nullptr, 0.f, 1.f, n, epochs, batch, lr, seed, stream);
}

// As dtl_mlp_train, with the prediction-dependent weight of k_mlp_grad when
// d_best is given (temp > 0, in the units of the targets).
int dtl_mlp_train_ex(dtl_mlp* m, const float* d_x, const float* d_t, const float* d_w, const float* d_best,
                     float dyn, float temp, uint32_t n, uint32_t epochs, uint32_t batch, float lr, uint32_t seed,
                     dtl_stream_t stream) {
  if (!m || !d_x || !d_t || n == 0 || batch == 0 || !(temp > 0.f)) return DTL_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  (void)hipGetLastError();
  float* params = m->d;
  float* grad = m->d + kParams;
  float* am = m->d + 2 * kParams;
  float* av = m->d + 3 * kParams;
  uint64_t s = 0xD1B54A32D192ED03ull ^ (static_cast<uint64_t>(seed) << 1);
  auto next = [&]() {
    s ^= s << 13;
    s ^= s >> 7;
    s ^= s << 17;
    return s;
  };
  auto gcd = [](uint64_t a, uint64_t b) {
    while (b) {
      const uint64_t t = a % b;
      a = b;
      b = t;
    }
    return a;
  };
  hipLaunchKernelGGL(k_fill_zero, dim3((kParams + 255) / 256), dim3(256), 0, st, grad, kParams);
  for (uint32_t e = 0; e < epochs; ++e) {
    uint64_t a = next() % n;
    while (a == 0 || gcd(a, n) != 1) a = (a + 1) % n + (n == 1 ? 1 : 0);
    const uint64_t c = next() % n;
    for (uint32_t b0 = 0; b0 < n; b0 += batch) {
      const uint32_t cnt = std::min(batch, n - b0);
      const uint32_t blocks = (cnt + kRows - 1) / kRows;
      hipLaunchKernelGGL(k_mlp_grad, dim3(blocks), dim3(256), 0, st, params, d_x, d_t, d_w, d_best, dyn,
                         1.f / temp, n, a, c, b0, cnt, grad);
      m->step += 1;
      const float b1t = std::pow(0.9f, static_cast<float>(m->step));
      const float b2t = std::pow(0.999f, static_cast<float>(m->step));
      hipLaunchKernelGGL(k_mlp_adam, dim3((kParams + 255) / 256), dim3(256), 0, st, params, grad, am,
                         av, lr, b1t, b2t);
    }
  }
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    set_last_error(err);
    return DTL_EHIP;
  }
  return DTL_OK;
}

// Gradient of 0.5 * mean(w * (f(x) - t)^2) over all n rows (w = d_w[i], or
// 1 if d_w is null) into d_grad (kParams floats, device); the model's
// parameters are unchanged.
int dtl_mlp_gradient(const dtl_mlp* m, const float* d_x, const float* d_t, const float* d_w, uint32_t n,
                     float* d_grad, dtl_stream_t stream) {
  if (!m || !d_x || !d_t || !d_grad || n == 0) return DTL_EINVAL;
  hipStream_t st = static_cast<hipStream_t>(stream);
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_fill_zero, dim3((kParams + 255) / 256), dim3(256), 0, st, d_grad, kParams);
  hipLaunchKernelGGL(k_mlp_grad, dim3((n + kRows - 1) / kRows), dim3(256), 0, st, m->d, d_x, d_t, d_w,
                     static_cast<const float*>(nullptr), 0.f, 1.f, n, uint64_t(1), uint64_t(0), 0u, n, d_grad);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    set_last_error(err);
    return DTL_EHIP;
  }
  return DTL_OK;
}

int dtl_mlp_predict(const dtl_mlp* m, const float* d_x, uint32_t n, float* d_y, dtl_stream_t stream) {
  if (!m || !d_x || !d_y) return DTL_EINVAL;
  if (n == 0) return DTL_OK;
  (void)hipGetLastError();
  hipLaunchKernelGGL(k_mlp_forward, dim3((n + 255) / 256), dim3(256), 0, static_cast<hipStream_t>(stream),
                     m->d, d_x, n, d_y);
  const hipError_t err = hipGetLastError();
  if (err != hipSuccess) {
    set_last_error(err);
    return DTL_EHIP;
  }
  return DTL_OK;
}

}  // extern "C"
