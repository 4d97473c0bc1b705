// Matching-order network: a 32-64-64-64-10 ReLU multilayer perceptron with hand-written forward
// (f32 and bf16), backward and Adam kernels.  It has no counterpart in the reference (whose join
// fixes identity order); the layout and loss are those declared in bms2.h.
//
// Kernels, all one thread per row except the weight-gradient and Adam kernels:
//   forward_kernel<Save, Bf16>  weights staged in LDS; per-thread activations kept in an LDS
//                               column [feature][thread] so the layer loops index them freely.
//   backward_kernel             output deltas from the loss, then deltas of every hidden layer.
//   wgrad_kernel                one thread per parameter: dW[i][j] = sum_r a_i(r) dz_j(r), summed
//                               in f64 over rows in a fixed order (deterministic gradients).
//   adam_kernel                 elementwise Adam update.
// Saved activations and deltas are feature-major ([feature][row]) so the per-row kernels write
// them coalesced and the weight-gradient kernel streams them.
#include <cmath>
#include <cstring>
#include <vector>

#include "runtime.hpp"

struct bms2_mlp_s {
  bms2_queue q;
  float* d_p = nullptr;  // parameters
  float* d_g = nullptr;  // gradient
  float* d_m = nullptr;  // Adam first moment
  float* d_v = nullptr;  // Adam second moment
  float* d_act = nullptr;    // [3][64][cap] hidden activations
  float* d_delta = nullptr;  // [4][64][cap] deltas of hidden layers 1..3 and the output
  float* d_y = nullptr;      // [cap][10]
  float* d_loss = nullptr;   // [cap / kRows] per-block loss partials
  uint32_t cap = 0;
  uint64_t step = 0;
};

namespace bms2 {
namespace {

constexpr int kIn = BMS2_MLP_IN, kH = BMS2_MLP_HIDDEN, kOut = BMS2_MLP_OUT;
constexpr int oW1 = 0, oB1 = oW1 + kIn * kH, oW2 = oB1 + kH, oB2 = oW2 + kH * kH;
constexpr int oW3 = oB2 + kH, oB3 = oW3 + kH * kH, oW4 = oB3 + kH, oB4 = oW4 + kH * kOut;
constexpr int kParams = oB4 + kOut;
constexpr int kParamsLds = (kParams + 3) / 4 * 4;
constexpr int kRows = 64;  // rows (threads) per block of the per-row kernels (LDS: 60.6 KB)
constexpr size_t kLdsBytes = sizeof(float) * (kParamsLds + size_t(kH) * kRows);

static_assert(oW2 % 4 == 0 && oW3 % 4 == 0 && oW4 % 4 == 0, "weight rows must be float4 aligned");

// Round to the nearest bf16 value, ties to even (NaN passes through with its payload's top bits).
__device__ __host__ inline float round_bf16(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  if ((u & 0x7f800000u) != 0x7f800000u) u += 0x7fffu + ((u >> 16) & 1u);
  u &= 0xffff0000u;
  float r;
  memcpy(&r, &u, 4);
  return r;
}

__device__ inline void stage_params(const float* __restrict__ P, float* W, bool bf16) {
  for (int i = threadIdx.x; i < kParams; i += blockDim.x) W[i] = bf16 ? round_bf16(P[i]) : P[i];
}

// acc[j] = b[j] + sum_{i < in} H[i][t] W[i][j]
__device__ inline void dense(const float* W, const float* b, int in, const float* H, int t,
                             float (&acc)[kH]) {
#pragma unroll
  for (int j = 0; j < kH; ++j) acc[j] = b[j];
  for (int i = 0; i < in; ++i) {
    const float a = H[i * kRows + t];
    const float4* w = reinterpret_cast<const float4*>(W + i * kH);
#pragma unroll
    for (int j4 = 0; j4 < kH / 4; ++j4) {
      const float4 v = w[j4];
      acc[4 * j4 + 0] = fmaf(a, v.x, acc[4 * j4 + 0]);
      acc[4 * j4 + 1] = fmaf(a, v.y, acc[4 * j4 + 1]);
      acc[4 * j4 + 2] = fmaf(a, v.z, acc[4 * j4 + 2]);
      acc[4 * j4 + 3] = fmaf(a, v.w, acc[4 * j4 + 3]);
    }
  }
}

template <bool kSave, bool kBf16>
__global__ __launch_bounds__(kRows) void forward_kernel(const float* __restrict__ P,
                                                        const float* __restrict__ X, uint32_t rows,
                                                        uint32_t cap, float* __restrict__ Y,
                                                        float* __restrict__ act) {
  extern __shared__ float4 smem4[];
  float* W = reinterpret_cast<float*>(smem4);
  float* H = W + kParamsLds;
  const int t = threadIdx.x;
  const uint32_t r = blockIdx.x * kRows + t;
  const bool live = r < rows;
  stage_params(P, W, kBf16);
  for (int i = 0; i < kIn; ++i) {
    const float x = live ? X[size_t(r) * kIn + i] : 0.f;
    H[i * kRows + t] = kBf16 ? round_bf16(x) : x;
  }
  __syncthreads();
  float acc[kH];
  const int woff[3] = {oW1, oW2, oW3}, boff[3] = {oB1, oB2, oB3}, ins[3] = {kIn, kH, kH};
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    dense(W + woff[l], W + boff[l], ins[l], H, t, acc);
#pragma unroll
    for (int j = 0; j < kH; ++j) {
      float h = fmaxf(acc[j], 0.f);
      if (kBf16) h = round_bf16(h);
      H[j * kRows + t] = h;  // own column only: no other thread reads it
      if (kSave && r < cap) act[(size_t(l) * kH + j) * cap + r] = live ? h : 0.f;
    }
  }
  float y[kOut];
#pragma unroll
  for (int j = 0; j < kOut; ++j) y[j] = W[oB4 + j];
  for (int i = 0; i < kH; ++i) {
    const float a = H[i * kRows + t];
#pragma unroll
    for (int j = 0; j < kOut; ++j) y[j] = fmaf(a, W[oW4 + i * kOut + j], y[j]);
  }
  if (live) {
#pragma unroll
    for (int j = 0; j < kOut; ++j) Y[size_t(r) * kOut + j] = y[j];
  }
}

// Output deltas from the loss (scaled by 1 / sum of weights) and hidden-layer deltas.
__global__ __launch_bounds__(kRows) void backward_kernel(const float* __restrict__ P,
                                                         const float* __restrict__ T, uint32_t rows,
                                                         uint32_t cap, float inv_wsum,
                                                         float depth_weight,
                                                         const float* __restrict__ Y,
                                                         const float* __restrict__ act,
                                                         float* __restrict__ delta,
                                                         float* __restrict__ loss_part) {
  extern __shared__ float4 smem4[];
  float* W = reinterpret_cast<float*>(smem4);
  float* Dz = W + kParamsLds;  // [64][kRows] deltas of the layer above, own column per thread
  __shared__ float red[kRows];
  const int t = threadIdx.x;
  const uint32_t r = blockIdx.x * kRows + t;
  const bool live = r < rows;
  stage_params(P, W, false);
  __syncthreads();
  float dy[kOut];
  float loss = 0.f;
  if (live) {
    const float* y = Y + size_t(r) * kOut;
    const float* tr = T + size_t(r) * BMS2_MLP_TARGETS;
    const float w = tr[3] * inv_wsum;
    const float e0 = y[0] - tr[0], e1 = y[1] - tr[1];
    dy[0] = 2.f * e0 * w;
    dy[1] = 2.f * e1 * w;
    int tk = int(tr[2]);
    tk = tk < 0 ? 0 : (tk > 7 ? 7 : tk);
    float mx = y[2];
#pragma unroll
    for (int k = 1; k < 8; ++k) mx = fmaxf(mx, y[2 + k]);
    float ex[8], se = 0.f;
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      ex[k] = expf(y[2 + k] - mx);
      se += ex[k];
    }
#pragma unroll
    for (int k = 0; k < 8; ++k) dy[2 + k] = depth_weight * w * (ex[k] / se - (k == tk ? 1.f : 0.f));
    const float ce = logf(se) + mx - y[2 + tk];
    loss = w * (e0 * e0 + e1 * e1 + depth_weight * ce);
  } else {
#pragma unroll
    for (int j = 0; j < kOut; ++j) dy[j] = 0.f;
  }
#pragma unroll
  for (int j = 0; j < kH; ++j) delta[(size_t(3) * kH + j) * cap + r] = j < kOut ? dy[j] : 0.f;
  // hidden layer 3: dz3[i] = [h3_i > 0] sum_j dy[j] W4[i][j]
  for (int i = 0; i < kH; ++i) {
    float s = 0.f;
#pragma unroll
    for (int j = 0; j < kOut; ++j) s = fmaf(dy[j], W[oW4 + i * kOut + j], s);
    const float h = act[(size_t(2) * kH + i) * cap + r];
    const float dz = h > 0.f ? s : 0.f;
    Dz[i * kRows + t] = dz;
    delta[(size_t(2) * kH + i) * cap + r] = dz;
  }
  // hidden layers 2 and 1: dz_l[i] = [h_l,i > 0] sum_j dz_{l+1}[j] W_{l+1}[i][j]
  const int woff[2] = {oW3, oW2};
#pragma unroll
  for (int s2 = 0; s2 < 2; ++s2) {
    const int l = 1 - s2;  // 1 (layer 2), then 0 (layer 1)
    float up[kH];
#pragma unroll
    for (int j = 0; j < kH; ++j) up[j] = Dz[j * kRows + t];
    for (int i = 0; i < kH; ++i) {
      const float4* w = reinterpret_cast<const float4*>(W + woff[s2] + i * kH);
      float s = 0.f;
#pragma unroll
      for (int j4 = 0; j4 < kH / 4; ++j4) {
        const float4 v = w[j4];
        s = fmaf(up[4 * j4 + 0], v.x, s);
        s = fmaf(up[4 * j4 + 1], v.y, s);
        s = fmaf(up[4 * j4 + 2], v.z, s);
        s = fmaf(up[4 * j4 + 3], v.w, s);
      }
      const float h = act[(size_t(l) * kH + i) * cap + r];
      const float dz = h > 0.f ? s : 0.f;
      Dz[i * kRows + t] = dz;  // `up` already holds the layer above
      delta[(size_t(l) * kH + i) * cap + r] = dz;
    }
  }
  // block loss partial in a fixed order
  red[t] = loss;
  __syncthreads();
  for (int off = kRows / 2; off > 0; off >>= 1) {
    if (t < off) red[t] += red[t + off];
    __syncthreads();
  }
  if (t == 0) loss_part[blockIdx.x] = red[0];
}

// G[p] = sum_r a_i(r) dz_j(r) for weight p = (i, j) of layer l (a = 1 for biases), rows in order.
__global__ void wgrad_kernel(const float* __restrict__ X, uint32_t rows, uint32_t cap,
                             const float* __restrict__ act, const float* __restrict__ delta,
                             float* __restrict__ G) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kParams) return;
  int l, i, j, out;
  bool bias;
  if (p < oB1) { l = 0; out = kH; i = p / kH; j = p % kH; bias = false; }
  else if (p < oW2) { l = 0; out = kH; i = 0; j = p - oB1; bias = true; }
  else if (p < oB2) { l = 1; out = kH; i = (p - oW2) / kH; j = (p - oW2) % kH; bias = false; }
  else if (p < oW3) { l = 1; out = kH; i = 0; j = p - oB2; bias = true; }
  else if (p < oB3) { l = 2; out = kH; i = (p - oW3) / kH; j = (p - oW3) % kH; bias = false; }
  else if (p < oW4) { l = 2; out = kH; i = 0; j = p - oB3; bias = true; }
  else if (p < oB4) { l = 3; out = kOut; i = (p - oW4) / kOut; j = (p - oW4) % kOut; bias = false; }
  else { l = 3; out = kOut; i = 0; j = p - oB4; bias = true; }
  (void)out;
  const float* dz = delta + (size_t(l) * kH + j) * cap;
  // f32 products summed in f64: a batch of thousands of rows with cancelling terms keeps the
  // gradient accurate to f32 rounding of its inputs
  double s = 0.0;
  if (bias) {
    for (uint32_t r = 0; r < rows; ++r) s += double(dz[r]);
  } else if (l == 0) {
    for (uint32_t r = 0; r < rows; ++r) s += double(X[size_t(r) * kIn + i] * dz[r]);
  } else {
    const float* a = act + (size_t(l - 1) * kH + i) * cap;
    for (uint32_t r = 0; r < rows; ++r) s += double(a[r] * dz[r]);
  }
  G[p] = float(s);
}

__global__ void adam_kernel(float* __restrict__ P, const float* __restrict__ G, float* __restrict__ M,
                            float* __restrict__ V, float lr, float b1, float b2, float eps,
                            float c1, float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= kParams) return;
  const float g = G[p];
  const float m = b1 * M[p] + (1.f - b1) * g;
  const float v = b2 * V[p] + (1.f - b2) * g * g;
  M[p] = m;
  V[p] = v;
  P[p] -= lr * (m / c1) / (sqrtf(v / c2) + eps);
}

uint64_t splitmix64(uint64_t& s) {
  uint64_t z = (s += 0x9e3779b97f4a7c15ull);
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

int ensure_capacity(bms2_mlp_s* m, uint32_t rows) {
  const uint32_t need = (rows + kRows - 1) / kRows * kRows;
  if (need <= m->cap) return BMS2_OK;
  (void)hipStreamSynchronize(m->q->stream);
  (void)hipFree(m->d_act);
  (void)hipFree(m->d_delta);
  (void)hipFree(m->d_y);
  (void)hipFree(m->d_loss);
  m->d_act = m->d_delta = m->d_y = m->d_loss = nullptr;
  m->cap = 0;
  if (int rc = device_alloc((void**)&m->d_act, 4ull * 3 * kH * need)) return rc;
  if (int rc = device_alloc((void**)&m->d_delta, 4ull * 4 * kH * need)) return rc;
  if (int rc = device_alloc((void**)&m->d_y, 4ull * kOut * need)) return rc;
  if (int rc = device_alloc((void**)&m->d_loss, 4ull * (need / kRows))) return rc;
  m->cap = need;
  return BMS2_OK;
}

// Forward (saving activations), backward and weight gradient; the loss is summed on the host
// from the block partials in block order.
int loss_grad(bms2_mlp_s* m, const float* d_x, const float* d_t, uint32_t rows,
              float depth_weight, float* loss) {
  hipStream_t s = m->q->stream;
  if (int rc = ensure_capacity(m, rows)) return rc;
  // sum of row weights (host side: targets are small)
  std::vector<float> t(size_t(rows) * BMS2_MLP_TARGETS);
  BMS2_HIP(hipMemcpyAsync(t.data(), d_t, 4ull * t.size(), hipMemcpyDeviceToHost, s));
  BMS2_HIP(hipStreamSynchronize(s));
  double wsum = 0;
  for (uint32_t r = 0; r < rows; ++r) wsum += t[size_t(r) * BMS2_MLP_TARGETS + 3];
  if (!(wsum > 0)) return set_error(BMS2_ERR_INVALID, "row weights must sum to a positive value");
  const uint32_t blocks = m->cap / kRows;
  hipLaunchKernelGGL((forward_kernel<true, false>), dim3(blocks), dim3(kRows), kLdsBytes, s, m->d_p,
                     d_x, rows, m->cap, m->d_y, m->d_act);
  BMS2_CHECK_LAUNCH("forward_kernel");
  hipLaunchKernelGGL(backward_kernel, dim3(blocks), dim3(kRows), kLdsBytes, s, m->d_p, d_t, rows,
                     m->cap, float(1.0 / wsum), depth_weight, m->d_y, m->d_act, m->d_delta,
                     m->d_loss);
  BMS2_CHECK_LAUNCH("backward_kernel");
  hipLaunchKernelGGL(wgrad_kernel, dim3((kParams + 255) / 256), dim3(256), 0, s, d_x, rows,
                     m->cap, m->d_act, m->d_delta, m->d_g);
  BMS2_CHECK_LAUNCH("wgrad_kernel");
  if (loss) {
    std::vector<float> part(blocks);
    BMS2_HIP(hipMemcpyAsync(part.data(), m->d_loss, 4ull * blocks, hipMemcpyDeviceToHost, s));
    BMS2_HIP(hipStreamSynchronize(s));
    double acc = 0;
    for (float v : part) acc += v;
    *loss = float(acc);
  }
  return BMS2_OK;
}

}  // namespace
}  // namespace bms2

using namespace bms2;

extern "C" uint64_t bms2_mlp_num_params(void) { return kParams; }

extern "C" int bms2_mlp_destroy(bms2_mlp m) {
  if (!m) return BMS2_OK;
  if (int rc = use_device(m->q)) return rc;
  (void)hipStreamSynchronize(m->q->stream);
  for (float* p : {m->d_p, m->d_g, m->d_m, m->d_v, m->d_act, m->d_delta, m->d_y, m->d_loss})
    (void)hipFree(p);
  delete m;
  return BMS2_OK;
}

extern "C" int bms2_mlp_create(bms2_queue q, uint64_t seed, bms2_mlp* out) {
  if (int rc = use_device(q)) return rc;
  if (!out) return set_error(BMS2_ERR_INVALID, "null output");
  auto* m = new bms2_mlp_s{};
  m->q = q;
  int rc = BMS2_OK;
  if ((rc = device_alloc((void**)&m->d_p, 4ull * kParams)) ||
      (rc = device_alloc((void**)&m->d_g, 4ull * kParams)) ||
      (rc = device_alloc((void**)&m->d_m, 4ull * kParams)) ||
      (rc = device_alloc((void**)&m->d_v, 4ull * kParams))) {
    bms2_mlp_destroy(m);
    return rc;
  }
  std::vector<float> p(kParams, 0.f);
  uint64_t st = seed;
  auto init = [&](int off, int in, int outn) {
    const double lim = std::sqrt(6.0 / in);
    for (int k = 0; k < in * outn; ++k) {
      const double u = double(splitmix64(st) >> 11) * (1.0 / 9007199254740992.0);
      p[off + k] = float((2.0 * u - 1.0) * lim);
    }
  };
  init(oW1, kIn, kH);
  init(oW2, kH, kH);
  init(oW3, kH, kH);
  init(oW4, kH, kOut);
  if ((rc = bms2_mlp_set_params(m, p.data(), kParams))) {
    bms2_mlp_destroy(m);
    return rc;
  }
  *out = m;
  return BMS2_OK;
}

extern "C" int bms2_mlp_get_params(bms2_mlp m, float* host, uint64_t n) {
  if (!m || !host) return set_error(BMS2_ERR_INVALID, "null argument");
  if (n != uint64_t(kParams)) return set_error(BMS2_ERR_INVALID, "expected %d parameters", kParams);
  if (int rc = use_device(m->q)) return rc;
  BMS2_HIP(hipMemcpyAsync(host, m->d_p, 4ull * kParams, hipMemcpyDeviceToHost, m->q->stream));
  BMS2_HIP(hipStreamSynchronize(m->q->stream));
  return BMS2_OK;
}

extern "C" int bms2_mlp_set_params(bms2_mlp m, const float* host, uint64_t n) {
  if (!m || !host) return set_error(BMS2_ERR_INVALID, "null argument");
  if (n != uint64_t(kParams)) return set_error(BMS2_ERR_INVALID, "expected %d parameters", kParams);
  if (int rc = use_device(m->q)) return rc;
  hipStream_t s = m->q->stream;
  BMS2_HIP(hipMemcpyAsync(m->d_p, host, 4ull * kParams, hipMemcpyHostToDevice, s));
  BMS2_HIP(hipMemsetAsync(m->d_m, 0, 4ull * kParams, s));
  BMS2_HIP(hipMemsetAsync(m->d_v, 0, 4ull * kParams, s));
  BMS2_HIP(hipStreamSynchronize(s));
  m->step = 0;
  return BMS2_OK;
}

extern "C" int bms2_mlp_forward(bms2_mlp m, const float* d_x, uint32_t rows, float* d_y, int bf16) {
  if (!m || (rows && (!d_x || !d_y))) return set_error(BMS2_ERR_INVALID, "null argument");
  if (int rc = use_device(m->q)) return rc;
  if (!rows) return BMS2_OK;
  const uint32_t blocks = (rows + kRows - 1) / kRows;
  hipStream_t s = m->q->stream;
  if (bf16)
    hipLaunchKernelGGL((forward_kernel<false, true>), dim3(blocks), dim3(kRows), kLdsBytes, s,
                       m->d_p, d_x, rows, 0u, d_y, nullptr);
  else
    hipLaunchKernelGGL((forward_kernel<false, false>), dim3(blocks), dim3(kRows), kLdsBytes, s,
                       m->d_p, d_x, rows, 0u, d_y, nullptr);
  BMS2_CHECK_LAUNCH("forward_kernel");
  return BMS2_OK;
}

extern "C" int bms2_mlp_loss_grad(bms2_mlp m, const float* d_x, const float* d_t, uint32_t rows,
                                  float depth_weight, float* loss, float* grad_host) {
  if (!m || !d_x || !d_t || !rows) return set_error(BMS2_ERR_INVALID, "null argument or no rows");
  if (int rc = use_device(m->q)) return rc;
  if (int rc = loss_grad(m, d_x, d_t, rows, depth_weight, loss)) return rc;
  if (grad_host) {
    BMS2_HIP(hipMemcpyAsync(grad_host, m->d_g, 4ull * kParams, hipMemcpyDeviceToHost, m->q->stream));
    BMS2_HIP(hipStreamSynchronize(m->q->stream));
  }
  return BMS2_OK;
}

extern "C" int bms2_mlp_train_step(bms2_mlp m, const float* d_x, const float* d_t, uint32_t rows,
                                   float depth_weight, const bms2_adam* opt, float* loss) {
  if (!m || !d_x || !d_t || !rows || !opt) return set_error(BMS2_ERR_INVALID, "null argument or no rows");
  if (int rc = use_device(m->q)) return rc;
  if (int rc = loss_grad(m, d_x, d_t, rows, depth_weight, loss)) return rc;
  m->step += 1;
  const float c1 = float(1.0 - std::pow(double(opt->beta1), double(m->step)));
  const float c2 = float(1.0 - std::pow(double(opt->beta2), double(m->step)));
  hipLaunchKernelGGL(adam_kernel, dim3((kParams + 255) / 256), dim3(256), 0, m->q->stream, m->d_p,
                     m->d_g, m->d_m, m->d_v, opt->lr, opt->beta1, opt->beta2, opt->eps, c1, c2);
  BMS2_CHECK_LAUNCH("adam_kernel");
  return BMS2_OK;
}
