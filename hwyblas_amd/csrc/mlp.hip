// Training and inference of the grouped-GEMM cost model (mlp.h) on the GPU,
// with hand-written kernels:
//  - ForwardKernel: one thread per sample, parameters staged in LDS,
//    activations in registers; optionally keeps the hidden activations;
//  - BackwardKernel: one thread per sample, the loss and the error signal
//    of every layer (ReLU derivative: h > 0);
//  - WeightGradKernel: one thread per weight (and per bias), the sum over
//    the samples of error x input, in sample order (deterministic);
//  - AdamKernel: the Adam update (Kingma and Ba, 2015) with bias correction.
// Loss per sample s: 0.5 (out[s][cfg_s] - y_s)^2, averaged over a batch;
// only the output of the configuration that was timed is trained.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <vector>

#include "mlp.h"

namespace hwy_amd {
namespace mlp {
namespace {

constexpr int kThreads = 256;
constexpr int kLdsParams = NumParams(kMaxOutputs);

__device__ __forceinline__ void StageParams(const float* __restrict__ params, int np, float* w) {
  for (int i = threadIdx.x; i < np; i += blockDim.x) w[i] = params[i];
  __syncthreads();
}

// y = relu(W x + b), W row-major [kWidth][IN], sums in input order.
template <int IN>
__device__ __forceinline__ void DenseRelu(const float* W, const float* b, const float (&x)[IN],
                                          float (&y)[kWidth]) {
#pragma unroll
  for (int j = 0; j < kWidth; ++j) {
    float acc = b[j];
#pragma unroll
    for (int i = 0; i < IN; ++i) acc = __builtin_fmaf(W[j * IN + i], x[i], acc);
    y[j] = acc > 0.0f ? acc : 0.0f;
  }
}

template <int N>
__device__ __forceinline__ void StoreRow(float* dst, const float (&v)[N]) {
#pragma unroll
  for (int j = 0; j < N; ++j) dst[j] = v[j];
}

template <int N>
__device__ __forceinline__ void LoadRow(const float* src, float (&v)[N]) {
#pragma unroll
  for (int j = 0; j < N; ++j) v[j] = src[j];
}

__global__ void __launch_bounds__(kThreads)
    ForwardKernel(const float* __restrict__ params, int outputs, const float* __restrict__ x,
                  int n, float* __restrict__ h1, float* __restrict__ h2, float* __restrict__ h3,
                  float* __restrict__ out) {
  __shared__ float w[kLdsParams];
  StageParams(params, NumParams(outputs), w);
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  float in[kInputs];
  LoadRow(x + int64_t(s) * kInputs, in);
  float a[kWidth], b[kWidth];
  DenseRelu<kInputs>(w + OffW1(), w + OffB1(), in, a);
  if (h1) StoreRow(h1 + int64_t(s) * kWidth, a);
  DenseRelu<kWidth>(w + OffW2(), w + OffB2(), a, b);
  if (h2) StoreRow(h2 + int64_t(s) * kWidth, b);
  DenseRelu<kWidth>(w + OffW3(), w + OffB3(), b, a);
  if (h3) StoreRow(h3 + int64_t(s) * kWidth, a);
  for (int c = 0; c < outputs; ++c) {
    const float* wc = w + OffW4() + c * kWidth;
    float acc = w[OffB4(outputs) + c];
#pragma unroll
    for (int j = 0; j < kWidth; ++j) acc = __builtin_fmaf(wc[j], a[j], acc);
    out[int64_t(s) * outputs + c] = acc;
  }
}

// Error signals d_l = dLoss/dz_l (z_l: pre-activation of layer l) for a
// batch whose loss is scale * sum_s 0.5 r_s^2.
__global__ void __launch_bounds__(kThreads)
    BackwardKernel(const float* __restrict__ params, int outputs, const float* __restrict__ h1,
                   const float* __restrict__ h2, const float* __restrict__ h3,
                   const float* __restrict__ out, const int* __restrict__ cfg,
                   const float* __restrict__ y, int n, float scale, float* __restrict__ d1,
                   float* __restrict__ d2, float* __restrict__ d3, float* __restrict__ d4,
                   float* __restrict__ loss) {
  __shared__ float w[kLdsParams];
  StageParams(params, NumParams(outputs), w);
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int c = cfg[s];
  const float r = out[int64_t(s) * outputs + c] - y[s];
  loss[s] = 0.5f * r * r * scale;
  const float g = r * scale;
  for (int cc = 0; cc < outputs; ++cc) d4[int64_t(s) * outputs + cc] = cc == c ? g : 0.0f;
  float hv[kWidth], da[kWidth], db[kWidth];
  // Layer 3 (only output c carries an error).
  LoadRow(h3 + int64_t(s) * kWidth, hv);
  const float* w4c = w + OffW4() + c * kWidth;
#pragma unroll
  for (int j = 0; j < kWidth; ++j) da[j] = hv[j] > 0.0f ? w4c[j] * g : 0.0f;
  StoreRow(d3 + int64_t(s) * kWidth, da);
  // Layer 2: d2 = (W3^T d3) * relu'(z2).
  LoadRow(h2 + int64_t(s) * kWidth, hv);
#pragma unroll
  for (int i = 0; i < kWidth; ++i) {
    float acc = 0.0f;
#pragma unroll
    for (int j = 0; j < kWidth; ++j) acc = __builtin_fmaf(w[OffW3() + j * kWidth + i], da[j], acc);
    db[i] = hv[i] > 0.0f ? acc : 0.0f;
  }
  StoreRow(d2 + int64_t(s) * kWidth, db);
  // Layer 1: d1 = (W2^T d2) * relu'(z1).
  LoadRow(h1 + int64_t(s) * kWidth, hv);
#pragma unroll
  for (int i = 0; i < kWidth; ++i) {
    float acc = 0.0f;
#pragma unroll
    for (int j = 0; j < kWidth; ++j) acc = __builtin_fmaf(w[OffW2() + j * kWidth + i], db[j], acc);
    da[i] = hv[i] > 0.0f ? acc : 0.0f;
  }
  StoreRow(d1 + int64_t(s) * kWidth, da);
}

// gW[o][i] = sum_s D[s][o] X[s][i]; gb[o] = sum_s D[s][o].
__global__ void __launch_bounds__(kThreads)
    WeightGradKernel(const float* __restrict__ d, int dout, const float* __restrict__ x, int din,
                     int n, float* __restrict__ gw, float* __restrict__ gb) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < dout * din) {
    const int o = t / din, i = t % din;
    float acc = 0.0f;
    for (int s = 0; s < n; ++s)
      acc = __builtin_fmaf(d[int64_t(s) * dout + o], x[int64_t(s) * din + i], acc);
    gw[t] = acc;
  } else if (t < dout * din + dout) {
    const int o = t - dout * din;
    float acc = 0.0f;
    for (int s = 0; s < n; ++s) acc += d[int64_t(s) * dout + o];
    gb[o] = acc;
  }
}

__global__ void AdamKernel(float* __restrict__ p, const float* __restrict__ g,
                           float* __restrict__ m, float* __restrict__ v, int count, float lr,
                           float beta1, float beta2, float eps, float corr1, float corr2) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= count) return;
  const float gi = g[i];
  const float mi = beta1 * m[i] + (1.0f - beta1) * gi;
  const float vi = beta2 * v[i] + (1.0f - beta2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / corr1) / (sqrtf(vi / corr2) + eps);
}

__global__ void GatherKernel(const float* __restrict__ x, const int* __restrict__ cfg,
                             const float* __restrict__ y, const int* __restrict__ idx, int n,
                             float* __restrict__ xb, int* __restrict__ cfgb,
                             float* __restrict__ yb) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= n) return;
  const int j = idx[s];
#pragma unroll
  for (int i = 0; i < kInputs; ++i) xb[int64_t(s) * kInputs + i] = x[int64_t(j) * kInputs + i];
  cfgb[s] = cfg[j];
  yb[s] = y[j];
}

unsigned Blocks(int64_t n) { return unsigned(std::max<int64_t>(1, (n + kThreads - 1) / kThreads)); }

uint64_t SplitMix(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

// Device buffers of one forward/backward pass over up to `cap` samples.
struct Pass {
  int cap = 0, outputs = 0;
  float *x = nullptr, *y = nullptr, *h1 = nullptr, *h2 = nullptr, *h3 = nullptr;
  float *out = nullptr, *d1 = nullptr, *d2 = nullptr, *d3 = nullptr, *d4 = nullptr;
  float* loss = nullptr;
  int* cfg = nullptr;

  Pass(int n, int outs) : cap(std::max(n, 1)), outputs(outs) {
    auto f = [&](float** p, int64_t per) {
      HWY_HIP_CHECK(hipMalloc(p, size_t(cap) * per * sizeof(float)));
    };
    f(&x, kInputs);
    f(&y, 1);
    f(&h1, kWidth);
    f(&h2, kWidth);
    f(&h3, kWidth);
    f(&out, outputs);
    f(&d1, kWidth);
    f(&d2, kWidth);
    f(&d3, kWidth);
    f(&d4, outputs);
    f(&loss, 1);
    HWY_HIP_CHECK(hipMalloc(&cfg, size_t(cap) * sizeof(int)));
  }
  ~Pass() {
    for (float* p : {x, y, h1, h2, h3, out, d1, d2, d3, d4, loss}) (void)hipFree(p);
    (void)hipFree(cfg);
  }
};

// Forward, backward and the parameter gradients (into model.grads) of the
// samples in pass.{x, cfg, y}; loss per sample into pass.loss.
void Gradients(const Model& model, Pass& pass, int n, hipStream_t s) {
  const int o = model.outputs;
  hipLaunchKernelGGL(ForwardKernel, dim3(Blocks(n)), dim3(kThreads), 0, s, model.params, o,
                     pass.x, n, pass.h1, pass.h2, pass.h3, pass.out);
  hipLaunchKernelGGL(BackwardKernel, dim3(Blocks(n)), dim3(kThreads), 0, s, model.params, o,
                     pass.h1, pass.h2, pass.h3, pass.out, pass.cfg, pass.y, n, 1.0f / float(n),
                     pass.d1, pass.d2, pass.d3, pass.d4, pass.loss);
  float* g = model.grads;
  auto wgrad = [&](const float* d, int dout, const float* in, int din, int off_w, int off_b) {
    hipLaunchKernelGGL(WeightGradKernel, dim3(Blocks(int64_t(dout) * (din + 1))), dim3(kThreads),
                       0, s, d, dout, in, din, n, g + off_w, g + off_b);
  };
  wgrad(pass.d1, kWidth, pass.x, kInputs, OffW1(), OffB1());
  wgrad(pass.d2, kWidth, pass.h1, kWidth, OffW2(), OffB2());
  wgrad(pass.d3, kWidth, pass.h2, kWidth, OffW3(), OffB3());
  wgrad(pass.d4, o, pass.h3, kWidth, OffW4(), OffB4(o));
  HWY_HIP_CHECK(hipGetLastError());
}

std::atomic<Model*>& Installed() {
  static std::atomic<Model*> m{nullptr};
  return m;
}

bool SamplesOk(const Model* m, int n, const float* feats, const int* cfg, const float* y) {
  if (!m || n < 0 || (n > 0 && (!feats || !cfg || !y))) return false;
  for (int i = 0; i < n; ++i)
    if (cfg[i] < 0 || cfg[i] >= m->outputs) return false;
  return true;
}

}  // namespace

const Model* InstalledModel() { return Installed().load(); }

void Predict(const Model& model, int n, const float* features, float* out, hipStream_t s) {
  if (n <= 0) return;
  float *dx = nullptr, *dout = nullptr;
  HWY_HIP_CHECK(hipMalloc(&dx, size_t(n) * kInputs * sizeof(float)));
  HWY_HIP_CHECK(hipMalloc(&dout, size_t(n) * model.outputs * sizeof(float)));
  HWY_HIP_CHECK(hipMemcpyAsync(dx, features, size_t(n) * kInputs * sizeof(float),
                               hipMemcpyHostToDevice, s));
  hipLaunchKernelGGL(ForwardKernel, dim3(Blocks(n)), dim3(kThreads), 0, s, model.params,
                     model.outputs, dx, n, nullptr, nullptr, nullptr, dout);
  HWY_HIP_CHECK(hipGetLastError());
  HWY_HIP_CHECK(hipMemcpyAsync(out, dout, size_t(n) * model.outputs * sizeof(float),
                               hipMemcpyDeviceToHost, s));
  HWY_HIP_CHECK(hipStreamSynchronize(s));
  HWY_HIP_CHECK(hipFree(dx));
  HWY_HIP_CHECK(hipFree(dout));
}

}  // namespace mlp
}  // namespace hwy_amd

using namespace hwy_amd;
using namespace hwy_amd::mlp;

namespace {
hipStream_t StreamOr(void* stream) {
  return stream ? static_cast<hipStream_t>(stream) : Context().stream;
}
}  // namespace

extern "C" {

HWYBLAS_AMD_API void* hwyblas_mlp_create(int outputs, uint64_t seed) {
  if (outputs < 1 || outputs > kMaxOutputs) return nullptr;
  Model* m = new Model();
  m->outputs = outputs;
  HWY_HIP_CHECK(hipGetDevice(&m->device));
  const int np = NumParams(outputs);
  for (float** p : {&m->params, &m->grads, &m->adam_m, &m->adam_v})
    HWY_HIP_CHECK(hipMalloc(p, size_t(np) * sizeof(float)));
  // He-uniform weights (bound sqrt(6 / fan_in)), zero biases.
  std::vector<float> host(np, 0.0f);
  uint64_t state = SplitMix(seed ^ 0x243F6A8885A308D3ull);
  auto fill = [&](int off, int rows, int cols) {
    const float bound = std::sqrt(6.0f / float(cols));
    for (int i = 0; i < rows * cols; ++i) {
      state = SplitMix(state);
      const float u = float((state >> 40) * (1.0 / 16777216.0));  // [0, 1)
      host[off + i] = (2.0f * u - 1.0f) * bound;
    }
  };
  fill(OffW1(), kWidth, kInputs);
  fill(OffW2(), kWidth, kWidth);
  fill(OffW3(), kWidth, kWidth);
  fill(OffW4(), outputs, kWidth);
  HWY_HIP_CHECK(hipMemcpy(m->params, host.data(), size_t(np) * sizeof(float),
                          hipMemcpyHostToDevice));
  std::vector<float> zeros(np, 0.0f);
  HWY_HIP_CHECK(hipMemcpy(m->adam_m, zeros.data(), size_t(np) * sizeof(float),
                          hipMemcpyHostToDevice));
  HWY_HIP_CHECK(hipMemcpy(m->adam_v, zeros.data(), size_t(np) * sizeof(float),
                          hipMemcpyHostToDevice));
  return m;
}

HWYBLAS_AMD_API void hwyblas_mlp_destroy(void* model) {
  Model* m = static_cast<Model*>(model);
  if (!m) return;
  Model* expected = m;
  Installed().compare_exchange_strong(expected, nullptr);
  for (float* p : {m->params, m->grads, m->adam_m, m->adam_v}) (void)hipFree(p);
  delete m;
}

HWYBLAS_AMD_API int hwyblas_mlp_num_params(void* model) {
  return model ? NumParams(static_cast<Model*>(model)->outputs) : 0;
}

HWYBLAS_AMD_API int hwyblas_mlp_outputs(void* model) {
  return model ? static_cast<Model*>(model)->outputs : 0;
}

HWYBLAS_AMD_API void hwyblas_mlp_get_params(void* model, float* params) {
  const Model* m = static_cast<Model*>(model);
  HWY_HIP_CHECK(hipMemcpy(params, m->params, size_t(NumParams(m->outputs)) * sizeof(float),
                          hipMemcpyDeviceToHost));
}

HWYBLAS_AMD_API void hwyblas_mlp_set_params(void* model, const float* params) {
  Model* m = static_cast<Model*>(model);
  const size_t bytes = size_t(NumParams(m->outputs)) * sizeof(float);
  HWY_HIP_CHECK(hipMemcpy(m->params, params, bytes, hipMemcpyHostToDevice));
  std::vector<float> zeros(NumParams(m->outputs), 0.0f);
  HWY_HIP_CHECK(hipMemcpy(m->adam_m, zeros.data(), bytes, hipMemcpyHostToDevice));
  HWY_HIP_CHECK(hipMemcpy(m->adam_v, zeros.data(), bytes, hipMemcpyHostToDevice));
  m->step = 0;
}

HWYBLAS_AMD_API int hwyblas_mlp_predict(void* model, int n, const float* features, float* out,
                                        void* stream) {
  const Model* m = static_cast<Model*>(model);
  if (!m || n < 0 || (n > 0 && (!features || !out))) return -1;
  Predict(*m, n, features, out, StreamOr(stream));
  return 0;
}

// Loss (mean over the samples) and its gradient w.r.t. every parameter.
HWYBLAS_AMD_API int hwyblas_mlp_gradients(void* model, int n, const float* features,
                                          const int* cfg, const float* target, float* grads,
                                          float* loss, void* stream) {
  Model* m = static_cast<Model*>(model);
  if (!SamplesOk(m, n, features, cfg, target) || n == 0 || !grads) return -1;
  hipStream_t s = StreamOr(stream);
  Pass pass(n, m->outputs);
  HWY_HIP_CHECK(hipMemcpyAsync(pass.x, features, size_t(n) * kInputs * sizeof(float),
                               hipMemcpyHostToDevice, s));
  HWY_HIP_CHECK(hipMemcpyAsync(pass.cfg, cfg, size_t(n) * sizeof(int), hipMemcpyHostToDevice, s));
  HWY_HIP_CHECK(hipMemcpyAsync(pass.y, target, size_t(n) * sizeof(float), hipMemcpyHostToDevice,
                               s));
  Gradients(*m, pass, n, s);
  std::vector<float> l(n);
  HWY_HIP_CHECK(hipMemcpyAsync(grads, m->grads, size_t(NumParams(m->outputs)) * sizeof(float),
                               hipMemcpyDeviceToHost, s));
  HWY_HIP_CHECK(hipMemcpyAsync(l.data(), pass.loss, size_t(n) * sizeof(float),
                               hipMemcpyDeviceToHost, s));
  HWY_HIP_CHECK(hipStreamSynchronize(s));
  if (loss) {
    double sum = 0;
    for (float v : l) sum += v;
    *loss = float(sum);
  }
  return 0;
}

// Adam over shuffled mini-batches; epoch_loss[e] = mean mini-batch loss.
HWYBLAS_AMD_API int hwyblas_mlp_train(void* model, int n, const float* features, const int* cfg,
                                      const float* target, int epochs, int batch, float lr,
                                      uint64_t seed, float* epoch_loss, void* stream) {
  Model* m = static_cast<Model*>(model);
  if (!SamplesOk(m, n, features, cfg, target) || n == 0 || epochs < 0 || batch < 1 ||
      !(lr > 0.0f))
    return -1;
  hipStream_t s = StreamOr(stream);
  const int np = NumParams(m->outputs);
  float *dx = nullptr, *dy = nullptr;
  int *dcfg = nullptr, *didx = nullptr;
  HWY_HIP_CHECK(hipMalloc(&dx, size_t(n) * kInputs * sizeof(float)));
  HWY_HIP_CHECK(hipMalloc(&dy, size_t(n) * sizeof(float)));
  HWY_HIP_CHECK(hipMalloc(&dcfg, size_t(n) * sizeof(int)));
  HWY_HIP_CHECK(hipMalloc(&didx, size_t(n) * sizeof(int)));
  HWY_HIP_CHECK(hipMemcpyAsync(dx, features, size_t(n) * kInputs * sizeof(float),
                               hipMemcpyHostToDevice, s));
  HWY_HIP_CHECK(hipMemcpyAsync(dy, target, size_t(n) * sizeof(float), hipMemcpyHostToDevice, s));
  HWY_HIP_CHECK(hipMemcpyAsync(dcfg, cfg, size_t(n) * sizeof(int), hipMemcpyHostToDevice, s));
  const int bsz = std::min(batch, n);
  Pass pass(bsz, m->outputs);
  float* dloss = nullptr;
  HWY_HIP_CHECK(hipMalloc(&dloss, size_t(n) * sizeof(float)));
  std::vector<int> order(n);
  std::vector<float> losses(n);
  uint64_t state = SplitMix(seed);
  const float beta1 = 0.9f, beta2 = 0.999f, eps = 1e-8f;
  for (int e = 0; e < epochs; ++e) {
    for (int i = 0; i < n; ++i) order[i] = i;
    for (int i = n - 1; i > 0; --i) {  // Fisher-Yates
      state = SplitMix(state);
      std::swap(order[i], order[int(state % uint64_t(i + 1))]);
    }
    HWY_HIP_CHECK(hipMemcpyAsync(didx, order.data(), size_t(n) * sizeof(int),
                                 hipMemcpyHostToDevice, s));
    int nbatches = 0;
    for (int b0 = 0; b0 < n; b0 += bsz) {
      const int nb = std::min(bsz, n - b0);
      hipLaunchKernelGGL(GatherKernel, dim3(Blocks(nb)), dim3(kThreads), 0, s, dx, dcfg, dy,
                         didx + b0, nb, pass.x, pass.cfg, pass.y);
      Gradients(*m, pass, nb, s);
      HWY_HIP_CHECK(hipMemcpyAsync(dloss + b0, pass.loss, size_t(nb) * sizeof(float),
                                   hipMemcpyDeviceToDevice, s));
      ++m->step;
      const float corr1 = 1.0f - std::pow(beta1, float(m->step));
      const float corr2 = 1.0f - std::pow(beta2, float(m->step));
      hipLaunchKernelGGL(AdamKernel, dim3(Blocks(np)), dim3(kThreads), 0, s, m->params, m->grads,
                         m->adam_m, m->adam_v, np, lr, beta1, beta2, eps, corr1, corr2);
      HWY_HIP_CHECK(hipGetLastError());
      ++nbatches;
    }
    if (epoch_loss) {
      HWY_HIP_CHECK(hipMemcpyAsync(losses.data(), dloss, size_t(n) * sizeof(float),
                                   hipMemcpyDeviceToHost, s));
      HWY_HIP_CHECK(hipStreamSynchronize(s));
      double sum = 0;
      for (float v : losses) sum += v;
      epoch_loss[e] = float(sum / nbatches);
    }
  }
  HWY_HIP_CHECK(hipStreamSynchronize(s));
  for (void* p : {static_cast<void*>(dx), static_cast<void*>(dy), static_cast<void*>(dcfg),
                  static_cast<void*>(didx), static_cast<void*>(dloss)})
    HWY_HIP_CHECK(hipFree(p));
  return 0;
}

HWYBLAS_AMD_API void hwyblas_grouped_set_model(void* model) {
  Installed().store(static_cast<Model*>(model));
}

}  // extern "C"
