// GEMM dispatch: picks a tile configuration of the plain MFMA kernel
// (gemm_kernel.h) for a shape, plus the C := beta C kernel used by the
// quick returns, and the config-indexed entry points the autotuner uses.
#include <algorithm>

#include "common.h"
#include "gemm_configs.h"

namespace hwy_amd {

// part: 0 all of C, 1 its lower triangle, 2 its upper triangle; krange: a
// triangular factor's zero k-range skipped per tile (GemmParams::krange).
void GemmLaunchF64(int cfg, bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha,
                   const double* a, int64_t lda, const double* b, int64_t ldb, double beta,
                   double* c, int64_t ldc, hipStream_t stream, int part = 0,
                   int krange = 0, const int* gate = nullptr, int64_t ksplit = 0,
                   int64_t c_split = 0);
void GemmLaunchF32(int cfg, bool ta, bool tb, int64_t m, int64_t n, int64_t k, float alpha,
                   const float* a, int64_t lda, const float* b, int64_t ldb, float beta,
                   float* c, int64_t ldc, hipStream_t stream, int part = 0,
                   int krange = 0, int64_t ksplit = 0, int64_t c_split = 0);

namespace {

int64_t Tiles(int64_t m, int64_t n, int bm, int bn) {
  return ((m + bm - 1) / bm) * ((n + bn - 1) / bn);
}

}  // namespace

// Default choice without tuning data (from a sweep of the table on MI355X at
// m = n = k in 4096..16384, all transpositions): FP64 runs best on 128 x 128
// tiles with two waves per SIMD (~81 % of the FP64 MFMA peak), FP32 on
// 256 x 128 x 8 (~76 %). Shapes with too few tiles to fill the 256 CUs fall
// to 64 x 64 and then 32 x 32 single-wave tiles; those are latency-bound,
// and a 32-deep k slice halves their k steps (FP64 512^3: 37.0 -> 32.6 us,
// 64 x 4096 x 4096: 249 -> 215 us).
int ChooseRegularGemmConfig(int dtype_bytes, int64_t m, int64_t n, int64_t k) {
  if (dtype_bytes == 8) {
    if (Tiles(m, n, 128, 128) >= 512) return 8;
  } else {
    if (Tiles(m, n, 256, 128) >= 256) return 10;
    if (Tiles(m, n, 128, 128) >= 512) return 8;
  }
  if (Tiles(m, n, 64, 64) >= 256) return 0;
  return k > 16 ? 11 : 7;
}

// FP64: whether the DMA-fed 128 x 128 kernel beats the plain 64 x 64 one,
// by a per-CU time model of both fitted on MI355X (256 CUs; NN products of
// 2048^3, 3072^3, 1024 x 4096 x 4096, 2048 x 3072 x 3072, 1536 x 2048 x 2048,
// 4096 x 1024 x 1024 and k = 128 .. 256 shapes): the busiest CU holds j
// tiles; a pair of DMA tiles shares a CU in 256.5 ns per unit of k and a
// single one takes 154 ns, plus 20 us of pipeline fill; plain 64 x 64 tiles
// take 46 ns each per unit of k, plus 8 us. A tile count just past a
// multiple of the CU count is where the finer tiles win (3072^3: 576 DMA
// tiles, 3 on the busiest CU).
bool DmaF64Faster(int64_t m, int64_t n, int64_t k) {
  constexpr int64_t kCus = 256;
  const int64_t j15 = (Tiles(m, n, 128, 128) + kCus - 1) / kCus;
  const int64_t j0 = (Tiles(m, n, 64, 64) + kCus - 1) / kCus;
  const double t15 = 20e3 + double(k) * (double(j15 / 2) * 256.5 + double(j15 % 2) * 154.0);
  const double t0 = 8e3 + double(k) * 46.0 * double(j0);
  return t15 < t0;
}

// Large products take the DMA-fed kernel (MI355X, m = n = k = 16384, NN):
// FP64 on 128 x 128 x 16 tiles through a two-slice LDS ring, 86 % of the
// FP64 MFMA peak against 80 % for the best plain configuration; FP32 on
// 256 x 256 x 16 tiles of 16 waves, 86 % against 75 %. Calls it cannot take
// (and its edge strips) go to the plain rule above.
int ChooseGemmConfig(int dtype_bytes, bool ta, bool tb, int64_t m, int64_t n, int64_t k) {
  (void)ta;
  (void)tb;
  if (dtype_bytes == 8 && Tiles(m, n, 128, 128) >= 128) return DmaF64Faster(m, n, k) ? 15 : 0;
  if (dtype_bytes == 4 && Tiles(m, n, 256, 256) >= 256) return 16;
  return ChooseRegularGemmConfig(dtype_bytes, m, n, k);
}

template <>
void GemmDevice<double>(bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha,
                        const double* a, int64_t lda, const double* b, int64_t ldb,
                        double beta, double* c, int64_t ldc, hipStream_t stream) {
  const int cfg = ChooseGemmConfig(8, ta, tb, m, n, k);
  GemmLaunchF64(cfg, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream);
}

template <>
void GemmDevice<float>(bool ta, bool tb, int64_t m, int64_t n, int64_t k, float alpha,
                       const float* a, int64_t lda, const float* b, int64_t ldb, float beta,
                       float* c, int64_t ldc, hipStream_t stream) {
  const int cfg = ChooseGemmConfig(4, ta, tb, m, n, k);
  GemmLaunchF32(cfg, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream);
}

template <>
void GemmDeviceGated<double>(const int* gate, bool ta, bool tb, int64_t m, int64_t n,
                             int64_t k, double alpha, const double* a, int64_t lda,
                             const double* b, int64_t ldb, double beta, double* c, int64_t ldc,
                             hipStream_t stream) {
  const int cfg = ChooseGemmConfig(8, ta, tb, m, n, k);
  GemmLaunchF64(cfg, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream, 0, 0, gate);
}

namespace {

// C = alpha sum_s P_s + beta C (hwyBLAS's Store rules on the sum), the
// partials P_s (m x n, leading dimension m, s < splits, P_s at p + s m n)
// added in ascending s.
template <typename T>
__global__ void SplitSumKernel(const T* __restrict__ p, int splits, int64_t m, int64_t n,
                               T alpha, T beta, T* __restrict__ c, int64_t ldc) {
  const int64_t total = m * n, plane = m * n;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
       t += int64_t(gridDim.x) * blockDim.x) {
    T acc = p[t];
    for (int s = 1; s < splits; ++s) acc += p[t + s * plane];
    const int64_t i = t % m, j = t / m;
    T* cij = c + i + j * ldc;
    T v = acc * alpha;
    if (beta != T(0)) v = beta == T(1) ? v + *cij : __builtin_fma(*cij, beta, v);
    *cij = v;
  }
}

}  // namespace

int SplitKCount(int64_t m, int64_t n, int64_t k) {
  // Worth it when the 32 x 32 tiles alone cannot fill the device (about
  // 4096 single-wave tiles in all) and each split keeps at least 256 of k.
  const int64_t tiles = Tiles(m, n, 32, 32);
  if (tiles >= 2048 || k < 1024) return 1;
  const int64_t want = (4096 + tiles - 1) / tiles;
  return int(std::max<int64_t>(1, std::min<int64_t>({want, k / 256, kMaxSplitK})));
}

namespace {

// Products whose 128 x 128 tiles cover C exactly but give at most one per
// CU: the DMA-fed kernel over k ranges of at least 1024, as many ranges as
// keep every CU at two tiles or fewer (one tile alone runs a CU at about 60 %
// of the rate of two; a third would queue behind them).
int DmaSplitKCount(int64_t m, int64_t n, int64_t k) {
  if (m % 128 != 0 || n % 128 != 0 || k % 16 != 0 || k < 2048) return 1;
  const int64_t tiles = Tiles(m, n, 128, 128);
  if (tiles > 256) return 1;
  const int64_t want = 512 / tiles;
  return int(std::max<int64_t>(1, std::min<int64_t>({want, k / 1024, kMaxSplitK})));
}

// The k ranges of a split product (1: no split). `entry`: for the GEMM entry
// points, which split only where the 32 x 32 tiles alone cannot fill the
// device (SplitKCount), so that products of common sizes keep one k chain.
template <typename T>
int GemmSplitCount(int64_t m, int64_t n, int64_t k, int64_t capacity, bool* dma,
                   bool entry = false) {
  *dma = false;
  if (entry && SplitKCount(m, n, k) <= 1) return 1;
  const int64_t fit = m * n > 0 ? capacity / (m * n) : 1;
  const int dma_splits =
      sizeof(T) == 8 ? int(std::min<int64_t>(DmaSplitKCount(m, n, k), fit)) : 1;
  *dma = dma_splits > 1;
  return *dma ? dma_splits : int(std::min<int64_t>(SplitKCount(m, n, k), fit));
}

template <typename T>
void GemmSplitK(bool ta, bool tb, int64_t m, int64_t n, int64_t k, T alpha, const T* a,
                int64_t lda, const T* b, int64_t ldb, T beta, T* c, int64_t ldc, T* partials,
                int64_t capacity, hipStream_t stream, bool entry = false) {
  bool dma = false;
  const int splits = GemmSplitCount<T>(m, n, k, capacity, &dma, entry);
  if (splits <= 1) {
    GemmDevice<T>(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream);
    return;
  }
  // Chunks of whole 32-deep slices (16-deep for the DMA kernel); each partial
  // is written with alpha 1, beta 0.
  const int64_t step = dma ? 16 : 32;
  const int64_t chunk = ((k + splits - 1) / splits + step - 1) / step * step;
  const int used = int((k + chunk - 1) / chunk);
  const int cfg = dma ? 15 : ChooseRegularGemmConfig(int(sizeof(T)), m, n, chunk);
  if constexpr (sizeof(T) == 8)
    GemmLaunchF64(cfg, ta, tb, m, n, k, 1.0, a, lda, b, ldb, 0.0, partials, m, stream, 0, 0,
                  nullptr, chunk, m * n);
  else
    GemmLaunchF32(cfg, ta, tb, m, n, k, 1.0f, a, lda, b, ldb, 0.0f, partials, m, stream, 0, 0,
                  chunk, m * n);
  const int64_t total = m * n;
  hipLaunchKernelGGL(SplitSumKernel<T>, dim3(unsigned(std::min<int64_t>((total + 255) / 256, 4096))),
                     dim3(256), 0, stream, partials, used, m, n, alpha, beta, c, ldc);
  HWY_HIP_CHECK(hipGetLastError());
}

}  // namespace

void GemmSplitKDevice(bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha,
                      const double* a, int64_t lda, const double* b, int64_t ldb, double beta,
                      double* c, int64_t ldc, double* partials, int64_t capacity,
                      hipStream_t stream) {
  GemmSplitK<double>(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, partials, capacity,
                     stream);
}

template <typename T>
bool GemmSplitPays(int64_t m, int64_t n, int64_t k) {
  bool dma = false;
  return GemmSplitCount<T>(m, n, k, int64_t(kMaxSplitK) * m * n, &dma, true) > 1;
}
template bool GemmSplitPays<double>(int64_t, int64_t, int64_t);
template bool GemmSplitPays<float>(int64_t, int64_t, int64_t);

template <typename T>
void GemmSplitKPublic(bool ta, bool tb, int64_t m, int64_t n, int64_t k, T alpha, const T* a,
                      int64_t lda, const T* b, int64_t ldb, T beta, T* c, int64_t ldc,
                      hipStream_t stream) {
  bool dma = false;
  const int splits = GemmSplitCount<T>(m, n, k, int64_t(kMaxSplitK) * m * n, &dma, true);
  const size_t bytes = size_t(splits) * size_t(m) * size_t(n) * sizeof(T);
  T* partials = static_cast<T*>(SplitPartials(bytes, stream));
  GemmSplitK<T>(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, partials,
                int64_t(splits) * m * n, stream, true);
  SplitPartialsDone(stream);
}
template void GemmSplitKPublic<double>(bool, bool, int64_t, int64_t, int64_t, double,
                                       const double*, int64_t, const double*, int64_t, double,
                                       double*, int64_t, hipStream_t);
template void GemmSplitKPublic<float>(bool, bool, int64_t, int64_t, int64_t, float, const float*,
                                      int64_t, const float*, int64_t, float, float*, int64_t,
                                      hipStream_t);

namespace {

// A triangle of C runs on square tiles, so that only the tiles meeting it
// are launched (the 128 x 128 x 16 two-waves-per-SIMD tile otherwise).
int PartConfig(int cfg, int part) {
  if (part == 0 || kGemmConfigs[cfg].bm == kGemmConfigs[cfg].bn) return cfg;
  return 8;
}

}  // namespace

template <>
void GemmDevicePart<double>(int part, bool ta, bool tb, int64_t m, int64_t n, int64_t k,
                            double alpha, const double* a, int64_t lda, const double* b,
                            int64_t ldb, double beta, double* c, int64_t ldc,
                            hipStream_t stream) {
  const int cfg = PartConfig(ChooseGemmConfig(8, ta, tb, m, n, k), part);
  GemmLaunchF64(cfg, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream, part);
}

template <>
void GemmDevicePart<float>(int part, bool ta, bool tb, int64_t m, int64_t n, int64_t k,
                           float alpha, const float* a, int64_t lda, const float* b,
                           int64_t ldb, float beta, float* c, int64_t ldc,
                           hipStream_t stream) {
  const int cfg = PartConfig(ChooseGemmConfig(4, ta, tb, m, n, k), part);
  GemmLaunchF32(cfg, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream, part);
}

template <>
void GemmDeviceKRange<double>(int krange, bool ta, bool tb, int64_t m, int64_t n, int64_t k,
                              double alpha, const double* a, int64_t lda, const double* b,
                              int64_t ldb, double beta, double* c, int64_t ldc,
                              hipStream_t stream) {
  const int cfg = ChooseGemmConfig(8, ta, tb, m, n, k);
  GemmLaunchF64(cfg, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream, 0, krange);
}

template <>
void GemmDeviceKRange<float>(int krange, bool ta, bool tb, int64_t m, int64_t n, int64_t k,
                             float alpha, const float* a, int64_t lda, const float* b,
                             int64_t ldb, float beta, float* c, int64_t ldc,
                             hipStream_t stream) {
  const int cfg = ChooseGemmConfig(4, ta, tb, m, n, k);
  GemmLaunchF32(cfg, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream, 0, krange);
}

namespace {

template <typename T>
__global__ void ScaleKernel(int64_t m, int64_t n, T beta, T* c, int64_t ldc) {
  const int64_t total = m * n;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
       idx += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = idx % m, j = idx / m;
    T* x = c + i + j * ldc;
    *x = (beta == T(0)) ? T(0) : *x * beta;
  }
}

}  // namespace

template <typename T>
void ScaleDevice(int64_t m, int64_t n, T beta, T* c, int64_t ldc, hipStream_t stream) {
  if (m <= 0 || n <= 0) return;
  const int64_t total = m * n;
  const int64_t blocks = std::min<int64_t>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(ScaleKernel<T>, dim3(unsigned(blocks)), dim3(256), 0, stream, m, n, beta,
                     c, ldc);
  HWY_HIP_CHECK(hipGetLastError());
}

template void ScaleDevice<double>(int64_t, int64_t, double, double*, int64_t, hipStream_t);
template void ScaleDevice<float>(int64_t, int64_t, float, float*, int64_t, hipStream_t);

}  // namespace hwy_amd

using namespace hwy_amd;

extern "C" {

HWYBLAS_AMD_API int hwyblas_gemm_config_count(void) { return kNumGemmConfigs; }

HWYBLAS_AMD_API int hwyblas_gemm_config_info(int cfg, int* bm, int* bn, int* bk, int* waves) {
  if (cfg < 0 || cfg >= kNumGemmConfigs) return -1;
  const GemmConfigInfo& g = kGemmConfigs[cfg];
  *bm = g.bm;
  *bn = g.bn;
  *bk = g.bk;
  *waves = g.waves_m * g.waves_n;
  return 0;
}

HWYBLAS_AMD_API int hwyblas_gemm_config_stages(int cfg) {
  if (cfg < 0 || cfg >= kNumGemmConfigs) return -1;
  return kGemmConfigs[cfg].stages;
}

HWYBLAS_AMD_API int hwyblas_gemm_default_config(int dtype_bytes, char transa, char transb,
                                                int m, int n, int k) {
  Op oa, ob;
  DecodeOp(transa, &oa);
  DecodeOp(transb, &ob);
  return ChooseGemmConfig(dtype_bytes, oa == Op::kTrans, ob == Op::kTrans, m, n, k);
}

namespace {
// Shape and leading-dimension checks of the config-indexed entries (no
// error reporting: they return -1).
bool CfgArgsOk(int cfg, char transa, char transb, int m, int n, int k, int lda, int ldb,
               int ldc, Op* oa, Op* ob) {
  if (cfg < 0 || cfg >= kNumGemmConfigs) return false;
  if (!DecodeOp(transa, oa) || !DecodeOp(transb, ob)) return false;
  if (m <= 0 || n <= 0 || k <= 0) return false;
  const int ra = *oa == Op::kNoTrans ? m : k, rb = *ob == Op::kNoTrans ? k : n;
  return lda >= ra && ldb >= rb && ldc >= m;
}
}  // namespace

HWYBLAS_AMD_API int hwyblas_dgemm_dev_cfg(int cfg, char transa, char transb, int m, int n,
                                          int k, double alpha, const double* a, int lda,
                                          const double* b, int ldb, double beta, double* c,
                                          int ldc, void* stream) {
  Op oa, ob;
  if (!CfgArgsOk(cfg, transa, transb, m, n, k, lda, ldb, ldc, &oa, &ob)) return -1;
  GemmLaunchF64(cfg, oa == Op::kTrans, ob == Op::kTrans, m, n, k, alpha, a, lda, b, ldb, beta,
                c, ldc, static_cast<hipStream_t>(stream));
  return 0;
}

HWYBLAS_AMD_API int hwyblas_sgemm_dev_cfg(int cfg, char transa, char transb, int m, int n,
                                          int k, float alpha, const float* a, int lda,
                                          const float* b, int ldb, float beta, float* c,
                                          int ldc, void* stream) {
  Op oa, ob;
  if (!CfgArgsOk(cfg, transa, transb, m, n, k, lda, ldb, ldc, &oa, &ob)) return -1;
  GemmLaunchF32(cfg, oa == Op::kTrans, ob == Op::kTrans, m, n, k, alpha, a, lda, b, ldb, beta,
                c, ldc, static_cast<hipStream_t>(stream));
  return 0;
}

}  // extern "C"
