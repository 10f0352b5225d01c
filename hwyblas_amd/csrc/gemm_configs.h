// Tile configurations of the plain MFMA GEMM kernel. The autotuner times
// them per shape, the cost model predicts among them, and GemmDevice picks
// one by the rules in gemm.hip when neither has an opinion.
//
// X(id, BM, BN, BK, WAVES_M, WAVES_N, OCC): macro-tile BM x BN, k slice BK,
// WAVES_M x WAVES_N waves (each computing (BM/WAVES_M) x (BN/WAVES_N)), and
// the minimum waves per SIMD the register allocation must leave room for.
#ifndef HWYBLAS_AMD_CSRC_GEMM_CONFIGS_H_
#define HWYBLAS_AMD_CSRC_GEMM_CONFIGS_H_

#define HWY_GEMM_CONFIGS(X)      \
  X(0, 64, 64, 16, 2, 2, 1)      \
  X(1, 128, 128, 16, 2, 2, 1)    \
  X(2, 128, 64, 16, 2, 2, 1)     \
  X(3, 64, 128, 16, 2, 2, 1)     \
  X(4, 128, 128, 8, 2, 2, 1)     \
  X(5, 256, 128, 16, 4, 2, 1)    \
  X(6, 128, 256, 16, 2, 4, 1)    \
  X(7, 32, 32, 16, 1, 1, 1)      \
  X(8, 128, 128, 16, 2, 2, 2)    \
  X(9, 128, 128, 32, 2, 2, 1)    \
  X(10, 256, 128, 8, 4, 2, 1)    \
  X(11, 32, 32, 32, 1, 1, 1)     \
  X(12, 32, 32, 64, 1, 1, 1)

namespace hwy_amd {

struct GemmConfigInfo {
  int bm, bn, bk, waves_m, waves_n;
};

constexpr GemmConfigInfo kGemmConfigs[] = {
#define HWY_GEMM_INFO(ID, BM, BN, BK, WM, WN, OCC) {BM, BN, BK, WM, WN},
    HWY_GEMM_CONFIGS(HWY_GEMM_INFO)
#undef HWY_GEMM_INFO
};
constexpr int kNumGemmConfigs = sizeof(kGemmConfigs) / sizeof(kGemmConfigs[0]);

}  // namespace hwy_amd

#endif  // HWYBLAS_AMD_CSRC_GEMM_CONFIGS_H_
