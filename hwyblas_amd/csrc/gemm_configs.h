// Tile configurations of the plain MFMA GEMM kernel. The autotuner times
// them per shape, the cost model predicts among them, and GemmDevice picks
// one by the rules in gemm.hip when neither has an opinion.
//
// X(id, BM, BN, BK, WAVES_M, WAVES_N, OCC): macro-tile BM x BN, k slice BK,
// WAVES_M x WAVES_N waves (each computing (BM/WAVES_M) x (BN/WAVES_N)), and
// the minimum waves per SIMD the register allocation must leave room for.
#ifndef HWYBLAS_AMD_CSRC_GEMM_CONFIGS_H_
#define HWYBLAS_AMD_CSRC_GEMM_CONFIGS_H_

#include <stdint.h>

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
  X(12, 32, 32, 64, 1, 1, 1)     \
  X(13, 256, 128, 16, 4, 4, 4)   \
  X(14, 256, 256, 16, 4, 4, 4)

// DMA-fed configurations (GemmDmaKernel): the operand slices are streamed
// into a STAGES-deep ring of LDS buffers by global_load_lds_dwordx4, so
// STAGES - 1 slices are in flight while one is multiplied. They run the
// full tiles of a plain GEMM whose k is a multiple of BK and whose operands
// are 16-byte aligned with leading dimensions of whole 16-byte chunks; the
// remaining row and column strips (and any other call) go to a plain
// configuration, with the same k order, so the bits do not depend on it.
// A configuration whose LDS ring or accumulators do not fit for a type
// (DmaBuilt) runs the plain rule for that type.
// X(id, BM, BN, BK, WAVES_M, WAVES_N, OCC, STAGES)
#define HWY_GEMM_DMA_CONFIGS(X)            \
  X(15, 128, 128, 16, 2, 2, 2, 2)          \
  X(16, 256, 256, 16, 4, 4, 4, 2)          \
  X(17, 256, 128, 16, 4, 2, 2, 3)

namespace hwy_amd {

struct GemmConfigInfo {
  int bm, bn, bk, waves_m, waves_n, stages;  // stages: 0 for a plain configuration
};

constexpr GemmConfigInfo kGemmConfigs[] = {
#define HWY_GEMM_INFO(ID, BM, BN, BK, WM, WN, OCC) {BM, BN, BK, WM, WN, 0},
    HWY_GEMM_CONFIGS(HWY_GEMM_INFO)
#undef HWY_GEMM_INFO
#define HWY_GEMM_DMA_INFO(ID, BM, BN, BK, WM, WN, OCC, ST) {BM, BN, BK, WM, WN, ST},
    HWY_GEMM_DMA_CONFIGS(HWY_GEMM_DMA_INFO)
#undef HWY_GEMM_DMA_INFO
};
constexpr int kNumGemmConfigs = sizeof(kGemmConfigs) / sizeof(kGemmConfigs[0]);

// The plain (register-staged) configuration the built-in rule picks for a
// shape; the DMA configurations hand their strips and unsupported calls to it.
int ChooseRegularGemmConfig(int dtype_bytes, int64_t m, int64_t n, int64_t k);

}  // namespace hwy_amd

#endif  // HWYBLAS_AMD_CSRC_GEMM_CONFIGS_H_
