// Grouped (batched) GEMM/TRSM plans: tile configurations and entry points
// shared by grouped.hip and the autotuner / cost model.
#ifndef HWYBLAS_AMD_CSRC_GROUPED_H_
#define HWYBLAS_AMD_CSRC_GROUPED_H_

#include "common.h"

// Wave tiles, X(id, BM, BN, UK): one wave computes a BM x BN tile, loading
// its MFMA operands straight from memory, k advancing 4 * UK per step.
#define HWY_GROUPED_WAVE_CONFIGS(X) \
  X(0, 16, 16, 1)                   \
  X(1, 16, 16, 4)                   \
  X(2, 32, 32, 1)                   \
  X(3, 32, 32, 4)                   \
  X(4, 64, 32, 1)                   \
  X(5, 64, 32, 4)                   \
  X(6, 32, 64, 1)                   \
  X(7, 32, 64, 4)                   \
  X(8, 64, 64, 1)                   \
  X(9, 64, 64, 2)

// Workgroup tiles, X(id, BM, BN, BK): the four waves of a workgroup compute
// a BM x BN tile (2 x 2 waves) with the plain kernel's LDS-staged macro-tile,
// BK of k per stage. (A 128 x 128 tile would not fit the persistent kernel's
// registers at two waves per SIMD beside the wave tiles.)
#define HWY_GROUPED_BLOCK_CONFIGS(X) \
  X(10, 64, 64, 16)                  \
  X(11, 128, 64, 8)                  \
  X(12, 64, 128, 8)                  \
  X(13, 64, 64, 32)

namespace hwy_amd {
namespace grouped {

struct GemmConfig {
  int bm, bn;
  int uk;     // wave tiles: k steps of 4 per load batch; workgroup tiles: BK
  int block;  // 1: a workgroup tile
};

constexpr GemmConfig kGroupedGemmConfigs[] = {
#define HWY_GROUPED_WAVE_INFO(ID, BM, BN, UK) {BM, BN, UK, 0},
    HWY_GROUPED_WAVE_CONFIGS(HWY_GROUPED_WAVE_INFO)
#undef HWY_GROUPED_WAVE_INFO
#define HWY_GROUPED_BLOCK_INFO(ID, BM, BN, BK) {BM, BN, BK, 1},
    HWY_GROUPED_BLOCK_CONFIGS(HWY_GROUPED_BLOCK_INFO)
#undef HWY_GROUPED_BLOCK_INFO
};
constexpr int kNumGroupedGemmConfigs = sizeof(kGroupedGemmConfigs) / sizeof(kGroupedGemmConfigs[0]);

struct Plan;
void Enqueue(const Plan* plan, hipStream_t stream);
void LaunchGraph(Plan* plan, hipStream_t stream);
void Destroy(Plan* plan);
int DefaultGroupedConfig(int dtype_bytes, int m, int n, int k);

}  // namespace grouped
}  // namespace hwy_amd

#endif  // HWYBLAS_AMD_CSRC_GROUPED_H_
