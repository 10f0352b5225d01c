// Grouped (batched) GEMM/TRSM plans: tile configurations and entry points
// shared by grouped.hip and the autotuner / cost model.
#ifndef HWYBLAS_AMD_CSRC_GROUPED_H_
#define HWYBLAS_AMD_CSRC_GROUPED_H_

#include "common.h"

// X(id, BM, BN, UK): one wave computes a BM x BN tile, k advancing 4 * UK
// per step.
#define HWY_GROUPED_GEMM_CONFIGS(X) \
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

namespace hwy_amd {
namespace grouped {

struct GemmConfig {
  int bm, bn, uk;
};

constexpr GemmConfig kGroupedGemmConfigs[] = {
#define HWY_GROUPED_INFO(ID, BM, BN, UK) {BM, BN, UK},
    HWY_GROUPED_GEMM_CONFIGS(HWY_GROUPED_INFO)
#undef HWY_GROUPED_INFO
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
