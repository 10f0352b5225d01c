// FP32 instantiations of the plain MFMA GEMM (v_mfma_f32_16x16x4_f32).
#include "gemm_kernel.h"

namespace hwy_amd {

void GemmLaunchF32(int cfg, bool ta, bool tb, int64_t m, int64_t n, int64_t k, float alpha,
                   const float* a, int64_t lda, const float* b, int64_t ldb, float beta,
                   float* c, int64_t ldc, hipStream_t stream, int part, int krange,
                   int64_t ksplit, int64_t c_split) {
  gemm_detail::GemmParams<float> prm{a, b, c, m, n, k, lda, ldb, ldc, alpha, beta, 0, 0, part};
  prm.krange = krange;
  prm.ksplit = ksplit;
  prm.c_split = c_split;
  gemm_detail::LaunchAny<float>(cfg, ta, tb, prm, stream);
}

}  // namespace hwy_amd
