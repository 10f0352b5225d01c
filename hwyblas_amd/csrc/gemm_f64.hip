// FP64 instantiations of the plain MFMA GEMM (v_mfma_f64_16x16x4_f64).
#include "gemm_kernel.h"

namespace hwy_amd {

void GemmLaunchF64(int cfg, bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha,
                   const double* a, int64_t lda, const double* b, int64_t ldb, double beta,
                   double* c, int64_t ldc, hipStream_t stream, int part, int krange,
                   const int* gate, int64_t ksplit, int64_t c_split) {
  gemm_detail::GemmParams<double> prm{a, b, c, m, n, k, lda, ldb, ldc, alpha, beta, 0, 0, part};
  prm.krange = krange;
  prm.gate = gate;
  prm.ksplit = ksplit;
  prm.c_split = c_split;
  gemm_detail::LaunchAny<double>(cfg, ta, tb, prm, stream);
}

}  // namespace hwy_amd
