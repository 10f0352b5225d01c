// Fortran LAPACK entries of the LU solver: dgetrf_, dgetrs_, dgesv_ (hwyBLAS
// src/lapack/getrf.cc:112-339 for the routines, :341-367 for xGESV's checks,
// :500-540 for the entries) and dgerfs_ (src/lapack/gesvx.cc:333-521), over
// the device path (hwyblas_dgetrf_dev, hwyblas_dgetrs_dev,
// hwyblas_dgerfs_dev; FP64 trailing updates). Host or device operands: host
// matrices and pivots are staged in the calling thread's device buffers and
// copied back. Each call synchronizes.
#include <algorithm>

#include "common.h"

using hwy_amd::blasint;

namespace hwy_amd {
namespace {

void CopyIn2D(void* dst, const void* src, int64_t rows_bytes, int64_t cols, int64_t ld_bytes,
              hipStream_t s) {
  HWY_HIP_CHECK(hwy_amd::Memcpy2DAsync(dst, rows_bytes, src, ld_bytes, rows_bytes, cols,
                                       hipMemcpyHostToDevice, s));
}

void CopyOut2D(void* dst, int64_t ld_bytes, const void* src, int64_t rows_bytes, int64_t cols,
               hipStream_t s) {
  HWY_HIP_CHECK(hwy_amd::Memcpy2DAsync(dst, ld_bytes, src, rows_bytes, rows_bytes, cols,
                                       hipMemcpyDeviceToHost, s));
}

// A host-or-device matrix of doubles (rows x cols, ld); a host one is staged
// (copied in when `in`) and copied back by Finish when `out`.
struct Staged {
  double* host = nullptr;
  double* dev = nullptr;
  int64_t rows = 0, cols = 0, ld = 0, host_ld = 0;
  Staged(double* x, int64_t r, int64_t c, int64_t ldx, int slot, bool in, hipStream_t s)
      : rows(r), cols(c) {
    if (IsDevicePointer(x)) {
      dev = x;
      ld = ldx;
      return;
    }
    host = x;
    host_ld = ldx;
    ld = std::max<int64_t>(1, r);
    dev = static_cast<double*>(Staging(slot, size_t(ld) * std::max<int64_t>(1, c) * 8));
    if (in && r > 0 && c > 0) CopyIn2D(dev, x, r * 8, c, ldx * 8, s);
  }
  void Finish(hipStream_t s) {
    if (host && rows > 0 && cols > 0) CopyOut2D(host, host_ld * 8, dev, rows * 8, cols, s);
  }
};

// Pivots: device storage for k entries (staged when host memory).
struct StagedPivots {
  int* host = nullptr;
  int* dev = nullptr;
  int64_t k = 0;
  StagedPivots(int* p, int64_t k_, bool in, hipStream_t s) : k(k_) {
    if (IsDevicePointer(p)) {
      dev = p;
      return;
    }
    host = p;
    dev = static_cast<int*>(Staging(kSlotD, size_t(std::max<int64_t>(1, k)) * sizeof(int)));
    if (in && k > 0)
      HWY_HIP_CHECK(hipMemcpyAsync(dev, p, size_t(k) * sizeof(int), hipMemcpyHostToDevice, s));
  }
  void Finish(hipStream_t s) {
    if (host && k > 0)
      HWY_HIP_CHECK(hipMemcpyAsync(host, dev, size_t(k) * sizeof(int), hipMemcpyDeviceToHost, s));
  }
};

blasint Getrf(blasint m, blasint n, double* a, blasint lda, blasint* ipiv) {
  if (m < 0) return -1;
  if (n < 0) return -2;
  if (lda < std::max(1, m)) return -4;
  if (m == 0 || n == 0) return 0;
  hipStream_t s = Context().stream;
  Staged A(a, m, n, lda, kSlotA, true, s);
  StagedPivots P(ipiv, std::min(m, n), false, s);
  // Trailing updates on the emulated DGEMM when the DGEMM mode asks for it.
  const int info = hwyblas_dgetrf_dev(m, n, A.dev, int(A.ld), P.dev, DgemmMode() == 1 ? 1 : 0, s);
  A.Finish(s);
  P.Finish(s);
  HWY_HIP_CHECK(hipStreamSynchronize(s));
  return info;
}

blasint Getrs(char trans, blasint n, blasint nrhs, const double* a, blasint lda,
              const blasint* ipiv, double* b, blasint ldb) {
  Op op;
  if (!DecodeOp(trans, &op)) return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < std::max(1, n)) return -5;
  if (ldb < std::max(1, n)) return -8;
  if (n == 0 || nrhs == 0) return 0;
  hipStream_t s = Context().stream;
  Staged A(const_cast<double*>(a), n, n, lda, kSlotA, true, s);
  StagedPivots P(const_cast<blasint*>(ipiv), n, true, s);
  Staged B(b, n, nrhs, ldb, kSlotB, true, s);
  const int info = hwyblas_dgetrs_dev(trans, n, nrhs, A.dev, int(A.ld), P.dev, B.dev, int(B.ld), s);
  B.Finish(s);
  HWY_HIP_CHECK(hipStreamSynchronize(s));
  return info;
}

blasint Gerfs(char trans, blasint n, blasint nrhs, const double* a, blasint lda,
              const double* af, blasint ldaf, const blasint* ipiv, const double* b, blasint ldb,
              double* x, blasint ldx, double* ferr, double* berr) {
  Op op;
  if (!DecodeOp(trans, &op)) return -1;
  if (n < 0) return -2;
  if (nrhs < 0) return -3;
  if (lda < std::max(1, n)) return -5;
  if (ldaf < std::max(1, n)) return -7;
  if (ldb < std::max(1, n)) return -10;
  if (ldx < std::max(1, n)) return -12;
  if (n == 0 || nrhs == 0) {
    for (blasint j = 0; j < nrhs; ++j) ferr[j] = berr[j] = 0.0;
    return 0;
  }
  hipStream_t s = Context().stream;
  Staged A(const_cast<double*>(a), n, n, lda, kSlotA, true, s);
  Staged AF(const_cast<double*>(af), n, n, ldaf, kSlotPackA, true, s);
  StagedPivots P(const_cast<blasint*>(ipiv), n, true, s);
  Staged B(const_cast<double*>(b), n, nrhs, ldb, kSlotB, true, s);
  Staged X(x, n, nrhs, ldx, kSlotC, true, s);
  const int info = hwyblas_dgerfs_dev(trans, n, nrhs, A.dev, int(A.ld), AF.dev, int(AF.ld),
                                      P.dev, B.dev, int(B.ld), X.dev, int(X.ld), ferr, berr, s);
  X.Finish(s);
  HWY_HIP_CHECK(hipStreamSynchronize(s));
  return info;
}

}  // namespace
}  // namespace hwy_amd

extern "C" {

HWYBLAS_AMD_API void dgetrf_(const hwyblas_int* m, const hwyblas_int* n, double* a,
                             const hwyblas_int* lda, hwyblas_int* ipiv, hwyblas_int* info) {
  *info = hwy_amd::Getrf(*m, *n, a, *lda, ipiv);
  if (*info == -1000) hwy_amd::ReportInternal("DGETRF", -1000);
  else if (*info < 0) hwy_amd::Reject("DGETRF", -*info);
}

HWYBLAS_AMD_API void dgetrs_(const char* trans, const hwyblas_int* n, const hwyblas_int* nrhs,
                             const double* a, const hwyblas_int* lda, const hwyblas_int* ipiv,
                             double* b, const hwyblas_int* ldb, hwyblas_int* info) {
  *info = hwy_amd::Getrs(*trans, *n, *nrhs, a, *lda, ipiv, b, *ldb);
  if (*info < 0) hwy_amd::Reject("DGETRS", -*info);
}

HWYBLAS_AMD_API void dgerfs_(const char* trans, const hwyblas_int* n, const hwyblas_int* nrhs,
                             const double* a, const hwyblas_int* lda, const double* af,
                             const hwyblas_int* ldaf, const hwyblas_int* ipiv, const double* b,
                             const hwyblas_int* ldb, double* x, const hwyblas_int* ldx,
                             double* ferr, double* berr, double* work, hwyblas_int* iwork,
                             hwyblas_int* info) {
  (void)work;  // the GPU path keeps its work arrays on the device
  (void)iwork;
  *info = hwy_amd::Gerfs(*trans, *n, *nrhs, a, *lda, af, *ldaf, ipiv, b, *ldb, x, *ldx, ferr,
                         berr);
  if (*info < 0) hwy_amd::Reject("DGERFS", -*info);
}

HWYBLAS_AMD_API void dgesv_(const hwyblas_int* n, const hwyblas_int* nrhs, double* a,
                            const hwyblas_int* lda, hwyblas_int* ipiv, double* b,
                            const hwyblas_int* ldb, hwyblas_int* info) {
  const blasint N = *n, R = *nrhs;
  blasint e = 0;
  if (N < 0) e = -1;
  else if (R < 0) e = -2;
  else if (*lda < std::max(1, N)) e = -4;
  else if (*ldb < std::max(1, N)) e = -7;
  if (e != 0) {
    *info = e;
    hwy_amd::Reject("DGESV", -e);
    return;
  }
  *info = hwy_amd::Getrf(N, N, a, *lda, ipiv);
  if (*info == -1000) hwy_amd::ReportInternal("DGESV", -1000);
  if (*info == 0) *info = hwy_amd::Getrs('N', N, R, a, *lda, ipiv, b, *ldb);
}

}  // extern "C"
