// Fortran LAPACK entries of the LU solver: dgetrf_, dgetrs_, dgesv_ (hwyBLAS
// src/lapack/getrf.cc:112-339 for the routines, :341-367 for xGESV's checks,
// :500-540 for the entries), over the device path (hwyblas_dgetrf_dev,
// hwyblas_dgetrs_dev; FP64 trailing updates). Host or device operands: host
// matrices and pivots are staged in the calling thread's device buffers and
// copied back. Each call synchronizes.
#include <algorithm>

#include "common.h"

using hwy_amd::blasint;

namespace hwy_amd {
namespace {

void CopyIn2D(void* dst, const void* src, int64_t rows_bytes, int64_t cols, int64_t ld_bytes,
              hipStream_t s) {
  HWY_HIP_CHECK(hipMemcpy2DAsync(dst, rows_bytes, src, ld_bytes, rows_bytes, cols,
                                 hipMemcpyHostToDevice, s));
}

void CopyOut2D(void* dst, int64_t ld_bytes, const void* src, int64_t rows_bytes, int64_t cols,
               hipStream_t s) {
  HWY_HIP_CHECK(hipMemcpy2DAsync(dst, ld_bytes, src, rows_bytes, rows_bytes, cols,
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
  const int info = hwyblas_dgetrf_dev(m, n, A.dev, int(A.ld), P.dev, 0, s);
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

}  // namespace
}  // namespace hwy_amd

extern "C" {

HWYBLAS_AMD_API void dgetrf_(const hwyblas_int* m, const hwyblas_int* n, double* a,
                             const hwyblas_int* lda, hwyblas_int* ipiv, hwyblas_int* info) {
  *info = hwy_amd::Getrf(*m, *n, a, *lda, ipiv);
  if (*info < 0 && *info != -1000) hwy_amd::Reject("DGETRF", -*info);
}

HWYBLAS_AMD_API void dgetrs_(const char* trans, const hwyblas_int* n, const hwyblas_int* nrhs,
                             const double* a, const hwyblas_int* lda, const hwyblas_int* ipiv,
                             double* b, const hwyblas_int* ldb, hwyblas_int* info) {
  *info = hwy_amd::Getrs(*trans, *n, *nrhs, a, *lda, ipiv, b, *ldb);
  if (*info < 0) hwy_amd::Reject("DGETRS", -*info);
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
  if (*info == 0) *info = hwy_amd::Getrs('N', N, R, a, *lda, ipiv, b, *ldb);
}

}  // extern "C"
