// Cholesky factorization and solve on the GPU: ?potrf_, ?potrs_, ?posv_
// (hwyBLAS src/lapack/potrf.cc, posv.cc: argument checks, info positions and
// quick returns). The factorization is the recursion of xPOTRF2
// (src/lapack/potrf.cc:49-122), right-looking: factor A11, solve the block
// beside it with the library's TRSM, update A22 with the triangle-only GEMM
// (SYRK), factor A22. Diagonal blocks of at most kCholLeaf are factored by one
// wave in LDS, column by column (l_jj = sqrt(a_jj), l_ij = a_ij / l_jj, then
// the trailing update of the block). hwyBLAS's blocked xPOTRF runs the same
// arithmetic in a left-looking order, so results agree up to rounding order.
// info > 0: the leading minor of that order is not positive (or is NaN); the
// leaves after it do nothing, and the entries past that column are not
// specified.
#include <algorithm>
#include <climits>

#include "common.h"

namespace hwy_amd {
namespace {

constexpr int kCholLeaf = 64;

// In-place Cholesky of the n x n (n <= kCholLeaf) diagonal block at a: the
// lower triangle (A = L L^T) or the upper one (A = U^T U, handled as the
// lower factor of its transpose). A non-positive or NaN pivot j records
// offset + j + 1 in *info (atomic min) and stops; a block whose *info is
// already set does nothing.
template <typename T>
__global__ void __launch_bounds__(kCholLeaf)
    CholLeafKernel(T* __restrict__ a, int64_t lda, int n, bool upper, int offset,
                   int* __restrict__ info) {
  __shared__ T t[kCholLeaf][kCholLeaf + 1];
  __shared__ int failed;
  const int i = threadIdx.x;
  if (i == 0) failed = (*info != INT_MAX);
  __syncthreads();
  if (failed) return;
  // t[r][c], r >= c: the block's lower factor in the making.
  for (int c = 0; c < n; ++c)
    if (i < n) t[i][c] = (i >= c) ? (upper ? a[c + int64_t(i) * lda] : a[i + int64_t(c) * lda]) : T(0);
  __syncthreads();
  int fail_at = -1;
  for (int j = 0; j < n; ++j) {
    const T d = t[j][j];
    if (!(d > T(0))) {
      fail_at = j;
      break;
    }
    const T ljj = sqrt(d);
    if (i > j && i < n) t[i][j] = t[i][j] / ljj;
    __syncthreads();
    if (i == j) t[j][j] = ljj;
    if (i > j && i < n) {
      const T lij = t[i][j];
      for (int k = j + 1; k <= i; ++k) t[i][k] = __builtin_fma(-lij, t[k][j], t[i][k]);
    }
    __syncthreads();
  }
  if (fail_at >= 0 && i == 0) atomicMin(info, offset + fail_at + 1);
  // Write back every column before the failure (all of them on success).
  const int done = fail_at >= 0 ? fail_at : n;
  for (int c = 0; c < done; ++c) {
    if (i < n && i >= c) {
      if (upper) a[c + int64_t(i) * lda] = t[i][c];
      else a[i + int64_t(c) * lda] = t[i][c];
    }
  }
}

template <typename T>
void CholRecurse(bool upper, int n, T* a, int64_t lda, int offset, int* info, hipStream_t s) {
  if (n <= kCholLeaf) {
    hipLaunchKernelGGL(CholLeafKernel<T>, dim3(1), dim3(kCholLeaf), 0, s, a, lda, n, upper, offset,
                       info);
    HWY_HIP_CHECK(hipGetLastError());
    return;
  }
  const int n1 = std::max(kCholLeaf, (n / 2) / kCholLeaf * kCholLeaf);
  const int n2 = n - n1;
  CholRecurse<T>(upper, n1, a, lda, offset, info, s);
  T* a22 = a + n1 + int64_t(n1) * lda;
  if (upper) {
    // A12 := U11^-T A12; A22 -= A12^T A12 (upper triangle).
    T* a12 = a + int64_t(n1) * lda;
    TrsmDevice<T>(true, false, true, false, n1, n2, a, lda, a12, lda, s);
    GemmDevicePart<T>(2, true, false, n2, n2, n1, T(-1), a12, lda, a12, lda, T(1), a22, lda, s);
  } else {
    // A21 := A21 L11^-T; A22 -= A21 A21^T (lower triangle).
    T* a21 = a + n1;
    TrsmDevice<T>(false, true, true, false, n2, n1, a, lda, a21, lda, s);
    GemmDevicePart<T>(1, false, true, n2, n2, n1, T(-1), a21, lda, a21, lda, T(1), a22, lda, s);
  }
  CholRecurse<T>(upper, n2, a22, lda, offset + n1, info, s);
}

// Factors the device matrix in place; returns info (0 or the first failing
// minor). Synchronizes.
template <typename T>
int PotrfDevice(bool upper, int n, T* a, int64_t lda, hipStream_t s) {
  int* dinfo = static_cast<int*>(Staging(kSlotD, 256));
  const int init = INT_MAX;
  HWY_HIP_CHECK(hipMemcpyAsync(dinfo, &init, sizeof(int), hipMemcpyHostToDevice, s));
  CholRecurse<T>(upper, n, a, lda, 0, dinfo, s);
  int info = INT_MAX;
  HWY_HIP_CHECK(hipMemcpyAsync(&info, dinfo, sizeof(int), hipMemcpyDeviceToHost, s));
  HWY_HIP_CHECK(hipStreamSynchronize(s));
  return info == INT_MAX ? 0 : info;
}

// B := A^-1 B with the factor (lower: L^-T L^-1 B; upper: U^-1 U^-T B).
template <typename T>
void PotrsDevice(bool upper, int n, int nrhs, const T* a, int64_t lda, T* b, int64_t ldb,
                 hipStream_t s) {
  if (upper) {
    TrsmDevice<T>(true, false, true, false, n, nrhs, a, lda, b, ldb, s);
    TrsmDevice<T>(true, false, false, false, n, nrhs, a, lda, b, ldb, s);
  } else {
    TrsmDevice<T>(true, true, false, false, n, nrhs, a, lda, b, ldb, s);
    TrsmDevice<T>(true, true, true, false, n, nrhs, a, lda, b, ldb, s);
  }
}

// A matrix operand on the device: the caller's, or a staged copy (written
// back by Finish when `out`).
template <typename T>
struct DevMatrix {
  T* host = nullptr;
  int64_t host_ld = 0, rows = 0, cols = 0;
  T* dev = nullptr;
  int64_t ld = 0;
  DevMatrix(T* x, int64_t r, int64_t c, int64_t ldx, int slot, hipStream_t s)
      : rows(r), cols(c) {
    if (IsDevicePointer(x)) {
      dev = x;
      ld = ldx;
      return;
    }
    host = x;
    host_ld = ldx;
    dev = static_cast<T*>(Staging(slot, size_t(std::max<int64_t>(1, r * c)) * sizeof(T)));
    ld = r;
    HWY_HIP_CHECK(hipMemcpy2DAsync(dev, size_t(r) * sizeof(T), x, size_t(ldx) * sizeof(T),
                                   size_t(r) * sizeof(T), c, hipMemcpyHostToDevice, s));
  }
  void CopyBack(hipStream_t s) {
    if (host)
      HWY_HIP_CHECK(hipMemcpy2DAsync(host, size_t(host_ld) * sizeof(T), dev, size_t(ld) * sizeof(T),
                                     size_t(rows) * sizeof(T), cols, hipMemcpyDeviceToHost, s));
  }
};

// src/lapack/potrf.cc:129-153 (info: 1 uplo, 2 n, 4 lda).
blasint PotrfInfo(char uplo, blasint n, blasint lda, bool* upper) {
  bool lower;
  if (!DecodeUplo(uplo, &lower)) return 1;
  *upper = !lower;
  if (n < 0) return 2;
  if (lda < std::max(1, n)) return 4;
  return 0;
}

// xPOTRS / xPOSV (info: 1 uplo, 2 n, 3 nrhs, 5 lda, 7 ldb).
blasint PosvInfo(char uplo, blasint n, blasint nrhs, blasint lda, blasint ldb, bool* upper) {
  bool lower;
  if (!DecodeUplo(uplo, &lower)) return 1;
  *upper = !lower;
  if (n < 0) return 2;
  if (nrhs < 0) return 3;
  if (lda < std::max(1, n)) return 5;
  if (ldb < std::max(1, n)) return 7;
  return 0;
}

template <typename T>
void PotrfHost(const char* name, const char* uplo, const blasint* n, T* a, const blasint* lda,
               blasint* info) {
  bool upper = false;
  *info = -PotrfInfo(*uplo, *n, *lda, &upper);
  if (*info != 0) return ReportError(name, -*info);
  if (*n == 0) return;
  hipStream_t s = Context().stream;
  DevMatrix<T> ma(a, *n, *n, *lda, kSlotA, s);
  *info = PotrfDevice<T>(upper, *n, ma.dev, ma.ld, s);
  ma.CopyBack(s);
  HWY_HIP_CHECK(hipStreamSynchronize(s));
}

template <typename T>
void PotrsHost(const char* name, const char* uplo, const blasint* n, const blasint* nrhs,
               const T* a, const blasint* lda, T* b, const blasint* ldb, blasint* info) {
  bool upper = false;
  *info = -PosvInfo(*uplo, *n, *nrhs, *lda, *ldb, &upper);
  if (*info != 0) return ReportError(name, -*info);
  if (*n == 0 || *nrhs == 0) return;
  hipStream_t s = Context().stream;
  DevMatrix<T> ma(const_cast<T*>(a), *n, *n, *lda, kSlotA, s);
  DevMatrix<T> mb(b, *n, *nrhs, *ldb, kSlotB, s);
  PotrsDevice<T>(upper, *n, *nrhs, ma.dev, ma.ld, mb.dev, mb.ld, s);
  mb.CopyBack(s);
  HWY_HIP_CHECK(hipStreamSynchronize(s));
}

// xPOSV: the factorization, then (if it succeeded) the solve.
template <typename T>
void PosvHost(const char* name, const char* uplo, const blasint* n, const blasint* nrhs, T* a,
              const blasint* lda, T* b, const blasint* ldb, blasint* info) {
  bool upper = false;
  *info = -PosvInfo(*uplo, *n, *nrhs, *lda, *ldb, &upper);
  if (*info != 0) return ReportError(name, -*info);
  if (*n == 0) return;
  hipStream_t s = Context().stream;
  DevMatrix<T> ma(a, *n, *n, *lda, kSlotA, s);
  *info = PotrfDevice<T>(upper, *n, ma.dev, ma.ld, s);
  ma.CopyBack(s);
  if (*info == 0 && *nrhs > 0) {
    DevMatrix<T> mb(b, *n, *nrhs, *ldb, kSlotB, s);
    PotrsDevice<T>(upper, *n, *nrhs, ma.dev, ma.ld, mb.dev, mb.ld, s);
    mb.CopyBack(s);
  }
  HWY_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace
}  // namespace hwy_amd

using hwy_amd::blasint;

extern "C" {

#define HWY_CHOL_ENTRIES(P, T, NAME_P)                                                        \
  HWYBLAS_AMD_API void P##potrf_(const char* uplo, const blasint* n, T* a, const blasint* lda, \
                                 blasint* info) {                                             \
    hwy_amd::PotrfHost<T>(NAME_P "POTRF", uplo, n, a, lda, info);                            \
  }                                                                                           \
  HWYBLAS_AMD_API void P##potrs_(const char* uplo, const blasint* n, const blasint* nrhs,    \
                                 const T* a, const blasint* lda, T* b, const blasint* ldb,     \
                                 blasint* info) {                                             \
    hwy_amd::PotrsHost<T>(NAME_P "POTRS", uplo, n, nrhs, a, lda, b, ldb, info);              \
  }                                                                                           \
  HWYBLAS_AMD_API void P##posv_(const char* uplo, const blasint* n, const blasint* nrhs, T* a, \
                                const blasint* lda, T* b, const blasint* ldb, blasint* info) { \
    hwy_amd::PosvHost<T>(NAME_P "POSV", uplo, n, nrhs, a, lda, b, ldb, info);                \
  }
HWY_CHOL_ENTRIES(s, float, "S")
HWY_CHOL_ENTRIES(d, double, "D")
#undef HWY_CHOL_ENTRIES

}  // extern "C"
