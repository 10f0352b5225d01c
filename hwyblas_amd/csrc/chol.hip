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
// Columns the leaf factors together in registers between LDS exchanges.
constexpr int kCholPanel = 8;

// Lane `lane`'s v in every lane (lane wave-uniform): v_readlane per 32 bits.
__device__ inline double LaneValue(double v, int lane) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(int(bits & 0xFFFFFFFFll), lane);
  const int hi = __builtin_amdgcn_readlane(int(bits >> 32), lane);
  return __longlong_as_double((static_cast<long long>(hi) << 32) |
                              static_cast<unsigned int>(lo));
}
__device__ inline float LaneValue(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// In-place Cholesky of the n x n (n <= kCholLeaf) diagonal block at a: the
// lower triangle (A = L L^T) or the upper one (A = U^T U, handled as the
// lower factor of its transpose). A non-positive or NaN pivot j records
// offset + j + 1 in *info (atomic min) and stops; a block whose *info is
// already set does nothing.
//
// One wave; lane i holds row i of the block in registers (every index a
// compile-time constant), padded to kCholLeaf with the identity. Columns are
// finished kCholPanel at a time, left-looking: a panel first takes the
// products of the finished columns (from LDS, kept column-major so a row's
// panel entries are one broadcast read), then factors its own columns with
// the diagonal and the column entries passed between lanes by v_readlane.
// Each entry sees the same sequence of fused multiply-adds as the
// right-looking column order: a_ij -= l_ik l_jk for k = 0, 1, ..., then
// l_ij = a_ij / sqrt(a_jj).
template <typename T, bool UPPER>
__global__ void __launch_bounds__(kCholLeaf)
    CholLeafKernel(T* __restrict__ a, int64_t lda, int n, int offset, int* __restrict__ info) {
  __shared__ __align__(16) T lcol[kCholLeaf][kCholLeaf];  // lcol[c][r] = l_rc
  if (__builtin_amdgcn_readfirstlane(*info) != INT_MAX) return;
  const int i = threadIdx.x;
  const int ri = min(i, n - 1);
  T r[kCholLeaf];
#pragma unroll
  for (int c = 0; c < kCholLeaf; ++c) {
    const int rc = min(c, n - 1);
    r[c] = UPPER ? a[rc + int64_t(ri) * lda] : a[ri + int64_t(rc) * lda];
  }
#pragma unroll
  for (int c = 0; c < kCholLeaf; ++c)
    if (i >= n || c >= n) r[c] = (i == c) ? T(1) : T(0);
  int done = kCholLeaf;
#pragma unroll
  for (int p = 0; p < kCholLeaf; p += kCholPanel) {
#pragma unroll 4
    for (int k = 0; k < p; ++k) {
      const T lik = lcol[k][i];
#pragma unroll
      for (int c = 0; c < kCholPanel; ++c)
        r[p + c] = __builtin_fma(-lik, lcol[k][p + c], r[p + c]);
    }
#pragma unroll
    for (int c = 0; c < kCholPanel; ++c) {
      const int j = p + c;
      const T d = LaneValue(r[j], j);
      if (!(d > T(0))) {
        done = j;
        break;
      }
      const T ljj = sqrt(d);
      const T scaled = r[j] / ljj;
      r[j] = i > j ? scaled : (i == j ? ljj : r[j]);
#pragma unroll
      for (int c2 = c + 1; c2 < kCholPanel; ++c2)
        r[p + c2] = __builtin_fma(-r[j], LaneValue(r[j], p + c2), r[p + c2]);
      lcol[j][i] = r[j];
    }
    if (done < kCholLeaf) break;
    __syncthreads();
  }
  if (done < n && i == 0) atomicMin(info, offset + done + 1);
  // Lane i writes its own row: every column before the failure (all of them
  // on success).
#pragma unroll
  for (int c = 0; c < kCholLeaf; ++c) {
    if (c < done && c < n && i >= c && i < n) {
      if (UPPER) a[c + int64_t(i) * lda] = r[c];
      else a[i + int64_t(c) * lda] = r[c];
    }
  }
}

template <typename T>
void CholRecurse(bool upper, int n, T* a, int64_t lda, int offset, int* info, hipStream_t s) {
  if (n <= kCholLeaf) {
    if (upper)
      hipLaunchKernelGGL((CholLeafKernel<T, true>), dim3(1), dim3(kCholLeaf), 0, s, a, lda, n,
                         offset, info);
    else
      hipLaunchKernelGGL((CholLeafKernel<T, false>), dim3(1), dim3(kCholLeaf), 0, s, a, lda, n,
                         offset, info);
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
    HWY_HIP_CHECK(hwy_amd::Memcpy2DAsync(dev, size_t(r) * sizeof(T), x, size_t(ldx) * sizeof(T),
                                         size_t(r) * sizeof(T), c, hipMemcpyHostToDevice, s));
  }
  void CopyBack(hipStream_t s) {
    if (host)
      HWY_HIP_CHECK(hwy_amd::Memcpy2DAsync(host, size_t(host_ld) * sizeof(T), dev,
                                           size_t(ld) * sizeof(T), size_t(rows) * sizeof(T), cols,
                                           hipMemcpyDeviceToHost, s));
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
  if (*info != 0) return Reject(name, -*info);
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
  if (*info != 0) return Reject(name, -*info);
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
  if (*info != 0) return Reject(name, -*info);
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
