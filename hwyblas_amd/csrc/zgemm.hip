// Complex GEMM, C := alpha op(A) op(B) + beta C with op = N, T or C
// (conjugate transpose): cgemm_, zgemm_ and their device-pointer forms
// (hwyBLAS src/gemm.cc:73-132 for the checks, info positions and quick
// returns, src/gemm.cc:160-176 for the entries).
//
// Arithmetic (hwyBLAS src/gemm_kernel-inl.h:228-424): the four real sums
// sum ar br, sum ai br, sum ar bi, sum ai bi are accumulated separately and
// combined once, ab = (ar br - ai bi, ai br + ar bi), then
// r = ab alpha (complex product (x sr - y si, y sr + x si)); beta == 1:
// r + c; beta != 0: r + c beta; beta == 0 never reads C. Conjugation is
// applied while packing, so the products are plain.
//
// On MI355X the four sums are ONE real MFMA GEMM: op(A) is packed as the
// real 2m x k matrix [Re; Im] and op(B) as the real k x 2n matrix [Re Im],
// so that their product (2m x 2n) holds [ar br, ar bi; ai br, ai bi] in its
// four quadrants, at the 8mnk flops of the complex product. A combining pass
// reads the quadrants and C and writes C. Workspace: 2(m + n)k + 4mn reals.
#include <algorithm>

#include "common.h"

namespace hwy_amd {
namespace {

template <typename R>
struct Vec2;
template <>
struct Vec2<double> {
  using type = double2;
};
template <>
struct Vec2<float> {
  using type = float2;
};

enum class ZOp { kN, kT, kC };

bool DecodeZOp(char c, ZOp* op) {
  switch (c) {
    case 'N': case 'n': *op = ZOp::kN; return true;
    case 'T': case 't': *op = ZOp::kT; return true;
    case 'C': case 'c': *op = ZOp::kC; return true;
    default: *op = ZOp::kN; return false;
  }
}

// Info of the first invalid argument, or 0 (src/gemm.cc:84-106).
blasint ZgemmInfo(char transa, char transb, blasint m, blasint n, blasint k, blasint lda,
                  blasint ldb, blasint ldc, ZOp* oa, ZOp* ob) {
  const bool va = DecodeZOp(transa, oa);
  const bool vb = DecodeZOp(transb, ob);
  const blasint rows_a = *oa == ZOp::kN ? m : k;
  const blasint rows_b = *ob == ZOp::kN ? k : n;
  if (!va) return 1;
  if (!vb) return 2;
  if (m < 0) return 3;
  if (n < 0) return 4;
  if (k < 0) return 5;
  if (lda < std::max(1, rows_a)) return 8;
  if (ldb < std::max(1, rows_b)) return 10;
  if (ldc < std::max(1, m)) return 13;
  return 0;
}

constexpr int kZThreads = 256;
constexpr unsigned kZMaxSlow = 1024;  // grid rows; longer slow extents loop

dim3 ZGrid(int64_t fast, int64_t slow) {
  return dim3(unsigned((fast + kZThreads - 1) / kZThreads),
              unsigned(std::min<int64_t>(slow, kZMaxSlow)));
}

// Element (r, q) of op(X) (rows x cols) into re[r + q ld] and im[r + q ld]
// (negated for a conjugate). !TRANS: X(r, q), threads along r; TRANS:
// X(q, r), threads along q, so that the reads stay coalesced.
template <typename R, bool TRANS>
__global__ void __launch_bounds__(kZThreads)
    ZPackKernel(const typename Vec2<R>::type* __restrict__ x, int64_t ldx, int64_t rows,
                int64_t cols, bool conj, R* __restrict__ re, R* __restrict__ im, int64_t ld) {
  const int64_t f = int64_t(blockIdx.x) * kZThreads + threadIdx.x;
  const int64_t fast = TRANS ? cols : rows, slow = TRANS ? rows : cols;
  if (f >= fast) return;
  for (int64_t s = blockIdx.y; s < slow; s += gridDim.y) {
    const auto v = x[f + s * ldx];
    const int64_t r = TRANS ? s : f, q = TRANS ? f : s;
    re[r + q * ld] = v.x;
    im[r + q * ld] = conj ? -v.y : v.y;
  }
}

// x s for complex x = (x.x, x.y) and s = (sr, si).
template <typename R>
__device__ __forceinline__ void CMulScalar(R xr, R xi, R sr, R si, R* outr, R* outi) {
  *outr = xr * sr - xi * si;
  *outi = xi * sr + xr * si;
}

// C(i, j) from the four quadrants of the real product p (2m x 2n, ld 2m).
template <typename R>
__global__ void __launch_bounds__(kZThreads)
    ZCombineKernel(const R* __restrict__ p, int64_t m, int64_t n, R ar, R ai, R br, R bi,
                   int beta_kind, typename Vec2<R>::type* __restrict__ c, int64_t ldc) {
  const int64_t i = int64_t(blockIdx.x) * kZThreads + threadIdx.x;
  if (i >= m) return;
  const int64_t ldp = 2 * m;
  for (int64_t j = blockIdx.y; j < n; j += gridDim.y) {
    const R rr = p[i + j * ldp];              // sum ar br
    const R ir = p[m + i + j * ldp];          // sum ai br
    const R ri = p[i + (n + j) * ldp];        // sum ar bi
    const R ii = p[m + i + (n + j) * ldp];    // sum ai bi
    const R abr = rr - ii, abi = ir + ri;
    R xr, xi;
    CMulScalar(abr, abi, ar, ai, &xr, &xi);
    auto& cij = c[i + j * ldc];
    if (beta_kind == 1) {
      const auto cv = cij;
      xr = xr + cv.x;
      xi = xi + cv.y;
    } else if (beta_kind == 2) {
      const auto cv = cij;
      R yr, yi;
      CMulScalar(cv.x, cv.y, br, bi, &yr, &yi);
      xr = xr + yr;
      xi = xi + yi;
    }
    typename Vec2<R>::type out;
    out.x = xr;
    out.y = xi;
    cij = out;
  }
}

// C := beta C (beta == 0 writes zeros without reading C).
template <typename R>
__global__ void __launch_bounds__(kZThreads)
    ZScaleKernel(int64_t m, int64_t n, R br, R bi, typename Vec2<R>::type* __restrict__ c,
                 int64_t ldc) {
  const int64_t i = int64_t(blockIdx.x) * kZThreads + threadIdx.x;
  if (i >= m) return;
  const bool zero = br == R(0) && bi == R(0);
  for (int64_t j = blockIdx.y; j < n; j += gridDim.y) {
    typename Vec2<R>::type out;
    if (zero) {
      out.x = R(0);
      out.y = R(0);
    } else {
      const auto cv = c[i + j * ldc];
      CMulScalar(cv.x, cv.y, br, bi, &out.x, &out.y);
    }
    c[i + j * ldc] = out;
  }
}

// Packs op(X) (rows x cols of the operation) into re/im planes.
template <typename R>
void ZPack(ZOp op, const typename Vec2<R>::type* x, int64_t ldx, int64_t rows, int64_t cols,
           R* re, R* im, int64_t ld, hipStream_t s) {
  const bool conj = op == ZOp::kC;
  if (op == ZOp::kN) {
    hipLaunchKernelGGL((ZPackKernel<R, false>), ZGrid(rows, cols), dim3(kZThreads), 0, s, x,
                       ldx, rows, cols, conj, re, im, ld);
  } else {
    hipLaunchKernelGGL((ZPackKernel<R, true>), ZGrid(cols, rows), dim3(kZThreads), 0, s, x,
                       ldx, rows, cols, conj, re, im, ld);
  }
  HWY_HIP_CHECK(hipGetLastError());
}

template <typename R>
void ZScale(int64_t m, int64_t n, const R* beta, typename Vec2<R>::type* c, int64_t ldc,
            hipStream_t s) {
  hipLaunchKernelGGL((ZScaleKernel<R>), ZGrid(m, n), dim3(kZThreads), 0, s, m, n, beta[0],
                     beta[1], c, ldc);
  HWY_HIP_CHECK(hipGetLastError());
}

// The product on device operands (m, n, k > 0, alpha != 0).
template <typename R>
void ZgemmDevice(ZOp oa, ZOp ob, int64_t m, int64_t n, int64_t k, const R* alpha,
                 const typename Vec2<R>::type* a, int64_t lda, const typename Vec2<R>::type* b,
                 int64_t ldb, const R* beta, typename Vec2<R>::type* c, int64_t ldc,
                 hipStream_t s) {
  R* pa = static_cast<R*>(Staging(kSlotPackA, size_t(2 * m) * k * sizeof(R)));
  R* pb = static_cast<R*>(Staging(kSlotPackB, size_t(2 * n) * k * sizeof(R)));
  R* p = static_cast<R*>(Staging(kSlotProd, size_t(4 * m) * n * sizeof(R)));
  ZPack<R>(oa, a, lda, m, k, pa, pa + m, 2 * m, s);
  ZPack<R>(ob, b, ldb, k, n, pb, pb + k * n, k, s);
  GemmDevice<R>(false, false, 2 * m, 2 * n, k, R(1), pa, 2 * m, pb, k, R(0), p, 2 * m, s);
  const int beta_kind = (beta[0] == R(0) && beta[1] == R(0)) ? 0
                        : (beta[0] == R(1) && beta[1] == R(0)) ? 1
                                                                : 2;
  hipLaunchKernelGGL((ZCombineKernel<R>), ZGrid(m, n), dim3(kZThreads), 0, s, p, m, n,
                     alpha[0], alpha[1], beta[0], beta[1], beta_kind, c, ldc);
  HWY_HIP_CHECK(hipGetLastError());
}

template <typename R>
bool IsZeroC(const R* z) {
  return z[0] == R(0) && z[1] == R(0);
}
template <typename R>
bool IsOneC(const R* z) {
  return z[0] == R(1) && z[1] == R(0);
}

template <typename T>
void CopyIn2D(T* dst, const T* src, int64_t rows, int64_t cols, int64_t ld, hipStream_t s) {
  HWY_HIP_CHECK(hipMemcpy2DAsync(dst, rows * sizeof(T), src, ld * sizeof(T), rows * sizeof(T),
                                 cols, hipMemcpyHostToDevice, s));
}

template <typename T>
void CopyOut2D(T* dst, int64_t ld, const T* src, int64_t rows, int64_t cols, hipStream_t s) {
  HWY_HIP_CHECK(hipMemcpy2DAsync(dst, ld * sizeof(T), src, rows * sizeof(T), rows * sizeof(T),
                                 cols, hipMemcpyDeviceToHost, s));
}

// Fortran entry: host or device operands; synchronizes.
template <typename R>
void ZgemmHost(const char* name, const char* transa, const char* transb, const blasint* mp,
               const blasint* np, const blasint* kp, const R* alpha, const R* a,
               const blasint* ldap, const R* b, const blasint* ldbp, const R* beta, R* c,
               const blasint* ldcp) {
  using C2 = typename Vec2<R>::type;
  const blasint m = *mp, n = *np, k = *kp, lda = *ldap, ldb = *ldbp, ldc = *ldcp;
  ZOp oa, ob;
  const blasint info = ZgemmInfo(*transa, *transb, m, n, k, lda, ldb, ldc, &oa, &ob);
  if (info != 0) {
    ReportError(name, info);
    return;
  }
  if (m == 0 || n == 0) return;
  const bool no_product = k == 0 || IsZeroC(alpha);
  if (no_product && IsOneC(beta)) return;
  ThreadContext& ctx = Context();
  hipStream_t s = ctx.stream;
  C2* dc = reinterpret_cast<C2*>(c);
  int64_t ldc_d = ldc;
  const bool c_dev = IsDevicePointer(c);
  if (!c_dev) {
    dc = static_cast<C2*>(Staging(kSlotC, size_t(m) * n * sizeof(C2)));
    ldc_d = m;
    if (!IsZeroC(beta)) CopyIn2D(dc, reinterpret_cast<const C2*>(c), m, n, ldc, s);
  }
  if (no_product) {
    ZScale<R>(m, n, beta, dc, ldc_d, s);
  } else {
    const int64_t ra = oa == ZOp::kN ? m : k, ca = oa == ZOp::kN ? k : m;
    const int64_t rb = ob == ZOp::kN ? k : n, cb = ob == ZOp::kN ? n : k;
    const C2* da = reinterpret_cast<const C2*>(a);
    const C2* db = reinterpret_cast<const C2*>(b);
    int64_t lda_d = lda, ldb_d = ldb;
    if (!IsDevicePointer(a)) {
      C2* buf = static_cast<C2*>(Staging(kSlotA, size_t(ra) * ca * sizeof(C2)));
      CopyIn2D(buf, da, ra, ca, lda, s);
      da = buf;
      lda_d = ra;
    }
    if (!IsDevicePointer(b)) {
      C2* buf = static_cast<C2*>(Staging(kSlotB, size_t(rb) * cb * sizeof(C2)));
      CopyIn2D(buf, db, rb, cb, ldb, s);
      db = buf;
      ldb_d = rb;
    }
    ZgemmDevice<R>(oa, ob, m, n, k, alpha, da, lda_d, db, ldb_d, beta, dc, ldc_d, s);
  }
  if (!c_dev) CopyOut2D(reinterpret_cast<C2*>(c), ldc, dc, m, n, s);
  HWY_HIP_CHECK(hipStreamSynchronize(s));
}

// Device-pointer form: asynchronous on `stream`; returns the info.
template <typename R>
int ZgemmDev(const char* name, char transa, char transb, int m, int n, int k, const R* alpha,
             const R* a, int lda, const R* b, int ldb, const R* beta, R* c, int ldc,
             hipStream_t s) {
  using C2 = typename Vec2<R>::type;
  ZOp oa, ob;
  const blasint info = ZgemmInfo(transa, transb, m, n, k, lda, ldb, ldc, &oa, &ob);
  if (info != 0) {
    ReportError(name, info);
    return info;
  }
  if (m == 0 || n == 0) return 0;
  const bool no_product = k == 0 || IsZeroC(alpha);
  if (no_product && IsOneC(beta)) return 0;
  if (no_product) {
    ZScale<R>(m, n, beta, reinterpret_cast<C2*>(c), ldc, s);
    return 0;
  }
  ZgemmDevice<R>(oa, ob, m, n, k, alpha, reinterpret_cast<const C2*>(a), lda,
                 reinterpret_cast<const C2*>(b), ldb, beta, reinterpret_cast<C2*>(c), ldc, s);
  return 0;
}

}  // namespace
}  // namespace hwy_amd

using hwy_amd::blasint;

extern "C" {

HWYBLAS_AMD_API void cgemm_(const char* transa, const char* transb, const blasint* m,
                            const blasint* n, const blasint* k, const float* alpha,
                            const float* a, const blasint* lda, const float* b,
                            const blasint* ldb, const float* beta, float* c,
                            const blasint* ldc) {
  hwy_amd::ZgemmHost<float>("CGEMM", transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c,
                            ldc);
}

HWYBLAS_AMD_API void zgemm_(const char* transa, const char* transb, const blasint* m,
                            const blasint* n, const blasint* k, const double* alpha,
                            const double* a, const blasint* lda, const double* b,
                            const blasint* ldb, const double* beta, double* c,
                            const blasint* ldc) {
  hwy_amd::ZgemmHost<double>("ZGEMM", transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c,
                             ldc);
}

HWYBLAS_AMD_API int hwyblas_cgemm_dev(char transa, char transb, int m, int n, int k,
                                      const float* alpha, const float* a, int lda,
                                      const float* b, int ldb, const float* beta, float* c,
                                      int ldc, void* stream) {
  return hwy_amd::ZgemmDev<float>("CGEMM", transa, transb, m, n, k, alpha, a, lda, b, ldb,
                                  beta, c, ldc, static_cast<hipStream_t>(stream));
}

HWYBLAS_AMD_API int hwyblas_zgemm_dev(char transa, char transb, int m, int n, int k,
                                      const double* alpha, const double* a, int lda,
                                      const double* b, int ldb, const double* beta, double* c,
                                      int ldc, void* stream) {
  return hwy_amd::ZgemmDev<double>("ZGEMM", transa, transb, m, n, k, alpha, a, lda, b, ldb,
                                   beta, c, ldc, static_cast<hipStream_t>(stream));
}

}  // extern "C"
