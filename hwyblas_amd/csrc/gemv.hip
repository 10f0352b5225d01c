// GEMV, y := alpha op(A) x + beta y (hwyBLAS src/gemv.cc:39-84 for the
// arguments, checks and quick returns; the arithmetic is HBM-bound: A is
// read once). op = N: the columns are split into chunks of kGemvChunk over
// a 2-D grid, one thread per row and chunk (every column's loads coalesced
// across the rows), partial sums written per chunk and added in chunk order
// by a finishing pass. op = T: one wave per four columns, the lanes
// striding down them, a fixed shuffle tree per sum. The store follows hwyBLAS's
// rule: r = alpha acc; beta == 1: r + y; beta != 0: fma(y, beta, r); beta == 0
// never reads y. Vectors with an increment other than 1 (or host vectors)
// are gathered into contiguous device buffers; a negative increment walks
// the vector from its end, as in the reference BLAS.
#include <algorithm>
#include <vector>

#include "entries.h"

namespace hwy_amd {
namespace {

constexpr int kGemvChunk = 256;  // columns per partial sum (op = N)
constexpr int kGemvRows = 256;   // rows (threads) per workgroup (op = N)
constexpr int kGemvUnroll = 8;
constexpr int kGemvTCols = 4;    // columns per wave (op = T)

// Element i of a BLAS vector of length len and increment inc.
__host__ __device__ inline int64_t VecIndex(int64_t i, int64_t len, int64_t inc) {
  return inc > 0 ? i * inc : (len - 1 - i) * (-inc);
}

template <typename T>
__global__ void GatherKernel(const T* __restrict__ x, int64_t len, int64_t inc, T* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < len) out[i] = x[VecIndex(i, len, inc)];
}

template <typename T>
__global__ void ScatterKernel(const T* __restrict__ y, int64_t len, int64_t inc, T* __restrict__ out) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < len) out[VecIndex(i, len, inc)] = y[i];
}

template <typename T>
__device__ __forceinline__ T Store(T acc, T alpha, T beta, const T* y) {
  const T r = acc * alpha;
  if (beta == T(0)) return r;
  return beta == T(1) ? r + *y : __builtin_fma(*y, beta, r);
}

// Partial sums p[chunk][i] = sum over the chunk's columns j (ascending) of
// a(i, j) x_j.
template <typename T>
__global__ void __launch_bounds__(kGemvRows)
    GemvNPartialKernel(int64_t m, int64_t n, const T* __restrict__ a, int64_t lda,
                       const T* __restrict__ x, T* __restrict__ p) {
  const int64_t i = int64_t(blockIdx.x) * kGemvRows + threadIdx.x;
  const int64_t j0 = int64_t(blockIdx.y) * kGemvChunk;
  const int64_t jn = min(int64_t(kGemvChunk), n - j0);
  if (i >= m) return;
  const T* ai = a + i + j0 * lda;
  const T* xj = x + j0;
  T acc = T(0);
  int64_t j = 0;
  for (; j + kGemvUnroll <= jn; j += kGemvUnroll) {
    T v[kGemvUnroll], w[kGemvUnroll];
#pragma unroll
    for (int u = 0; u < kGemvUnroll; ++u) {
      v[u] = ai[(j + u) * lda];
      w[u] = xj[j + u];
    }
#pragma unroll
    for (int u = 0; u < kGemvUnroll; ++u) acc = __builtin_fma(v[u], w[u], acc);
  }
  for (; j < jn; ++j) acc = __builtin_fma(ai[j * lda], xj[j], acc);
  p[int64_t(blockIdx.y) * m + i] = acc;
}

template <typename T>
__global__ void GemvNFinishKernel(int64_t m, int nchunks, const T* __restrict__ p, T alpha,
                                  T beta, T* __restrict__ y) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i >= m) return;
  T acc = T(0);
  for (int c = 0; c < nchunks; ++c) acc += p[int64_t(c) * m + i];
  y[i] = Store(acc, alpha, beta, y + i);
}

// y_j = alpha sum_i a(i, j) x_i + beta y_j, one wave per kGemvTCols
// columns: lane l sums rows l, l + 64, ... of each (ascending), then a
// shuffle tree per column. The columns of a wave share their loads of x
// (one x load per kGemvTCols column loads: 4.98 -> 5.77 TB/s at 16384^2
// FP64); each column's sum is the same as with one column per wave.
template <typename T>
__global__ void __launch_bounds__(256)
    GemvTKernel(int64_t m, int64_t n, const T* __restrict__ a, int64_t lda,
                const T* __restrict__ x, T alpha, T beta, T* __restrict__ y) {
  constexpr int kCols = kGemvTCols;
  const int lane = threadIdx.x & 63;
  const int64_t j0 = (int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6)) * kCols;
  if (j0 >= n) return;
  const T* aj[kCols];
  bool ok[kCols];
#pragma unroll
  for (int c = 0; c < kCols; ++c) {
    ok[c] = j0 + c < n;
    aj[c] = a + (ok[c] ? j0 + c : j0) * lda;
  }
  T acc[kCols];
#pragma unroll
  for (int c = 0; c < kCols; ++c) acc[c] = T(0);
  int64_t i = lane;
  for (; i + 64 * (kGemvUnroll - 1) < m; i += 64 * kGemvUnroll) {
    T w[kGemvUnroll], v[kCols][kGemvUnroll];
#pragma unroll
    for (int u = 0; u < kGemvUnroll; ++u) {
      w[u] = x[i + 64 * u];
#pragma unroll
      for (int c = 0; c < kCols; ++c) v[c][u] = aj[c][i + 64 * u];
    }
#pragma unroll
    for (int c = 0; c < kCols; ++c)
#pragma unroll
      for (int u = 0; u < kGemvUnroll; ++u) acc[c] = __builtin_fma(v[c][u], w[u], acc[c]);
  }
  for (; i < m; i += 64) {
    const T xi = x[i];
#pragma unroll
    for (int c = 0; c < kCols; ++c) acc[c] = __builtin_fma(aj[c][i], xi, acc[c]);
  }
#pragma unroll
  for (int c = 0; c < kCols; ++c) {
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) acc[c] += __shfl_xor(acc[c], off, 64);
    if (lane == 0 && ok[c]) y[j0 + c] = Store(acc[c], alpha, beta, y + j0 + c);
  }
}

unsigned Blocks(int64_t len, int per) { return unsigned((len + per - 1) / per); }

template <typename T>
void GemvDevice(bool trans, int64_t m, int64_t n, T alpha, const T* a, int64_t lda, const T* x,
                T beta, T* y, T* scratch, hipStream_t s) {
  if (!trans) {
    const int nchunks = int((n + kGemvChunk - 1) / kGemvChunk);
    hipLaunchKernelGGL(GemvNPartialKernel<T>, dim3(Blocks(m, kGemvRows), unsigned(nchunks)),
                       dim3(kGemvRows), 0, s, m, n, a, lda, x, scratch);
    hipLaunchKernelGGL(GemvNFinishKernel<T>, dim3(Blocks(m, 256)), dim3(256), 0, s, m, nchunks,
                       scratch, alpha, beta, y);
  } else {
    hipLaunchKernelGGL(GemvTKernel<T>, dim3(Blocks(n, 4 * kGemvTCols)), dim3(256), 0, s, m, n,
                       a, lda, x, alpha, beta, y);
  }
  HWY_HIP_CHECK(hipGetLastError());
}

// A BLAS vector on the device, contiguous: the caller's (device, inc 1),
// or a gathered copy in `slot` (host or strided).
template <typename T>
T* ContiguousVector(const T* v, int64_t len, int64_t inc, bool copy_in, int slot, hipStream_t s,
                    std::vector<T>* host_tmp) {
  const bool dev = IsDevicePointer(v);
  if (dev && inc == 1) return const_cast<T*>(v);
  T* buf = static_cast<T*>(Staging(slot, size_t(std::max<int64_t>(1, len)) * sizeof(T)));
  if (!copy_in) return buf;
  if (dev) {
    hipLaunchKernelGGL(GatherKernel<T>, dim3(Blocks(len, 256)), dim3(256), 0, s, v, len, inc,
                       buf);
    HWY_HIP_CHECK(hipGetLastError());
  } else {
    host_tmp->resize(size_t(len));
    for (int64_t i = 0; i < len; ++i) (*host_tmp)[size_t(i)] = v[VecIndex(i, len, inc)];
    HWY_HIP_CHECK(hipMemcpyAsync(buf, host_tmp->data(), size_t(len) * sizeof(T),
                                 hipMemcpyHostToDevice, s));
  }
  return buf;
}

template <typename T>
void GemvHost(const Caller& who, const char* transp, const blasint* mp, const blasint* np,
              const T* alphap, const T* a, const blasint* ldap, const T* x,
              const blasint* incxp, const T* betap, T* y, const blasint* incyp) {
  const blasint m = *mp, n = *np, lda = *ldap, incx = *incxp, incy = *incyp;
  Op op;
  blasint info = 0;
  if (!DecodeOp(*transp, &op)) info = 1;
  else if (m < 0) info = 2;
  else if (n < 0) info = 3;
  else if (lda < std::max(1, m)) info = 6;
  else if (incx == 0) info = 8;
  else if (incy == 0) info = 11;
  if (info != 0) return Reject(who, info);
  const T alpha = *alphap, beta = *betap;
  if (m == 0 || n == 0 || (alpha == T(0) && beta == T(1))) return;

  hipStream_t s = Context().stream;
  const bool trans = op == Op::kTrans;
  const int64_t len_x = trans ? m : n, len_y = trans ? n : m;
  std::vector<T> tmp_x, tmp_y;
  T* dy = ContiguousVector(y, len_y, incy, beta != T(0), kSlotC, s, &tmp_y);
  if (alpha == T(0)) {
    ScaleDevice<T>(len_y, 1, beta, dy, len_y, s);
  } else {
    const T* dx = ContiguousVector(x, len_x, incx, true, kSlotB, s, &tmp_x);
    const T* da = a;
    int64_t lda_d = lda;
    if (!IsDevicePointer(a)) {
      T* buf = static_cast<T*>(Staging(kSlotA, size_t(m) * n * sizeof(T)));
      HWY_HIP_CHECK(hwy_amd::Memcpy2DAsync(buf, size_t(m) * sizeof(T), a, size_t(lda) * sizeof(T),
                                           size_t(m) * sizeof(T), n, hipMemcpyHostToDevice, s));
      da = buf;
      lda_d = m;
    }
    T* part = nullptr;
    if (!trans) {
      const int64_t nchunks = (n + kGemvChunk - 1) / kGemvChunk;
      part = static_cast<T*>(Staging(kSlotD, size_t(nchunks) * m * sizeof(T)));
    }
    GemvDevice<T>(trans, m, n, alpha, da, lda_d, dx, beta, dy, part, s);
  }
  // Back to the caller's y (increment and place).
  if (dy != y) {
    if (IsDevicePointer(y)) {
      hipLaunchKernelGGL(ScatterKernel<T>, dim3(Blocks(len_y, 256)), dim3(256), 0, s, dy, len_y,
                         int64_t(incy), y);
      HWY_HIP_CHECK(hipGetLastError());
    } else {
      tmp_y.resize(size_t(len_y));
      HWY_HIP_CHECK(hipMemcpyAsync(tmp_y.data(), dy, size_t(len_y) * sizeof(T),
                                   hipMemcpyDeviceToHost, s));
      HWY_HIP_CHECK(hipStreamSynchronize(s));
      for (int64_t i = 0; i < len_y; ++i) y[VecIndex(i, len_y, incy)] = tmp_y[size_t(i)];
    }
  }
  HWY_HIP_CHECK(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// Complex GEMV (cgemv_, zgemv_; the CBLAS forms go through ComplexGemv). op:
// 'N', 'T', 'C', or 'R' = conjugated and not transposed (row-major CBLAS
// ConjTrans is 'R' on the column-major storage, so that x and y need no
// conjugated copies). Products are complex FMAs in ascending k; the store is
// r = alpha acc; beta == 1: r + y; beta != 0: r + beta y; beta == 0 never
// reads y.

template <typename R>
struct Cplx2;
template <>
struct Cplx2<double> {
  using type = double2;
};
template <>
struct Cplx2<float> {
  using type = float2;
};

template <typename C2>
__device__ __forceinline__ void CFma(C2 a, C2 x, bool conj_a, C2* acc) {
  const auto ai = conj_a ? -a.y : a.y;
  acc->x = __builtin_fma(a.x, x.x, __builtin_fma(-ai, x.y, acc->x));
  acc->y = __builtin_fma(a.x, x.y, __builtin_fma(ai, x.x, acc->y));
}

template <typename R, bool CONJ>
__global__ void __launch_bounds__(kGemvRows)
    ZGemvNPartialKernel(int64_t m, int64_t n, const typename Cplx2<R>::type* __restrict__ a,
                        int64_t lda, const typename Cplx2<R>::type* __restrict__ x,
                        typename Cplx2<R>::type* __restrict__ part) {
  const int64_t i = int64_t(blockIdx.x) * kGemvRows + threadIdx.x;
  if (i >= m) return;
  const int64_t j0 = int64_t(blockIdx.y) * kGemvChunk;
  const int64_t j1 = j0 + kGemvChunk < n ? j0 + kGemvChunk : n;
  typename Cplx2<R>::type acc;
  acc.x = R(0);
  acc.y = R(0);
#pragma unroll 4
  for (int64_t j = j0; j < j1; ++j) CFma(a[i + j * lda], x[j], CONJ, &acc);
  part[int64_t(blockIdx.y) * m + i] = acc;
}

template <typename C2, typename R>
__device__ __forceinline__ C2 ZStore(C2 acc, R ar, R ai, R br, R bi, int beta_kind, const C2* y) {
  C2 r;
  r.x = acc.x * ar - acc.y * ai;
  r.y = acc.y * ar + acc.x * ai;
  if (beta_kind == 1) {
    r.x = r.x + y->x;
    r.y = r.y + y->y;
  } else if (beta_kind == 2) {
    r.x = r.x + (y->x * br - y->y * bi);
    r.y = r.y + (y->y * br + y->x * bi);
  }
  return r;
}

template <typename R>
__global__ void __launch_bounds__(256)
    ZGemvNFinishKernel(int64_t m, int64_t nchunks, const typename Cplx2<R>::type* __restrict__ part,
                       R ar, R ai, R br, R bi, int beta_kind,
                       typename Cplx2<R>::type* __restrict__ y) {
  const int64_t i = int64_t(blockIdx.x) * 256 + threadIdx.x;
  if (i >= m) return;
  typename Cplx2<R>::type acc;
  acc.x = R(0);
  acc.y = R(0);
  for (int64_t c = 0; c < nchunks; ++c) {
    acc.x += part[c * m + i].x;
    acc.y += part[c * m + i].y;
  }
  y[i] = ZStore(acc, ar, ai, br, bi, beta_kind, &y[i]);
}

// op T or C: one wave per column, lanes striding down it.
template <typename R, bool CONJ>
__global__ void __launch_bounds__(256)
    ZGemvTKernel(int64_t m, int64_t n, const typename Cplx2<R>::type* __restrict__ a, int64_t lda,
                 const typename Cplx2<R>::type* __restrict__ x, R ar, R ai, R br, R bi,
                 int beta_kind, typename Cplx2<R>::type* __restrict__ y) {
  const int lane = threadIdx.x & 63;
  const int64_t j = int64_t(blockIdx.x) * 4 + (threadIdx.x >> 6);
  if (j >= n) return;
  typename Cplx2<R>::type acc;
  acc.x = R(0);
  acc.y = R(0);
  const auto* col = a + j * lda;
  for (int64_t i = lane; i < m; i += 64) CFma(col[i], x[i], CONJ, &acc);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    acc.x += __shfl_xor(acc.x, off);
    acc.y += __shfl_xor(acc.y, off);
  }
  if (lane == 0) y[j] = ZStore(acc, ar, ai, br, bi, beta_kind, &y[j]);
}

template <typename R>
void ZgemvHost(const Caller& who, char opc, blasint m, blasint n, const R* alpha, const R* a,
               blasint lda, const R* x, blasint incx, const R* beta, R* y, blasint incy) {
  using C2 = typename Cplx2<R>::type;
  const char oc = char(opc & ~0x20);
  blasint info = 0;
  if (oc != 'N' && oc != 'T' && oc != 'C' && oc != 'R') info = 1;
  else if (m < 0) info = 2;
  else if (n < 0) info = 3;
  else if (lda < std::max(1, m)) info = 6;
  else if (incx == 0) info = 8;
  else if (incy == 0) info = 11;
  if (info != 0) return Reject(who, info);
  const bool alpha0 = alpha[0] == R(0) && alpha[1] == R(0);
  const bool beta0 = beta[0] == R(0) && beta[1] == R(0);
  const bool beta1 = beta[0] == R(1) && beta[1] == R(0);
  if (m == 0 || n == 0 || (alpha0 && beta1)) return;
  const int beta_kind = beta0 ? 0 : beta1 ? 1 : 2;

  hipStream_t s = Context().stream;
  const bool trans = oc == 'T' || oc == 'C';
  const bool conj = oc == 'C' || oc == 'R';
  const int64_t len_x = trans ? m : n, len_y = trans ? n : m;
  std::vector<C2> tmp_x, tmp_y;
  C2* y2 = reinterpret_cast<C2*>(y);
  C2* dy = ContiguousVector(y2, len_y, incy, !beta0, kSlotC, s, &tmp_y);
  if (alpha0) {
    hipLaunchKernelGGL(ZGemvNFinishKernel<R>, dim3(Blocks(len_y, 256)), dim3(256), 0, s, len_y,
                       int64_t(0), static_cast<const C2*>(nullptr), alpha[0], alpha[1], beta[0],
                       beta[1], beta_kind, dy);
  } else {
    const C2* dx = ContiguousVector(reinterpret_cast<const C2*>(x), len_x, incx, true, kSlotB, s,
                                    &tmp_x);
    const C2* da = reinterpret_cast<const C2*>(a);
    int64_t lda_d = lda;
    if (!IsDevicePointer(a)) {
      C2* buf = static_cast<C2*>(Staging(kSlotA, size_t(m) * n * sizeof(C2)));
      HWY_HIP_CHECK(hwy_amd::Memcpy2DAsync(buf, size_t(m) * sizeof(C2), da,
                                           size_t(lda) * sizeof(C2), size_t(m) * sizeof(C2), n,
                                           hipMemcpyHostToDevice, s));
      da = buf;
      lda_d = m;
    }
    if (!trans) {
      const int64_t nchunks = (n + kGemvChunk - 1) / kGemvChunk;
      C2* part = static_cast<C2*>(Staging(kSlotD, size_t(nchunks) * m * sizeof(C2)));
      const dim3 grid(unsigned(Blocks(m, kGemvRows)), unsigned(nchunks));
      if (conj)
        hipLaunchKernelGGL((ZGemvNPartialKernel<R, true>), grid, dim3(kGemvRows), 0, s, m, n, da,
                           lda_d, dx, part);
      else
        hipLaunchKernelGGL((ZGemvNPartialKernel<R, false>), grid, dim3(kGemvRows), 0, s, m, n, da,
                           lda_d, dx, part);
      hipLaunchKernelGGL(ZGemvNFinishKernel<R>, dim3(Blocks(m, 256)), dim3(256), 0, s, m,
                         nchunks, part, alpha[0], alpha[1], beta[0], beta[1], beta_kind, dy);
    } else {
      if (conj)
        hipLaunchKernelGGL((ZGemvTKernel<R, true>), dim3(Blocks(n, 4)), dim3(256), 0, s, m, n,
                           da, lda_d, dx, alpha[0], alpha[1], beta[0], beta[1], beta_kind, dy);
      else
        hipLaunchKernelGGL((ZGemvTKernel<R, false>), dim3(Blocks(n, 4)), dim3(256), 0, s, m, n,
                           da, lda_d, dx, alpha[0], alpha[1], beta[0], beta[1], beta_kind, dy);
    }
  }
  HWY_HIP_CHECK(hipGetLastError());
  if (dy != y2) {
    if (IsDevicePointer(y)) {
      hipLaunchKernelGGL(ScatterKernel<C2>, dim3(Blocks(len_y, 256)), dim3(256), 0, s, dy, len_y,
                         int64_t(incy), y2);
      HWY_HIP_CHECK(hipGetLastError());
    } else {
      tmp_y.resize(size_t(len_y));
      HWY_HIP_CHECK(hipMemcpyAsync(tmp_y.data(), dy, size_t(len_y) * sizeof(C2),
                                   hipMemcpyDeviceToHost, s));
      HWY_HIP_CHECK(hipStreamSynchronize(s));
      for (int64_t i = 0; i < len_y; ++i) y2[VecIndex(i, len_y, incy)] = tmp_y[size_t(i)];
    }
  }
  HWY_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace

#define HWY_DEFINE_CORES(T)                                                                   \
  void Gemv(const Caller& who, const char* trans, const blasint* m, const blasint* n,         \
            const T* alpha, const T* a, const blasint* lda, const T* x, const blasint* incx,  \
            const T* beta, T* y, const blasint* incy) {                                       \
    GemvHost(who, trans, m, n, alpha, a, lda, x, incx, beta, y, incy);                        \
  }                                                                                           \
  void ComplexGemv(const Caller& who, const char* op, const blasint* m, const blasint* n,     \
                   const T* alpha, const T* a, const blasint* lda, const T* x,                \
                   const blasint* incx, const T* beta, T* y, const blasint* incy) {           \
    ZgemvHost<T>(who, *op, *m, *n, alpha, a, *lda, x, *incx, beta, y, *incy);                 \
  }
HWY_DEFINE_CORES(float)
HWY_DEFINE_CORES(double)
#undef HWY_DEFINE_CORES

}  // namespace hwy_amd

using hwy_amd::blasint;

extern "C" {

HWYBLAS_AMD_API void sgemv_(const char* trans, const blasint* m, const blasint* n,
                            const float* alpha, const float* a, const blasint* lda,
                            const float* x, const blasint* incx, const float* beta, float* y,
                            const blasint* incy) {
  hwy_amd::Gemv("SGEMV", trans, m, n, alpha, a, lda, x, incx, beta, y, incy);
}

HWYBLAS_AMD_API void dgemv_(const char* trans, const blasint* m, const blasint* n,
                            const double* alpha, const double* a, const blasint* lda,
                            const double* x, const blasint* incx, const double* beta,
                            double* y, const blasint* incy) {
  hwy_amd::Gemv("DGEMV", trans, m, n, alpha, a, lda, x, incx, beta, y, incy);
}

HWYBLAS_AMD_API void cgemv_(const char* trans, const blasint* m, const blasint* n,
                            const float* alpha, const float* a, const blasint* lda,
                            const float* x, const blasint* incx, const float* beta, float* y,
                            const blasint* incy) {
  // 'R' is internal to the CBLAS row-major mapping, not a Fortran option.
  const char t = (*trans == 'R' || *trans == 'r') ? '?' : *trans;
  hwy_amd::ComplexGemv("CGEMV", &t, m, n, alpha, a, lda, x, incx, beta, y, incy);
}

HWYBLAS_AMD_API void zgemv_(const char* trans, const blasint* m, const blasint* n,
                            const double* alpha, const double* a, const blasint* lda,
                            const double* x, const blasint* incx, const double* beta,
                            double* y, const blasint* incy) {
  const char t = (*trans == 'R' || *trans == 'r') ? '?' : *trans;
  hwy_amd::ComplexGemv("ZGEMV", &t, m, n, alpha, a, lda, x, incx, beta, y, incy);
}

}  // extern "C"
