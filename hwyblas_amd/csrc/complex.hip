// Complex Level-3 on the GPU: GEMM (cgemm_, zgemm_) and TRSM (ctrsm_,
// ztrsm_), with their device-pointer forms. op = N, T or C (conjugate
// transpose).
//
// GEMM, C := alpha op(A) op(B) + beta C (hwyBLAS src/gemm.cc:73-132 for the
// checks, info positions and quick returns, src/gemm.cc:160-176 for the
// entries).
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
//
// TRSM, op(A) X = alpha B or X op(A) = alpha B (hwyBLAS src/trsm.cc:87-180):
// B := alpha B first (alpha == 0 leaves zeros and never reads A), then the
// triangle is halved at multiples of 64 down to diagonal blocks of at most
// 64; each off-diagonal update is the complex GEMM above with alpha = -1,
// beta = 1, and each diagonal block is solved by one kernel: a wave per
// right-hand side, lane e holding solve row e, the block's triangle in LDS,
// x_i = (b_i - sum_{q<i} t(i, q) x_q) (1 / t(i, i)) with the inverse from a
// complex division (src/trsm-inl.h:60-70, conjugated diagonal for op = C).
#include <algorithm>

#include "entries.h"

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
// part: 0 all of C, 1 its lower triangle (i >= j), 2 its upper one; the
// other entries are not touched. real_diag: the diagonal's imaginary parts
// are set to zero (HERK). p == nullptr: no product (C := beta C).
template <typename R>
__global__ void __launch_bounds__(kZThreads)
    ZCombineKernel(const R* __restrict__ p, int64_t m, int64_t n, R ar, R ai, R br, R bi,
                   int beta_kind, typename Vec2<R>::type* __restrict__ c, int64_t ldc, int part,
                   bool real_diag) {
  const int64_t i = int64_t(blockIdx.x) * kZThreads + threadIdx.x;
  if (i >= m) return;
  const int64_t ldp = 2 * m;
  for (int64_t j = blockIdx.y; j < n; j += gridDim.y) {
    if ((part == 1 && i < j) || (part == 2 && i > j)) continue;
    R abr = R(0), abi = R(0);
    if (p) {
      const R rr = p[i + j * ldp];              // sum ar br
      const R ir = p[m + i + j * ldp];          // sum ai br
      const R ri = p[i + (n + j) * ldp];        // sum ar bi
      const R ii = p[m + i + (n + j) * ldp];    // sum ai bi
      abr = rr - ii;
      abi = ir + ri;
    }
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
    out.y = (real_diag && i == j) ? R(0) : xi;
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
                 hipStream_t s, int part = 0, bool real_diag = false) {
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
                     alpha[0], alpha[1], beta[0], beta[1], beta_kind, c, ldc, part, real_diag);
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
  HWY_HIP_CHECK(hwy_amd::Memcpy2DAsync(dst, rows * sizeof(T), src, ld * sizeof(T), rows * sizeof(T),
                                       cols, hipMemcpyHostToDevice, s));
}

template <typename T>
void CopyOut2D(T* dst, int64_t ld, const T* src, int64_t rows, int64_t cols, hipStream_t s) {
  HWY_HIP_CHECK(hwy_amd::Memcpy2DAsync(dst, ld * sizeof(T), src, rows * sizeof(T), rows * sizeof(T),
                                       cols, hipMemcpyDeviceToHost, s));
}

// Fortran entry: host or device operands; synchronizes.
template <typename R>
void ZgemmHost(const Caller& who, const char* transa, const char* transb, const blasint* mp,
               const blasint* np, const blasint* kp, const R* alpha, const R* a,
               const blasint* ldap, const R* b, const blasint* ldbp, const R* beta, R* c,
               const blasint* ldcp) {
  using C2 = typename Vec2<R>::type;
  const blasint m = *mp, n = *np, k = *kp, lda = *ldap, ldb = *ldbp, ldc = *ldcp;
  ZOp oa, ob;
  const blasint info = ZgemmInfo(*transa, *transb, m, n, k, lda, ldb, ldc, &oa, &ob);
  if (info != 0) {
    Reject(who, info);
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
int ZgemmDev(const Caller& who, char transa, char transb, int m, int n, int k, const R* alpha,
             const R* a, int lda, const R* b, int ldb, const R* beta, R* c, int ldc,
             hipStream_t s) {
  using C2 = typename Vec2<R>::type;
  ZOp oa, ob;
  const blasint info = ZgemmInfo(transa, transb, m, n, k, lda, ldb, ldc, &oa, &ob);
  if (info != 0) {
    Reject(who, info);
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


// ---------------------------------------------------------------------------
// TRSM.

constexpr int kZNB = 64;     // largest diagonal block
constexpr int kZWaves = 4;   // right-hand sides per workgroup (one per wave)

// 1 / z by Smith's scaling (no overflow for representable quotients).
template <typename C2>
__device__ __forceinline__ C2 CRecip(C2 z) {
  C2 out;
  if (fabs(z.x) >= fabs(z.y)) {
    const auto r = z.y / z.x, den = z.x + z.y * r;
    out.x = 1 / den;
    out.y = -r / den;
  } else {
    const auto r = z.x / z.y, den = z.x * r + z.y;
    out.x = r / den;
    out.y = -1 / den;
  }
  return out;
}

__device__ __forceinline__ double LaneBroadcast(double v, int i) {
  const long long bits = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane(int(bits & 0xFFFFFFFFll), i);
  const int hi = __builtin_amdgcn_readlane(int(bits >> 32), i);
  return __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
}
__device__ __forceinline__ float LaneBroadcast(float v, int i) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), i));
}

// The triangle op(A) as a strided view: element (i, j) = data[i rs + j cs],
// conjugated when conj.
template <typename R>
struct ZTri {
  const typename Vec2<R>::type* data;
  int64_t rs, cs;
  bool lower, unit, conj;
};

template <typename R>
ZTri<R> ZSub(const ZTri<R>& t, int64_t i0, int64_t j0) {
  ZTri<R> s = t;
  s.data += i0 * t.rs + j0 * t.cs;
  return s;
}

template <typename R>
struct ZDiagDesc {
  const typename Vec2<R>::type* a;
  int64_t rs, cs;
  typename Vec2<R>::type* b;
  int64_t ldb;
  int r, nrhs;
  bool asc, unit, conj;
};

// One diagonal block (order r <= kZNB). The solve matrix is L = T (left) or
// T^T (right), walked ascending for a lower L; a right-hand side is a column
// of B (left) or a row (right).
template <typename R, bool LEFT>
__global__ void __launch_bounds__(64 * kZWaves) ZDiagKernel(ZDiagDesc<R> d) {
  using C2 = typename Vec2<R>::type;
  __shared__ C2 st[kZNB][kZNB + 1];  // st[e][q] = L(e, q), q < e, in solve order
  __shared__ C2 sinv[kZNB];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = d.r;
  auto phys = [&](int e) { return d.asc ? e : r - 1 - e; };
  auto lval = [&](int e, int q) {
    const int64_t i = LEFT ? phys(e) : phys(q), j = LEFT ? phys(q) : phys(e);
    C2 v = d.a[i * d.rs + j * d.cs];
    if (d.conj) v.y = -v.y;
    return v;
  };
  for (int q = wave; q < kZNB; q += kZWaves) {
    C2 v;
    v.x = R(0);
    v.y = R(0);
    if (lane < r && q < lane) v = lval(lane, q);
    st[lane][q] = v;
  }
  if (wave == 0) {
    C2 inv;
    inv.x = R(1);
    inv.y = R(0);
    if (lane < r && !d.unit) inv = CRecip(lval(lane, lane));
    sinv[lane] = inv;
  }
  __syncthreads();
  const int c = int(blockIdx.x) * kZWaves + wave;
  if (c >= d.nrhs) return;
  const int64_t e_phys = lane < r ? phys(lane) : 0;
  C2* at = LEFT ? d.b + e_phys + int64_t(c) * d.ldb : d.b + c + e_phys * d.ldb;
  C2 x;
  x.x = R(0);
  x.y = R(0);
  if (lane < r) x = *at;
  const C2 myinv = sinv[lane];
  // All kZNB steps, unrolled: steps at or past r touch zeros only (padding
  // rows hold x = 0 and reciprocal 1, padding triangle entries are 0).
#pragma unroll
  for (int i = 0; i < kZNB; ++i) {
    if (lane == i) {
      const R xr = x.x * myinv.x - x.y * myinv.y;
      const R xi = x.y * myinv.x + x.x * myinv.y;
      x.x = xr;
      x.y = xi;
    }
    const R br = LaneBroadcast(x.x, i), bi = LaneBroadcast(x.y, i);
    if (lane > i) {
      const C2 t = st[lane][i];
      x.x = __builtin_fma(-t.x, br, __builtin_fma(t.y, bi, x.x));
      x.y = __builtin_fma(-t.x, bi, __builtin_fma(-t.y, br, x.y));
    }
  }
  if (lane < r) *at = x;
}

template <typename R>
void ZDiagSolve(bool left, const ZTri<R>& t, int r, typename Vec2<R>::type* b, int64_t ldb,
                int64_t other, hipStream_t s) {
  ZDiagDesc<R> d;
  d.a = t.data;
  d.rs = t.rs;
  d.cs = t.cs;
  d.b = b;
  d.ldb = ldb;
  d.r = r;
  d.nrhs = int(other);
  // L = T (left) or T^T (right); a lower L solves ascending.
  d.asc = left ? t.lower : !t.lower;
  d.unit = t.unit;
  d.conj = t.conj;
  const dim3 grid(unsigned((other + kZWaves - 1) / kZWaves));
  if (left)
    hipLaunchKernelGGL((ZDiagKernel<R, true>), grid, dim3(64 * kZWaves), 0, s, d);
  else
    hipLaunchKernelGGL((ZDiagKernel<R, false>), grid, dim3(64 * kZWaves), 0, s, d);
  HWY_HIP_CHECK(hipGetLastError());
}

// The GEMM operand and op of a strided view (one of rs, cs is 1).
template <typename R>
void ViewOperand(const ZTri<R>& t, ZOp* op, int64_t* ld) {
  if (t.rs == 1) {
    *op = ZOp::kN;
    *ld = t.cs;
  } else {
    *op = t.conj ? ZOp::kC : ZOp::kT;
    *ld = t.rs;
  }
}

template <typename R>
void ZRecurse(bool left, const ZTri<R>& t, int64_t r, int64_t other,
              typename Vec2<R>::type* b, int64_t ldb, hipStream_t s) {
  if (r <= kZNB) {
    ZDiagSolve<R>(left, t, int(r), b, ldb, other, s);
    return;
  }
  const R minus_one[2] = {R(-1), R(0)}, one[2] = {R(1), R(0)};
  const int64_t r1 = std::max<int64_t>(kZNB, (r / 2) / kZNB * kZNB);
  const int64_t r2 = r - r1;
  const ZTri<R> t22 = ZSub(t, r1, r1);
  const ZTri<R> tof = t.lower ? ZSub(t, r1, 0) : ZSub(t, 0, r1);
  ZOp op;
  int64_t ldt;
  ViewOperand(tof, &op, &ldt);
  if (left) {
    auto* b1 = b;
    auto* b2 = b + r1;
    if (t.lower) {
      ZRecurse<R>(left, t, r1, other, b1, ldb, s);
      ZgemmDevice<R>(op, ZOp::kN, r2, other, r1, minus_one, tof.data, ldt, b1, ldb, one, b2, ldb, s);
      ZRecurse<R>(left, t22, r2, other, b2, ldb, s);
    } else {
      ZRecurse<R>(left, t22, r2, other, b2, ldb, s);
      ZgemmDevice<R>(op, ZOp::kN, r1, other, r2, minus_one, tof.data, ldt, b2, ldb, one, b1, ldb, s);
      ZRecurse<R>(left, t, r1, other, b1, ldb, s);
    }
  } else {
    auto* b1 = b;
    auto* b2 = b + r1 * ldb;
    if (t.lower) {
      ZRecurse<R>(left, t22, r2, other, b2, ldb, s);
      ZgemmDevice<R>(ZOp::kN, op, other, r1, r2, minus_one, b2, ldb, tof.data, ldt, one, b1, ldb, s);
      ZRecurse<R>(left, t, r1, other, b1, ldb, s);
    } else {
      ZRecurse<R>(left, t, r1, other, b1, ldb, s);
      ZgemmDevice<R>(ZOp::kN, op, other, r2, r1, minus_one, b1, ldb, tof.data, ldt, one, b2, ldb, s);
      ZRecurse<R>(left, t22, r2, other, b2, ldb, s);
    }
  }
}

// Info of the first invalid TRSM argument, or 0 (src/trsm.cc:145-168).
blasint ZtrsmInfo(char side, char uplo, char transa, char diag, blasint m, blasint n,
                  blasint lda, blasint ldb, bool* left, bool* lower, ZOp* op, bool* unit) {
  const bool vs = DecodeSide(side, left);
  const bool vu = DecodeUplo(uplo, lower);
  const bool vt = DecodeZOp(transa, op);
  const bool vd = DecodeDiag(diag, unit);
  const blasint ka = *left ? m : n;
  if (!vs) return 1;
  if (!vu) return 2;
  if (!vt) return 3;
  if (!vd) return 4;
  if (m < 0) return 5;
  if (n < 0) return 6;
  if (lda < std::max(1, ka)) return 9;
  if (ldb < std::max(1, m)) return 11;
  return 0;
}

// B := alpha B, then the solve (m, n > 0; device operands).
template <typename R>
void ZtrsmDevice(bool left, bool lower, ZOp op, bool unit, int64_t m, int64_t n, const R* alpha,
                 const typename Vec2<R>::type* a, int64_t lda, typename Vec2<R>::type* b,
                 int64_t ldb, hipStream_t s) {
  if (!IsOneC(alpha)) ZScale<R>(m, n, alpha, b, ldb, s);
  if (IsZeroC(alpha)) return;
  ZTri<R> t;
  const bool trans = op != ZOp::kN;
  t.data = a;
  t.rs = trans ? lda : 1;
  t.cs = trans ? 1 : lda;
  t.lower = trans ? !lower : lower;
  t.unit = unit;
  t.conj = op == ZOp::kC;
  ZRecurse<R>(left, t, left ? m : n, left ? n : m, b, ldb, s);
}

template <typename R>
void ZtrsmHost(const Caller& who, const char* sidep, const char* uplop, const char* transp,
               const char* diagp, const blasint* mp, const blasint* np, const R* alpha,
               const R* a, const blasint* ldap, R* b, const blasint* ldbp) {
  using C2 = typename Vec2<R>::type;
  const blasint m = *mp, n = *np, lda = *ldap, ldb = *ldbp;
  bool left, lower, unit;
  ZOp op;
  const blasint info =
      ZtrsmInfo(*sidep, *uplop, *transp, *diagp, m, n, lda, ldb, &left, &lower, &op, &unit);
  if (info != 0) {
    Reject(who, info);
    return;
  }
  if (m == 0 || n == 0) return;
  ThreadContext& ctx = Context();
  hipStream_t s = ctx.stream;
  const bool b_dev = IsDevicePointer(b);
  C2* db = reinterpret_cast<C2*>(b);
  int64_t ldb_d = ldb;
  if (!b_dev) {
    db = static_cast<C2*>(Staging(kSlotB, size_t(m) * n * sizeof(C2)));
    ldb_d = m;
    if (!IsZeroC(alpha)) CopyIn2D(db, reinterpret_cast<const C2*>(b), m, n, ldb, s);
  }
  const C2* da = reinterpret_cast<const C2*>(a);
  int64_t lda_d = lda;
  if (!IsZeroC(alpha) && !IsDevicePointer(a)) {
    const int64_t ka = left ? m : n;
    C2* buf = static_cast<C2*>(Staging(kSlotA, size_t(ka) * ka * sizeof(C2)));
    CopyIn2D(buf, da, ka, ka, lda, s);
    da = buf;
    lda_d = ka;
  }
  ZtrsmDevice<R>(left, lower, op, unit, m, n, alpha, da, lda_d, db, ldb_d, s);
  if (!b_dev) CopyOut2D(reinterpret_cast<C2*>(b), ldb, db, m, n, s);
  HWY_HIP_CHECK(hipStreamSynchronize(s));
}

template <typename R>
int ZtrsmDev(const Caller& who, char side, char uplo, char transa, char diag, int m, int n,
             const R* alpha, const R* a, int lda, R* b, int ldb, hipStream_t s) {
  bool left, lower, unit;
  ZOp op;
  const blasint info =
      ZtrsmInfo(side, uplo, transa, diag, m, n, lda, ldb, &left, &lower, &op, &unit);
  if (info != 0) {
    Reject(who, info);
    return info;
  }
  if (m == 0 || n == 0) return 0;
  using C2 = typename Vec2<R>::type;
  ZtrsmDevice<R>(left, lower, op, unit, m, n, alpha, reinterpret_cast<const C2*>(a), lda,
                 reinterpret_cast<C2*>(b), ldb, s);
  return 0;
}

// ---------------------------------------------------------------------------
// SYRK and HERK (hwyBLAS src/level3.cc:78-135): C := alpha op(A) op(A)' +
// beta C on the uplo triangle, ' = ^T (SYRK: trans N or T, complex alpha and
// beta) or ^H (HERK: trans N or C, real alpha and beta, the diagonal's
// imaginary parts set to zero). The product is the complex GEMM above over
// all of n x n (twice the triangle's flops); the combining pass writes the
// triangle only.
template <typename R>
void ZSyrkHost(const Caller& who, bool herm, const char* uplop, const char* transp,
               const blasint* np, const blasint* kp, const R* alpha_in, const R* a,
               const blasint* ldap, const R* beta_in, R* c, const blasint* ldcp) {
  using C2 = typename Vec2<R>::type;
  const blasint n = *np, k = *kp, lda = *ldap, ldc = *ldcp;
  bool lower = false;
  const bool vu = DecodeUplo(*uplop, &lower);
  ZOp op;
  bool vt = DecodeZOp(*transp, &op);
  if (vt) vt = op != (herm ? ZOp::kT : ZOp::kC);
  const bool no_trans = op == ZOp::kN;
  const blasint rows_a = no_trans ? n : k;
  blasint info = 0;
  if (!vu) info = 1;
  else if (!vt) info = 2;
  else if (n < 0) info = 3;
  else if (k < 0) info = 4;
  else if (lda < std::max(1, rows_a)) info = 7;
  else if (ldc < std::max(1, n)) info = 10;
  if (info != 0) {
    Reject(who, info);
    return;
  }
  const R alpha[2] = {alpha_in[0], herm ? R(0) : alpha_in[1]};
  const R beta[2] = {beta_in[0], herm ? R(0) : beta_in[1]};
  const bool no_product = k == 0 || IsZeroC(alpha);
  if (n == 0 || (no_product && IsOneC(beta))) return;
  hipStream_t s = Context().stream;
  const bool c_dev = IsDevicePointer(c);
  C2* dc = reinterpret_cast<C2*>(c);
  int64_t ldc_d = ldc;
  if (!c_dev) {
    dc = static_cast<C2*>(Staging(kSlotC, size_t(n) * n * sizeof(C2)));
    ldc_d = n;
    CopyIn2D(dc, reinterpret_cast<const C2*>(c), n, n, ldc, s);  // the other triangle is kept
  }
  const int part = lower ? 1 : 2;
  if (no_product) {
    const int beta_kind = IsZeroC(beta) ? 0 : IsOneC(beta) ? 1 : 2;
    hipLaunchKernelGGL((ZCombineKernel<R>), ZGrid(n, n), dim3(kZThreads), 0, s,
                       static_cast<const R*>(nullptr), int64_t(n), int64_t(n), R(0), R(0),
                       beta[0], beta[1], beta_kind, dc, ldc_d, part, herm);
    HWY_HIP_CHECK(hipGetLastError());
  } else {
    const C2* da = reinterpret_cast<const C2*>(a);
    int64_t lda_d = lda;
    const int64_t ra = rows_a, ca = no_trans ? k : n;
    if (!IsDevicePointer(a)) {
      C2* buf = static_cast<C2*>(Staging(kSlotA, size_t(ra) * ca * sizeof(C2)));
      CopyIn2D(buf, da, ra, ca, lda, s);
      da = buf;
      lda_d = ra;
    }
    const ZOp second = herm ? ZOp::kC : ZOp::kT;
    ZgemmDevice<R>(no_trans ? ZOp::kN : second, no_trans ? second : ZOp::kN, n, n, k, alpha, da,
                   lda_d, da, lda_d, beta, dc, ldc_d, s, part, herm);
  }
  if (!c_dev) CopyOut2D(reinterpret_cast<C2*>(c), ldc, dc, n, n, s);
  HWY_HIP_CHECK(hipStreamSynchronize(s));
}

}  // namespace

#define HWY_DEFINE_CORES(R)                                                                   \
  void ComplexGemm(const Caller& who, const char* transa, const char* transb,                 \
                   const blasint* m, const blasint* n, const blasint* k, const R* alpha,      \
                   const R* a, const blasint* lda, const R* b, const blasint* ldb,            \
                   const R* beta, R* c, const blasint* ldc) {                                 \
    ZgemmHost<R>(who, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);          \
  }                                                                                           \
  void ComplexTrsm(const Caller& who, const char* side, const char* uplo,                     \
                   const char* transa, const char* diag, const blasint* m, const blasint* n,  \
                   const R* alpha, const R* a, const blasint* lda, R* b, const blasint* ldb) { \
    ZtrsmHost<R>(who, side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb);                 \
  }
HWY_DEFINE_CORES(float)
HWY_DEFINE_CORES(double)
#undef HWY_DEFINE_CORES

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

HWYBLAS_AMD_API void ctrsm_(const char* side, const char* uplo, const char* transa,
                            const char* diag, const blasint* m, const blasint* n,
                            const float* alpha, const float* a, const blasint* lda, float* b,
                            const blasint* ldb) {
  hwy_amd::ZtrsmHost<float>("CTRSM", side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb);
}

HWYBLAS_AMD_API void ztrsm_(const char* side, const char* uplo, const char* transa,
                            const char* diag, const blasint* m, const blasint* n,
                            const double* alpha, const double* a, const blasint* lda, double* b,
                            const blasint* ldb) {
  hwy_amd::ZtrsmHost<double>("ZTRSM", side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb);
}

HWYBLAS_AMD_API int hwyblas_ctrsm_dev(char side, char uplo, char transa, char diag, int m, int n,
                                      const float* alpha, const float* a, int lda, float* b,
                                      int ldb, void* stream) {
  return hwy_amd::ZtrsmDev<float>("CTRSM", side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb,
                                  static_cast<hipStream_t>(stream));
}

HWYBLAS_AMD_API int hwyblas_ztrsm_dev(char side, char uplo, char transa, char diag, int m, int n,
                                      const double* alpha, const double* a, int lda, double* b,
                                      int ldb, void* stream) {
  return hwy_amd::ZtrsmDev<double>("ZTRSM", side, uplo, transa, diag, m, n, alpha, a, lda, b,
                                   ldb, static_cast<hipStream_t>(stream));
}

HWYBLAS_AMD_API void csyrk_(const char* uplo, const char* trans, const blasint* n,
                            const blasint* k, const float* alpha, const float* a,
                            const blasint* lda, const float* beta, float* c, const blasint* ldc) {
  hwy_amd::ZSyrkHost<float>("CSYRK", false, uplo, trans, n, k, alpha, a, lda, beta, c, ldc);
}

HWYBLAS_AMD_API void zsyrk_(const char* uplo, const char* trans, const blasint* n,
                            const blasint* k, const double* alpha, const double* a,
                            const blasint* lda, const double* beta, double* c,
                            const blasint* ldc) {
  hwy_amd::ZSyrkHost<double>("ZSYRK", false, uplo, trans, n, k, alpha, a, lda, beta, c, ldc);
}

HWYBLAS_AMD_API void cherk_(const char* uplo, const char* trans, const blasint* n,
                            const blasint* k, const float* alpha, const float* a,
                            const blasint* lda, const float* beta, float* c, const blasint* ldc) {
  const float al[2] = {*alpha, 0.0f}, be[2] = {*beta, 0.0f};
  hwy_amd::ZSyrkHost<float>("CHERK", true, uplo, trans, n, k, al, a, lda, be, c, ldc);
}

HWYBLAS_AMD_API void zherk_(const char* uplo, const char* trans, const blasint* n,
                            const blasint* k, const double* alpha, const double* a,
                            const blasint* lda, const double* beta, double* c,
                            const blasint* ldc) {
  const double al[2] = {*alpha, 0.0}, be[2] = {*beta, 0.0};
  hwy_amd::ZSyrkHost<double>("ZHERK", true, uplo, trans, n, k, al, a, lda, be, c, ldc);
}

}  // extern "C"
