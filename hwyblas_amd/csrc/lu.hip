// LU factorization with partial pivoting, the solve with it, and FP64
// iterative refinement: the getrf -> getrs -> gerfs consumer of GEMM and
// TRSM (hwyBLAS's LAPACK port, src/lapack/getrf.cc and gesvx.cc:333-463),
// on device memory.
//
// getrf follows xGETRF2's recursion (src/lapack/getrf.cc:112-207): factor
// the left half of the columns, swap its pivots into the right half, solve
// A12 := L11^-1 A12 (TRSM, FP64), update A22 -= A21 A12 (GEMM), factor A22,
// swap its pivots into the left half. The recursion stops at panels of at
// most kPanel columns, factored column by column as xGETF2 / hwyBLAS's leaf
// kernel does (src/lapack/factor_kernels.cc:107-138): pivot = first entry of
// largest magnitude (xIAMAX), row swap, scaling by the reciprocal of the
// pivot when |pivot| >= sfmin (else division), rank-1 update. On the GPU a
// column step is one launch: every workgroup eliminates the current column
// from its rows and finds the largest entry of the next column among them;
// the last workgroup to finish (an atomic counter) reduces the candidates,
// records the pivot and swaps the two rows of the panel. Halving the
// columns keeps the GEMMs large, so the trailing updates can run on:
//  - update 0: the FP64 MFMA GEMM;
//  - update 1: the emulated FP64 GEMM (INT8 Ozaki scheme) when k is large
//    enough for it to pay, else FP64;
//  - update 2: FP32 (operands rounded to FP32, product accumulated in FP32,
//    subtracted in FP64).
// Panels, pivots, TRSMs and the refinement stay FP64.
//
// The refinement is gerfs's loop (gesvx.cc:395-463), all right-hand sides
// at once: R = B - A X and W = |A| |X| + |B| in one pass over A,
// berr_j = max_i |R_ij| / W_ij (with gerfs's safe1/safe2 guard), and a
// column keeps iterating (X_j += getrs(R_j)) while berr_j > eps, berr_j at
// most half its previous value, and at most 5 corrections.
#include <algorithm>
#include <climits>
#include <cmath>
#include <limits>
#include <vector>

#include "common.h"

namespace hwy_amd {
namespace ozaki {
int EmulatedDgemm(bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha,
                  const double* a, int64_t lda, const double* b, int64_t ldb, double beta,
                  double* c, int64_t ldc, int nmod, hipStream_t stream);
}  // namespace ozaki

namespace lu {
namespace {

constexpr int kPanel = 64;          // columns factored column by column
constexpr int kStepRows = 256;      // rows per workgroup in a column step
constexpr int64_t kEmulatedMinK = 1024;  // below this the emulation does not pay

struct Candidate {
  double v;  // |entry|
  int row;   // row within the panel
};

// IAMAX order: larger magnitude first, then the smaller row.
__device__ __forceinline__ bool Better(double v, int row, double bv, int brow) {
  return v > bv || (v == bv && row < brow);
}

// One column step of the panel factorization (panel p: m x nb, rows
// relative to the panel's first row, which is row `row0` of the matrix).
// For j >= 0: eliminates column j from rows j + 1 .. m - 1 (the pivot row j
// is already in place). For j + 1 < nb: finds the pivot of column j + 1;
// the last workgroup records it and swaps rows of the panel.
__global__ void __launch_bounds__(kStepRows)
    PanelStepKernel(double* __restrict__ p, int64_t lda, int m, int nb, int j, int row0,
                    int* __restrict__ ipiv, Candidate* __restrict__ cand,
                    unsigned* __restrict__ counter, int* __restrict__ info, double sfmin) {
  const int r = j + 1 + blockIdx.x * kStepRows + threadIdx.x;
  const bool mine = r < m;
  double next = 0.0;  // |p(r, j + 1)| after the elimination
  if (mine) {
    if (j >= 0) {
      const double piv = p[j + int64_t(j) * lda];
      double l = p[r + int64_t(j) * lda];
      if (piv != 0.0) {
        l = (fabs(piv) >= sfmin) ? l * (1.0 / piv) : l / piv;
        p[r + int64_t(j) * lda] = l;
      }
      for (int c = j + 1; c < nb; ++c) {
        double* x = p + r + int64_t(c) * lda;
        *x = __builtin_fma(-p[j + int64_t(c) * lda], l, *x);
      }
    }
    if (j + 1 < nb) next = fabs(p[r + int64_t(j + 1) * lda]);
  }
  // Pivots exist for the first min(m, nb) columns only.
  if (j + 1 >= min(m, nb)) return;

  // Block candidate (wave shuffles, then across the waves through LDS).
  // NaNs never win (xIAMAX skips them), except in the first row (below).
  double bv = (mine && !isnan(next)) ? next : -1.0;
  int brow = mine ? r : INT_MAX;
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_down(bv, off, 64);
    const int orow = __shfl_down(brow, off, 64);
    if (Better(ov, orow, bv, brow)) {
      bv = ov;
      brow = orow;
    }
  }
  __shared__ double sv[kStepRows / 64];
  __shared__ int srow[kStepRows / 64];
  __shared__ bool last;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    sv[wave] = bv;
    srow[wave] = brow;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < kStepRows / 64; ++w)
      if (Better(sv[w], srow[w], bv, brow)) {
        bv = sv[w];
        brow = srow[w];
      }
    cand[blockIdx.x] = {bv, brow};
    __threadfence();
    last = atomicAdd(counter, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;

  // The last workgroup: the pivot of column j + 1 and the row swap.
  __shared__ int piv_row;
  if (threadIdx.x == 0) {
    __threadfence();
    double gv = -1.0;
    int grow = INT_MAX;
    for (unsigned b = 0; b < gridDim.x; ++b) {
      const Candidate c = cand[b];
      if (Better(c.v, c.row, gv, grow)) {
        gv = c.v;
        grow = c.row;
      }
    }
    // xIAMAX picks the first row when it holds a NaN or the column is zero.
    if (isnan(p[(j + 1) + int64_t(j + 1) * lda]) || !(gv > 0.0)) {
      grow = j + 1;
      if (gv == 0.0) atomicMin(info, row0 + j + 2);
    }
    ipiv[j + 1] = grow + 1;
    piv_row = grow;
    *counter = 0;
  }
  __syncthreads();
  const int pr = piv_row;
  if (pr != j + 1) {
    for (int c = threadIdx.x; c < nb; c += kStepRows) {
      double* x = p + (j + 1) + int64_t(c) * lda;
      double* y = p + pr + int64_t(c) * lda;
      const double t = *x;
      *x = *y;
      *y = t;
    }
  }
}

// Row interchanges (xLASWP, src/lapack/getrf.cc:28-81) on ncols columns:
// rows k1 .. k2 - 1 in order (forward) or in reverse, row i with row
// ipiv[i] - 1 + shift. One thread per column; the swaps of a column are
// sequential.
__global__ void LaswpKernel(double* __restrict__ a, int64_t lda, int ncols, int k1, int k2,
                            const int* __restrict__ ipiv, int shift, int forward) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncols) return;
  double* col = a + int64_t(c) * lda;
  if (forward) {
    for (int i = k1; i < k2; ++i) {
      const int q = ipiv[i] - 1 + shift;
      if (q != i) {
        const double t = col[i];
        col[i] = col[q];
        col[q] = t;
      }
    }
  } else {
    for (int i = k2 - 1; i >= k1; --i) {
      const int q = ipiv[i] - 1 + shift;
      if (q != i) {
        const double t = col[i];
        col[i] = col[q];
        col[q] = t;
      }
    }
  }
}

__global__ void AddToPivots(int* __restrict__ ipiv, int count, int add) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) ipiv[i] += add;
}

__global__ void FillInt(int* __restrict__ x, int count, int v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) x[i] = v;
}

__global__ void FillUnsigned(unsigned* __restrict__ x, int count, unsigned v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) x[i] = v;
}

// dst (float, rows x cols, ld = rows) = src (double).
__global__ void ToFloat(const double* __restrict__ src, int64_t lds, int64_t rows, int64_t cols,
                        float* __restrict__ dst) {
  const int64_t total = rows * cols;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = t % rows, j = t / rows;
    dst[t] = float(src[i + j * lds]);
  }
}

// c (double) -= p (float, ld = rows).
__global__ void SubtractFloat(double* __restrict__ c, int64_t ldc, int64_t rows, int64_t cols,
                              const float* __restrict__ p) {
  const int64_t total = rows * cols;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = t % rows, j = t / rows;
    c[i + j * ldc] -= double(p[t]);
  }
}

unsigned GridFor(int64_t n, int threads = 256) {
  return unsigned(std::max<int64_t>(1, std::min<int64_t>((n + threads - 1) / threads, 1 << 16)));
}

struct Work {
  Candidate* cand = nullptr;
  unsigned* counter = nullptr;
  int* info = nullptr;
  float* f32 = nullptr;
  size_t f32_elems = 0;
};

void Laswp(double* a, int64_t lda, int ncols, int k1, int k2, const int* ipiv, int shift,
           bool forward, hipStream_t s) {
  if (ncols <= 0 || k2 <= k1) return;
  hipLaunchKernelGGL(LaswpKernel, dim3(GridFor(ncols, 64)), dim3(64), 0, s, a, lda, ncols, k1, k2,
                     ipiv, shift, forward ? 1 : 0);
}

void Panel(double* a, int64_t lda, int m, int nb, int row0, int* ipiv, Work& w, hipStream_t s) {
  const double sfmin = std::numeric_limits<double>::min();
  const int mn = std::min(m, nb);
  // Step j eliminates column j (j >= 0) and pivots column j + 1 (if any).
  for (int j = -1; j < mn; ++j) {
    const int rows = m - (j + 1);
    if (j + 1 >= mn && rows <= 0) break;
    const unsigned blocks = unsigned(std::max(1, (rows + kStepRows - 1) / kStepRows));
    hipLaunchKernelGGL(PanelStepKernel, dim3(blocks), dim3(kStepRows), 0, s, a, lda, m, nb, j,
                       row0, ipiv, w.cand, w.counter, w.info, sfmin);
  }
  HWY_HIP_CHECK(hipGetLastError());
}

void Update(int64_t m, int64_t n, int64_t k, const double* a21, const double* a12, double* a22,
            int64_t lda, int update, Work& w, hipStream_t s) {
  if (m <= 0 || n <= 0 || k <= 0) return;
  if (update == 1 && k >= kEmulatedMinK) {
    ozaki::EmulatedDgemm(false, false, m, n, k, -1.0, a21, lda, a12, lda, 1.0, a22, lda, 0, s);
    return;
  }
  if (update == 2) {
    const size_t need = size_t(m * k + k * n + m * n);
    if (need > w.f32_elems) {
      if (w.f32) {
        HWY_HIP_CHECK(hipStreamSynchronize(s));
        HWY_HIP_CHECK(hipFree(w.f32));
      }
      HWY_HIP_CHECK(hipMalloc(&w.f32, need * sizeof(float)));
      w.f32_elems = need;
    }
    float* fa = w.f32;
    float* fb = fa + m * k;
    float* fc = fb + k * n;
    hipLaunchKernelGGL(ToFloat, dim3(GridFor(m * k)), dim3(256), 0, s, a21, lda, m, k, fa);
    hipLaunchKernelGGL(ToFloat, dim3(GridFor(k * n)), dim3(256), 0, s, a12, lda, k, n, fb);
    GemmDevice<float>(false, false, m, n, k, 1.0f, fa, m, fb, k, 0.0f, fc, m, s);
    hipLaunchKernelGGL(SubtractFloat, dim3(GridFor(m * n)), dim3(256), 0, s, a22, lda, m, n, fc);
    HWY_HIP_CHECK(hipGetLastError());
    return;
  }
  GemmDevice<double>(false, false, m, n, k, -1.0, a21, lda, a12, lda, 1.0, a22, lda, s);
}

// xGETRF2's recursion on the m x n block at matrix row `row0` (ipiv
// relative to the block's first row).
void Recurse(double* a, int64_t lda, int m, int n, int row0, int* ipiv, int update, Work& w,
             hipStream_t s) {
  const int mn = std::min(m, n);
  if (n <= kPanel || mn <= kPanel) {
    Panel(a, lda, m, n, row0, ipiv, w, s);
    return;
  }
  int n1 = std::max(kPanel, (mn / 2) / kPanel * kPanel);
  const int n2 = n - n1;
  Recurse(a, lda, m, n1, row0, ipiv, update, w, s);
  Laswp(a + int64_t(n1) * lda, lda, n2, 0, n1, ipiv, 0, true, s);
  TrsmDevice<double>(true, true, false, true, n1, n2, a, lda, a + int64_t(n1) * lda, lda, s);
  Update(m - n1, n2, n1, a + n1, a + int64_t(n1) * lda, a + n1 + int64_t(n1) * lda, lda, update,
         w, s);
  Recurse(a + n1 + int64_t(n1) * lda, lda, m - n1, n2, row0 + n1, ipiv + n1, update, w, s);
  const int cnt = mn - n1;
  hipLaunchKernelGGL(AddToPivots, dim3(GridFor(cnt)), dim3(256), 0, s, ipiv + n1, cnt, n1);
  Laswp(a, lda, n1, n1, mn, ipiv, 0, true, s);
}

}  // namespace

// Returns info (0, or the first zero pivot, 1-based). Synchronizes.
int Getrf(int m, int n, double* a, int64_t lda, int* ipiv, int update, hipStream_t s) {
  if (m == 0 || n == 0) return 0;
  Work w;
  const int max_blocks = (m + kStepRows - 1) / kStepRows + 1;
  HWY_HIP_CHECK(hipMalloc(&w.cand, size_t(max_blocks) * sizeof(Candidate)));
  HWY_HIP_CHECK(hipMalloc(&w.counter, 256));
  HWY_HIP_CHECK(hipMalloc(&w.info, 256));
  hipLaunchKernelGGL(FillUnsigned, dim3(1), dim3(64), 0, s, w.counter, 1, 0u);
  hipLaunchKernelGGL(FillInt, dim3(1), dim3(64), 0, s, w.info, 1, INT_MAX);
  Recurse(a, lda, m, n, 0, ipiv, update, w, s);
  int info = INT_MAX;
  HWY_HIP_CHECK(hipMemcpyAsync(&info, w.info, sizeof(int), hipMemcpyDeviceToHost, s));
  HWY_HIP_CHECK(hipStreamSynchronize(s));
  for (void* p : {static_cast<void*>(w.cand), static_cast<void*>(w.counter),
                  static_cast<void*>(w.info), static_cast<void*>(w.f32)})
    if (p) HWY_HIP_CHECK(hipFree(p));
  return info == INT_MAX ? 0 : info;
}

// xGETRS (src/lapack/getrf.cc:287-339) with device ipiv.
void Getrs(bool trans, int n, int nrhs, const double* a, int64_t lda, const int* ipiv, double* b,
           int64_t ldb, hipStream_t s) {
  if (n == 0 || nrhs == 0) return;
  if (!trans) {
    Laswp(b, ldb, nrhs, 0, n, ipiv, 0, true, s);
    TrsmDevice<double>(true, true, false, true, n, nrhs, a, lda, b, ldb, s);
    TrsmDevice<double>(true, false, false, false, n, nrhs, a, lda, b, ldb, s);
  } else {
    TrsmDevice<double>(true, false, true, false, n, nrhs, a, lda, b, ldb, s);
    TrsmDevice<double>(true, true, true, true, n, nrhs, a, lda, b, ldb, s);
    Laswp(b, ldb, nrhs, 0, n, ipiv, 0, false, s);
  }
}

// ---------------------------------------------------------------------------
// Refinement

constexpr int kMaxRhs = 16;
constexpr int kResChunk = 512;  // columns of A per workgroup of the residual

// Partial sums over the columns [k0, k0 + kResChunk) of A: rp = sum a x,
// wp = sum |a| |x| for every row (one thread each) and right-hand side.
__global__ void __launch_bounds__(256)
    ResidualPartialKernel(const double* __restrict__ a, int64_t lda, int n,
                          const double* __restrict__ x, int64_t ldx, int nrhs,
                          double* __restrict__ rp, double* __restrict__ wp) {
  __shared__ double xs[kResChunk][kMaxRhs];
  const int k0 = blockIdx.y * kResChunk;
  const int kn = min(kResChunk, n - k0);
  for (int t = threadIdx.x; t < kn * kMaxRhs; t += blockDim.x) {
    const int kk = t / kMaxRhs, j = t % kMaxRhs;
    xs[kk][j] = j < nrhs ? x[(k0 + kk) + int64_t(j) * ldx] : 0.0;
  }
  __syncthreads();
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double r[kMaxRhs], w[kMaxRhs];
#pragma unroll
  for (int j = 0; j < kMaxRhs; ++j) r[j] = w[j] = 0.0;
  const double* ai = a + i + int64_t(k0) * lda;
  for (int kk = 0; kk < kn; ++kk) {
    const double v = ai[int64_t(kk) * lda];
    const double av = fabs(v);
#pragma unroll
    for (int j = 0; j < kMaxRhs; ++j) {
      r[j] = __builtin_fma(v, xs[kk][j], r[j]);
      w[j] = __builtin_fma(av, fabs(xs[kk][j]), w[j]);
    }
  }
  const int64_t base = (int64_t(blockIdx.y) * kMaxRhs) * n + i;
  for (int j = 0; j < nrhs; ++j) {
    rp[base + int64_t(j) * n] = r[j];
    wp[base + int64_t(j) * n] = w[j];
  }
}

// R = B - sum rp; W = |B| + sum wp; per workgroup and right-hand side the
// largest berr term and the largest |R| (columns of `stats`: 2 * nrhs).
__global__ void __launch_bounds__(256)
    ResidualFinishKernel(const double* __restrict__ rp, const double* __restrict__ wp,
                         int nchunks, const double* __restrict__ b, int64_t ldb, int n, int nrhs,
                         double* __restrict__ r_out, int64_t ldr, double safe1, double safe2,
                         double* __restrict__ stats) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  __shared__ double red[256];
  for (int j = 0; j < nrhs; ++j) {
    double berr = 0.0, rabs = 0.0;
    if (i < n) {
      double sr = 0.0, sw = 0.0;
      for (int c = 0; c < nchunks; ++c) {
        sr += rp[(int64_t(c) * kMaxRhs + j) * n + i];
        sw += wp[(int64_t(c) * kMaxRhs + j) * n + i];
      }
      const double bij = b[i + int64_t(j) * ldb];
      const double rij = bij - sr;
      const double wij = fabs(bij) + sw;
      r_out[i + int64_t(j) * ldr] = rij;
      berr = wij > safe2 ? fabs(rij) / wij : (fabs(rij) + safe1) / (wij + safe1);
      rabs = fabs(rij);
    }
    for (int which = 0; which < 2; ++which) {
      red[threadIdx.x] = which == 0 ? berr : rabs;
      __syncthreads();
      for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + h]);
        __syncthreads();
      }
      if (threadIdx.x == 0) stats[int64_t(blockIdx.x) * 2 * nrhs + which * nrhs + j] = red[0];
      __syncthreads();
    }
  }
}

// x(:, j) += d(:, j) for the columns with active[j] != 0.
__global__ void AddMaskedKernel(double* __restrict__ x, int64_t ldx, const double* __restrict__ d,
                                int64_t ldd, int n, int nrhs, const int* __restrict__ active) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int j = 0; j < nrhs; ++j)
    if (active[j]) x[i + int64_t(j) * ldx] += d[i + int64_t(j) * ldd];
}

// Largest row sum of |A| (the infinity norm).
__global__ void RowAbsSumKernel(const double* __restrict__ a, int64_t lda, int n,
                                double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int k = 0; k < n; ++k) s += fabs(a[i + int64_t(k) * lda]);
  out[i] = s;
}

}  // namespace lu
}  // namespace hwy_amd

using namespace hwy_amd;

namespace {
hipStream_t StreamOr(void* stream) {
  return stream ? static_cast<hipStream_t>(stream) : Context().stream;
}
}  // namespace

extern "C" {

HWYBLAS_AMD_API int hwyblas_dgetrf_dev(int m, int n, double* a, int lda, int* ipiv, int update,
                                       void* stream) {
  int info = 0;
  if (m < 0) info = -1;
  else if (n < 0) info = -2;
  else if (lda < std::max(1, m)) info = -4;
  else if (update < 0 || update > 2) info = -6;
  if (info != 0) {
    ReportError("DGETRF", -info);
    return info;
  }
  if (m == 0 || n == 0) return 0;
  return lu::Getrf(m, n, a, lda, ipiv, update, StreamOr(stream));
}

HWYBLAS_AMD_API int hwyblas_dgetrs_dev(char trans, int n, int nrhs, const double* a, int lda,
                                       const int* ipiv, double* b, int ldb, void* stream) {
  Op op;
  int info = 0;
  if (!DecodeOp(trans, &op)) info = -1;
  else if (n < 0) info = -2;
  else if (nrhs < 0) info = -3;
  else if (lda < std::max(1, n)) info = -5;
  else if (ldb < std::max(1, n)) info = -8;
  if (info != 0) {
    ReportError("DGETRS", -info);
    return info;
  }
  if (n == 0 || nrhs == 0) return 0;
  lu::Getrs(op == Op::kTrans, n, nrhs, a, lda, ipiv, b, ldb, StreamOr(stream));
  return 0;
}

// Solves A X = B: af := LU of A (trailing updates in the precision
// `update`), X := getrs(B), then FP64 refinement against A. berr, nbe
// (normwise backward error ||B - A X||inf / (||A||inf ||X||inf)) and
// iterations (corrections applied) per right-hand side, on the host.
// Returns getrf's info (> 0: singular, nothing solved) or < 0 on an
// invalid argument.
HWYBLAS_AMD_API int hwyblas_dgesv_ir_dev(int n, int nrhs, const double* a, int lda, double* af,
                                         int ldaf, int* ipiv, const double* b, int ldb, double* x,
                                         int ldx, int update, int itmax, double* berr,
                                         double* nbe, int* iterations, void* stream) {
  int info = 0;
  if (n < 0) info = -1;
  else if (nrhs < 0 || nrhs > lu::kMaxRhs) info = -2;
  else if (lda < std::max(1, n)) info = -4;
  else if (ldaf < std::max(1, n)) info = -6;
  else if (ldb < std::max(1, n)) info = -9;
  else if (ldx < std::max(1, n)) info = -11;
  else if (update < 0 || update > 2) info = -12;
  else if (itmax < 0) info = -13;
  if (info != 0) {
    ReportError("DGESVIR", -info);
    return info;
  }
  if (n == 0 || nrhs == 0) return 0;
  hipStream_t s = StreamOr(stream);
  HWY_HIP_CHECK(hipMemcpy2DAsync(af, size_t(ldaf) * 8, a, size_t(lda) * 8, size_t(n) * 8, n,
                                 hipMemcpyDeviceToDevice, s));
  info = lu::Getrf(n, n, af, ldaf, ipiv, update, s);
  if (info > 0) return info;
  HWY_HIP_CHECK(hipMemcpy2DAsync(x, size_t(ldx) * 8, b, size_t(ldb) * 8, size_t(n) * 8, nrhs,
                                 hipMemcpyDeviceToDevice, s));
  lu::Getrs(false, n, nrhs, af, ldaf, ipiv, x, ldx, s);

  const int nchunks = (n + lu::kResChunk - 1) / lu::kResChunk;
  const int rblocks = (n + 255) / 256;
  double *rp, *wp, *r, *stats, *rowsum;
  int* active_d;
  HWY_HIP_CHECK(hipMalloc(&rp, size_t(nchunks) * lu::kMaxRhs * n * 8));
  HWY_HIP_CHECK(hipMalloc(&wp, size_t(nchunks) * lu::kMaxRhs * n * 8));
  HWY_HIP_CHECK(hipMalloc(&r, size_t(n) * nrhs * 8));
  HWY_HIP_CHECK(hipMalloc(&stats, size_t(rblocks) * 2 * nrhs * 8));
  HWY_HIP_CHECK(hipMalloc(&rowsum, size_t(n) * 8));
  HWY_HIP_CHECK(hipMalloc(&active_d, lu::kMaxRhs * sizeof(int)));
  const double eps = std::numeric_limits<double>::epsilon() * 0.5;  // LAPACK's dlamch('E')
  const double safmin = std::numeric_limits<double>::min();
  const double safe1 = double(n + 1) * safmin, safe2 = safe1 / eps;
  std::vector<double> host_stats(size_t(rblocks) * 2 * nrhs);
  std::vector<double> lstres(nrhs, 3.0), rinf(nrhs, 0.0);
  std::vector<int> count(nrhs, 1), active(nrhs, 1), done(nrhs, 0);
  for (;;) {
    hipLaunchKernelGGL(lu::ResidualPartialKernel, dim3(rblocks, nchunks), dim3(256), 0, s, a,
                       int64_t(lda), n, x, int64_t(ldx), nrhs, rp, wp);
    hipLaunchKernelGGL(lu::ResidualFinishKernel, dim3(rblocks), dim3(256), 0, s, rp, wp, nchunks,
                       b, int64_t(ldb), n, nrhs, r, int64_t(n), safe1, safe2, stats);
    HWY_HIP_CHECK(hipGetLastError());
    HWY_HIP_CHECK(hipMemcpyAsync(host_stats.data(), stats, host_stats.size() * 8,
                                 hipMemcpyDeviceToHost, s));
    HWY_HIP_CHECK(hipStreamSynchronize(s));
    bool any = false;
    for (int j = 0; j < nrhs; ++j) {
      double be = 0.0, ri = 0.0;
      for (int blk = 0; blk < rblocks; ++blk) {
        be = std::max(be, host_stats[size_t(blk) * 2 * nrhs + j]);
        ri = std::max(ri, host_stats[size_t(blk) * 2 * nrhs + nrhs + j]);
      }
      active[j] = 0;
      if (done[j]) continue;
      berr[j] = be;
      rinf[j] = ri;
      if (be > eps && 2.0 * be <= lstres[j] && count[j] <= itmax) {
        active[j] = 1;
        lstres[j] = be;
        ++count[j];
        any = true;
      } else {
        done[j] = 1;
      }
    }
    if (!any) break;
    HWY_HIP_CHECK(hipMemcpyAsync(active_d, active.data(), nrhs * sizeof(int),
                                 hipMemcpyHostToDevice, s));
    lu::Getrs(false, n, nrhs, af, ldaf, ipiv, r, n, s);
    hipLaunchKernelGGL(lu::AddMaskedKernel, dim3(rblocks), dim3(256), 0, s, x, int64_t(ldx), r,
                       int64_t(n), n, nrhs, active_d);
    HWY_HIP_CHECK(hipGetLastError());
  }
  // Normwise backward error of the final X.
  hipLaunchKernelGGL(lu::RowAbsSumKernel, dim3(rblocks), dim3(256), 0, s, a, int64_t(lda), n,
                     rowsum);
  std::vector<double> rows(n), xs(size_t(n) * nrhs);
  HWY_HIP_CHECK(hipMemcpyAsync(rows.data(), rowsum, size_t(n) * 8, hipMemcpyDeviceToHost, s));
  HWY_HIP_CHECK(hipMemcpy2DAsync(xs.data(), size_t(n) * 8, x, size_t(ldx) * 8, size_t(n) * 8,
                                 nrhs, hipMemcpyDeviceToHost, s));
  HWY_HIP_CHECK(hipStreamSynchronize(s));
  const double anorm = *std::max_element(rows.begin(), rows.end());
  for (int j = 0; j < nrhs; ++j) {
    double xn = 0.0;
    for (int i = 0; i < n; ++i) xn = std::max(xn, std::fabs(xs[size_t(j) * n + i]));
    if (nbe) nbe[j] = (anorm > 0 && xn > 0) ? rinf[j] / (anorm * xn) : 0.0;
    if (iterations) iterations[j] = count[j] - 1;
  }
  for (void* p : {static_cast<void*>(rp), static_cast<void*>(wp), static_cast<void*>(r),
                  static_cast<void*>(stats), static_cast<void*>(rowsum),
                  static_cast<void*>(active_d)})
    HWY_HIP_CHECK(hipFree(p));
  return 0;
}

}  // extern "C"
