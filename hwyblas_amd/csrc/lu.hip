// LU factorization with partial pivoting, the solve with it, and FP64
// iterative refinement: the getrf -> getrs -> gerfs consumer of GEMM and
// TRSM (hwyBLAS's LAPACK port, src/lapack/getrf.cc and gesvx.cc:333-463),
// on device memory.
//
// getrf follows xGETRF2's recursion (src/lapack/getrf.cc:112-207): factor
// the left half of the columns, swap its pivots into the right half, solve
// A12 := L11^-1 A12 (TRSM, FP64, its 64 x 64 diagonal blocks applied through
// their inverses), update A22 -= A21 A12 (GEMM), factor A22, swap its pivots
// into the left half. The recursion stops at panels of at
// most kPanel columns, factored column by column as xGETF2 / hwyBLAS's leaf
// kernel does (src/lapack/factor_kernels.cc:107-138): pivot = first entry of
// largest magnitude (xIAMAX), row swap, scaling by the reciprocal of the
// pivot when |pivot| >= sfmin (else division), rank-1 update. On the GPU a
// column step is one launch: every workgroup eliminates the current column
// from its rows and finds the largest entry of the next column among them;
// the last workgroup to finish (an atomic counter) reduces the candidates,
// records the pivot and swaps the two rows across the whole matrix (xLASWP's
// interchanges, applied as they are found). Halving the columns keeps the
// GEMMs large, so the trailing updates can run on:
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
#include <cstdio>
#include <cstdlib>
#include <limits>
#include <string>
#include <vector>

#include "common.h"
#include "gemm_kernel.h"
#include "grid_sync.h"

namespace hwy_amd {
namespace lu {
namespace {

constexpr int kPanel = 64;          // columns factored column by column
constexpr int kGetrfInternalError = -1000;  // a panel barrier timed out
constexpr int kStepRows = 256;      // threads per workgroup in a column step
constexpr int64_t kEmulatedMinK = 1024;  // below this the emulation does not pay

struct Candidate {
  double v;  // |entry|
  int row;   // row within the panel
};

// IAMAX order: larger magnitude first, then the smaller row.
__device__ __forceinline__ bool Better(double v, int row, double bv, int brow) {
  return v > bv || (v == bv && row < brow);
}

constexpr int kStepGroups = 4;                  // column groups per row
constexpr int kGroupCols = kPanel / kStepGroups;  // 16 columns per thread
constexpr int kBlockRows = kStepRows / kStepGroups;

__device__ __forceinline__ void BlockBest(double& bv, int& brow, double* sv, int* srow) {
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_down(bv, off, 64);
    const int orow = __shfl_down(brow, off, 64);
    if (Better(ov, orow, bv, brow)) {
      bv = ov;
      brow = orow;
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (lane == 0) {
    sv[wave] = bv;
    srow[wave] = brow;
  }
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < kStepRows / 64; ++w)
      if (Better(sv[w], srow[w], bv, brow)) {
        bv = sv[w];
        brow = srow[w];
      }
}

// One column step of the panel factorization. The panel p is m x nb (rows
// relative to its first row, which is row `row0` of the matrix; `full`
// points at that row in column 0 of the matrix, ncols wide). A workgroup
// owns 64 rows; its four waves split the panel's columns (16 each).
//  - j >= 0: eliminates column j from rows j + 1 .. m - 1 (the pivot row j is
//    already in place): l = p(r, j) / pivot, p(r, c) -= p(j, c) l.
//  - j + 1 < min(m, nb): finds the pivot of column j + 1. The last workgroup
//    to finish (atomic counter) reduces the candidates, records the pivot,
//    and swaps the two rows of the panel. PanelSwapKernel then applies the
//    panel's interchanges to every other column of the matrix (xGETRF2
//    applies them to columns that do not change in between, so applying
//    them once the panel is done gives the same matrix).
__global__ void __launch_bounds__(kStepRows)
    PanelStepKernel(double* __restrict__ p, int64_t lda, int m, int nb, int j,
                    double* __restrict__ full, int ncols, int row0, int* __restrict__ ipiv,
                    Candidate* __restrict__ cand, unsigned* __restrict__ counter,
                    int* __restrict__ info, double sfmin) {
  const int rl = threadIdx.x % kBlockRows, cg = threadIdx.x / kBlockRows;
  const int r = j + 1 + blockIdx.x * kBlockRows + rl;
  const bool mine = r < m;
  const int c0 = cg * kGroupCols;
  double next = -1.0;  // |p(r, j + 1)| after the elimination, if this thread holds it
  if (j >= 0) {
    // Every load before any store (the multiplier column is read by the
    // four column groups and written by the first).
    const double piv = p[j + int64_t(j) * lda];
    const double lraw = mine ? p[r + int64_t(j) * lda] : 0.0;
    double u[kGroupCols], v[kGroupCols];
#pragma unroll
    for (int t = 0; t < kGroupCols; ++t) {
      const int c = c0 + t;
      const bool use = c > j && c < nb;
      u[t] = use ? p[j + int64_t(c) * lda] : 0.0;
      v[t] = (use && mine) ? p[r + int64_t(c) * lda] : 0.0;
    }
    __syncthreads();
    double l = lraw;
    if (piv != 0.0) l = (fabs(piv) >= sfmin) ? lraw * (1.0 / piv) : lraw / piv;
    if (mine) {
      if (cg == 0 && piv != 0.0) p[r + int64_t(j) * lda] = l;
#pragma unroll
      for (int t = 0; t < kGroupCols; ++t) {
        const int c = c0 + t;
        if (c > j && c < nb) {
          v[t] = __builtin_fma(-u[t], l, v[t]);
          p[r + int64_t(c) * lda] = v[t];
          if (c == j + 1) next = isnan(v[t]) ? -1.0 : fabs(v[t]);
        }
      }
    }
  } else if (mine && j + 1 >= c0 && j + 1 < c0 + kGroupCols) {
    const double v0 = p[r + int64_t(j + 1) * lda];
    next = isnan(v0) ? -1.0 : fabs(v0);
  }
  // Pivots exist for the first min(m, nb) columns only.
  if (j + 1 >= min(m, nb)) return;

  __shared__ double sv[kStepRows / 64];
  __shared__ int srow[kStepRows / 64];
  __shared__ bool last;
  double bv = next;
  int brow = next >= 0.0 ? r : INT_MAX;
  BlockBest(bv, brow, sv, srow);
  if (threadIdx.x == 0) {
    cand[blockIdx.x] = {bv, brow};
    __threadfence();
    last = atomicAdd(counter, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;

  // The last workgroup: the best candidate, the pivot and the row swap.
  __threadfence();
  double gv = -1.0;
  int grow = INT_MAX;
  for (unsigned b = threadIdx.x; b < gridDim.x; b += kStepRows) {
    const Candidate c = cand[b];
    if (Better(c.v, c.row, gv, grow)) {
      gv = c.v;
      grow = c.row;
    }
  }
  __syncthreads();
  BlockBest(gv, grow, sv, srow);
  __shared__ int piv_row;
  if (threadIdx.x == 0) {
    // xIAMAX picks the first row when it holds a NaN or the column is zero.
    if (isnan(p[(j + 1) + int64_t(j + 1) * lda]) || !(gv > 0.0)) {
      grow = j + 1;
      if (gv == 0.0) atomicMin(info, row0 + j + 2);
    }
    ipiv[j + 1] = row0 + grow + 1;  // 1-based row of the whole matrix
    piv_row = grow;
    *counter = 0;
  }
  __syncthreads();
  const int pr = piv_row;
  if (pr != j + 1 && threadIdx.x < nb) {
    double* x = p + (j + 1) + int64_t(threadIdx.x) * lda;
    double* y = p + pr + int64_t(threadIdx.x) * lda;
    const double t = *x;
    *x = *y;
    *y = t;
  }
}

constexpr int kMaxAffected = 2 * kPanel;

// Applies the row interchanges of one panel (count <= kPanel pivots, rows
// relative to the panel's first row) to every column of the matrix outside
// the panel's columns [pc0, pc0 + nb). The interchanges are composed into
// "row L[x] receives old row S[x]" over the affected rows (the panel's top
// rows and their partners); each thread then moves its column's affected
// entries through LDS, all loads before all stores.
__global__ void __launch_bounds__(64)
    PanelSwapKernel(double* __restrict__ full, int64_t lda, int ncols, int pc0, int nb,
                    const int* __restrict__ ipiv, int count, int row0, int rows) {
  __shared__ int L[kMaxAffected], S[kMaxAffected];
  __shared__ int naff;
  __shared__ double buf[kMaxAffected][64 + 1];
  const int lane = threadIdx.x;
  if (lane < count) {
    L[lane] = lane;
    S[lane] = lane;
  }
  // The whole wave composes the interchanges: lane t holds the t-th partner
  // row below the panel's top rows (a ballot finds a partner already seen).
  int extra_row = -1, extra_src = -1, na = count;
  __syncthreads();
  for (int i = 0; i < count; ++i) {
    // Relative to the panel's first row; a pivot outside the panel's rows
    // (not written by a panel kernel that gave up) is no interchange.
    int q = ipiv[i] - 1 - row0;
    if (q < i || q >= rows) q = i;
    if (q < count) {
      if (lane == 0) {
        const int tmp = S[i];
        S[i] = S[q];
        S[q] = tmp;
      }
    } else {
      const unsigned long long hit = __ballot(extra_row == q);
      const int t = hit ? __ffsll(static_cast<long long>(hit)) - 1 : na - count;
      if (!hit) {
        if (lane == t) {
          extra_row = q;
          extra_src = q;
        }
        ++na;
      }
      // swap(S[i], source of row q)
      const int si = __shfl(lane == 0 ? S[i] : 0, 0, 64);
      const int sq = __shfl(extra_src, t, 64);
      if (lane == t) extra_src = si;
      if (lane == 0) S[i] = sq;
    }
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    __builtin_amdgcn_wave_barrier();
  }
  if (lane < na - count) {
    L[count + lane] = extra_row;
    S[count + lane] = extra_src;
  }
  if (lane == 0) naff = na;
  __syncthreads();
  int c = blockIdx.x * 64 + lane;
  if (c >= pc0) c += nb;  // skip the panel's own columns
  if (c >= ncols) return;
  const int total = naff;
  double* col = full + int64_t(c) * lda;
  // Loads in batches of 16 in flight (each lane's rows are scattered: one
  // cache line each), then the stores.
  constexpr int kBatch = 16;
  for (int x0 = 0; x0 < total; x0 += kBatch) {
    double v[kBatch];
#pragma unroll
    for (int u = 0; u < kBatch; ++u) v[u] = (x0 + u < total) ? col[S[x0 + u]] : 0.0;
#pragma unroll
    for (int u = 0; u < kBatch; ++u)
      if (x0 + u < total) buf[x0 + u][lane] = v[u];
  }
  for (int x = 0; x < total; ++x) col[L[x]] = buf[x][lane];
}

// The row interchanges of xLASWP (src/lapack/getrf.cc:28-81) over rows
// [0, n), composed into one permutation: row i of the result is row src[i]
// of the input (forward order; the reverse order is the inverse). The
// composition is cheap on the host; moving the rows is a parallel gather
// through a scratch copy.
__global__ void GatherRows(const double* __restrict__ b, int64_t ldb, int n, int ncols,
                           const int* __restrict__ src, double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int r = src[i];
  for (int c = 0; c < ncols; ++c) out[i + int64_t(c) * n] = b[r + int64_t(c) * ldb];
}

__global__ void CopyBack(const double* __restrict__ in, int n, int ncols, double* __restrict__ b,
                         int64_t ldb) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  for (int c = 0; c < ncols; ++c) b[i + int64_t(c) * ldb] = in[i + int64_t(c) * n];
}


// ipiv[j] = j + 1: no interchange. Written before the factorization, so
// that a panel whose kernel gave up (a timed-out barrier) leaves in-range
// pivots for the row interchanges that follow it.
__global__ void IdentityPivotsKernel(int* __restrict__ ipiv, int count) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < count) ipiv[i] = i + 1;
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

// inv (64 x 64, dense, zeros above the diagonal) = the inverse of the r x r
// unit lower triangle of a (r <= 64). Thread j computes column j by forward
// substitution on e_j.
__global__ void __launch_bounds__(64)
    InvertUnitLowerKernel(const double* __restrict__ a, int64_t lda, int r,
                          double* __restrict__ inv) {
  __shared__ double l[kPanel][kPanel + 1];
  const int j = threadIdx.x;
  {
    // Column j of the triangle, every load in flight before the first store.
    double colv[kPanel];
#pragma unroll
    for (int i = 0; i < kPanel; ++i) colv[i] = (j < i && i < r) ? a[i + int64_t(j) * lda] : 0.0;
#pragma unroll
    for (int i = 0; i < kPanel; ++i) l[i][j] = colv[i];
  }
  __syncthreads();
  double x[kPanel];
#pragma unroll
  for (int i = 0; i < kPanel; ++i) {
    double v = (i == j) ? 1.0 : 0.0;
#pragma unroll
    for (int q = 0; q < i; ++q) v = __builtin_fma(-l[i][q], x[q], v);
    x[i] = v;
  }
#pragma unroll
  for (int i = 0; i < kPanel; ++i) inv[i + j * kPanel] = (i < r && j < r) ? x[i] : 0.0;
}

// The whole panel factorization in one cooperative launch: the panel stays
// in registers and every column step ends in one grid-wide barrier instead
// of a kernel boundary. Workgroup g (512 threads) owns rows
// [256 g, 256 g + 256): wave w holds the 16 columns 16 (w % 4) .. + 15 of
// the rows 128 (w / 4) + 64 q + lane, q < 2. Per column j, as
// PanelStepKernel: the waves holding column j form the multipliers, every
// wave updates its columns; the waves holding column j + 1 find the
// workgroup's best candidate (xIAMAX order), and the workgroup publishes a
// record {|value|, row, the row's 64 values}; the owner of row j + 1
// publishes that row too. After the barrier every workgroup loads all
// records in one pass (one memory round trip), reduces the candidates itself
// (the same result everywhere), takes the pivot row's values and swaps the
// two rows in its registers. Records alternate between two slots: a slot is
// rewritten two barriers after it was read. Records, the arrival counter and
// the polls are agent-scope atomic accesses (coherent across the XCDs'
// L2 caches without cache-wide write-back or invalidation fences, which
// would dominate the step); there are few, large workgroups because the
// barrier's cost grows with their number. Each wait is bounded (kSpinLimit
// polls); a timed-out wait sets
// *failed and every workgroup returns (the launch is cooperative, so all
// workgroups are resident and the bound is never reached in a working run).
constexpr int kPersistThreads = 512;
constexpr int kPersistRows = 256;               // rows per workgroup
constexpr int kPersistRpt = kPersistRows / 128;  // rows per thread (two row halves)
constexpr int kPersistMaxGroups = 128;          // records loaded through LDS
constexpr int kRecord = kPanel + 2;             // |value|, row, 64 values

struct PersistArgs {
  double* p;
  int64_t lda;
  int m, nb, row0;
  int* ipiv;
  int* info;
  double sfmin;
  double* rec;      // [2][gridDim.x + 1][kRecord]; slot gridDim.x: row j + 1
  unsigned* bar;    // arrivals (bar_base at launch: the count runs on over panels)
  unsigned* failed;
  double* linv;     // the inverse of the panel's unit lower triangle (kPanel x kPanel)
  unsigned bar_base;
};

// Wavefront reductions of the device library (ockl).
extern "C" __device__ unsigned long long __ockl_wfred_max_u64(unsigned long long);
extern "C" __device__ int __ockl_wfred_min_i32(int);

// The wave's best (value, row) in xIAMAX order: values are non-negative
// magnitudes or -1 (no candidate / NaN); ties go to the smaller row. The
// magnitude's bits order like its value, so the largest key
// (bits + 1, 0 for -1) picks the value and a minimum over the lanes holding
// it picks the row.
__device__ __forceinline__ void WaveBest(double& v, int& row) {
  const unsigned long long key =
      v < 0.0 ? 0ull : static_cast<unsigned long long>(__double_as_longlong(v)) + 1ull;
  const unsigned long long kmax = __ockl_wfred_max_u64(key);
  row = __ockl_wfred_min_i32(key == kmax ? row : INT_MAX);
  v = kmax == 0ull ? -1.0 : __longlong_as_double(static_cast<long long>(kmax - 1ull));
}

// x[t] for a run-time t < kGroupCols without indexing registers.
__device__ __forceinline__ double Pick(const double (&x)[kGroupCols], int t) {
  double r = 0.0;
#pragma unroll
  for (int i = 0; i < kGroupCols; ++i) r = (i == t) ? x[i] : r;
  return r;
}

// x[t] = y where `put`, for a run-time t < kGroupCols.
__device__ __forceinline__ void PutIf(double (&x)[kGroupCols], int t, bool put, double y) {
#pragma unroll
  for (int i = 0; i < kGroupCols; ++i) x[i] = (put && i == t) ? y : x[i];
}

__global__ void __launch_bounds__(kPersistThreads) PanelPersistentKernel(PersistArgs g) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int cg = wave % kStepGroups, half = wave / kStepGroups;
  const int c0 = cg * kGroupCols;
  const int G = gridDim.x, gid = blockIdx.x;
  const int m = g.m, nb = g.nb, mn = min(m, nb);
  const int64_t lda = g.lda;
  __shared__ double sl[kPersistRows];             // multipliers of the current column
  __shared__ double srec[kPersistMaxGroups + 1][kRecord];
  __shared__ double spub[2][kPanel];             // the candidate row, row jn
  __shared__ double hv[2];
  __shared__ int hrow[2];

  int rows[kPersistRpt];
  double v[kPersistRpt][kGroupCols];
#pragma unroll
  for (int q = 0; q < kPersistRpt; ++q) {
    rows[q] = gid * kPersistRows + half * (kPersistRows / 2) + q * 64 + lane;
#pragma unroll
    for (int t = 0; t < kGroupCols; ++t) {
      const int c = c0 + t;
      v[q][t] = (rows[q] < m && c < nb) ? g.p[rows[q] + int64_t(c) * lda] : 0.0;
    }
  }
  const int lbase = half * (kPersistRows / 2);  // this wave's rows in sl
  // The pivot row of the last step: record srec[pivot_rec] (its values from
  // srec[pivot_rec][2]); it stays in LDS until the next step's records land.
  int pivot_rec = 0;

  for (int j = -1; j < mn; ++j) {
    if (j >= 0) {
      // Eliminate column j with the pivot row (row j, in place).
      const double* su = &srec[pivot_rec][2];
      const double piv = su[j];
      if (cg == j / kGroupCols) {
#pragma unroll
        for (int q = 0; q < kPersistRpt; ++q) {
          const double lraw = Pick(v[q], j % kGroupCols);
          double l = lraw;
          if (piv != 0.0) l = (fabs(piv) >= g.sfmin) ? lraw * (1.0 / piv) : lraw / piv;
          const bool below = rows[q] > j && rows[q] < m;
          PutIf(v[q], j % kGroupCols, below && piv != 0.0, l);
          sl[lbase + q * 64 + lane] = below ? l : 0.0;
        }
      }
      __syncthreads();
      double u[kGroupCols];
#pragma unroll
      for (int t = 0; t < kGroupCols; ++t) u[t] = su[c0 + t];
#pragma unroll
      for (int q = 0; q < kPersistRpt; ++q) {
        if (rows[q] > j && rows[q] < m) {
          const double l = sl[lbase + q * 64 + lane];
#pragma unroll
          for (int t = 0; t < kGroupCols; ++t) {
            const int c = c0 + t;
            if (c > j && c < nb) v[q][t] = __builtin_fma(-u[t], l, v[q][t]);
          }
        }
      }
    }
    const int jn = j + 1;
    if (jn >= mn) break;
    const int buf = jn & 1;
    double* slot = g.rec + int64_t(buf) * (G + 1) * kRecord;
    // This workgroup's candidate for column jn: each of the two waves holding
    // the column reduces its rows, then the better of the two.
    if (cg == jn / kGroupCols) {
      double bv = -1.0;
      int brow = INT_MAX;
#pragma unroll
      for (int q = 0; q < kPersistRpt; ++q) {
        if (rows[q] >= jn && rows[q] < m) {
          const double x = Pick(v[q], jn % kGroupCols);
          const double val = isnan(x) ? -1.0 : fabs(x);
          if (Better(val, rows[q], bv, brow)) {
            bv = val;
            brow = rows[q];
          }
        }
      }
      WaveBest(bv, brow);
      if (lane == 0) {
        hv[half] = bv;
        hrow[half] = brow;
      }
    }
    __syncthreads();
    double bv = hv[0];
    int br = hrow[0];
    if (Better(hv[1], hrow[1], bv, br)) {
      bv = hv[1];
      br = hrow[1];
    }
    // The candidate row and row jn go through LDS, so that every thread
    // publishes at most one value (the compiler waits for earlier
    // agent-scope stores before each group of them).
#pragma unroll
    for (int q = 0; q < kPersistRpt; ++q) {
      if (rows[q] == br) {
#pragma unroll
        for (int t = 0; t < kGroupCols; ++t) spub[0][c0 + t] = v[q][t];
      }
      if (rows[q] == jn) {
#pragma unroll
        for (int t = 0; t < kGroupCols; ++t) spub[1][c0 + t] = v[q][t];
      }
    }
    __syncthreads();
    {
      const int t = threadIdx.x;
      const bool own_jn = jn >= gid * kPersistRows && jn < (gid + 1) * kPersistRows && jn < m;
      if (t < kPanel) {
        if (br != INT_MAX) PutRecord(slot + int64_t(gid) * kRecord + 2 + t, spub[0][t]);
      } else if (t < 2 * kPanel) {
        if (own_jn) PutRecord(slot + int64_t(G) * kRecord + 2 + (t - kPanel), spub[1][t - kPanel]);
      } else if (t == 2 * kPanel) {
        PutRecord(slot + int64_t(gid) * kRecord, bv);
      } else if (t == 2 * kPanel + 1) {
        PutRecord(slot + int64_t(gid) * kRecord + 1, double(br));
      }
    }
    if (!GridSync(g.bar, g.bar_base + unsigned(jn + 1) * unsigned(G), g.failed)) return;

    // Every record in one pass (all of a thread's loads in flight before
    // the first LDS store; the guard is uniform), then the pivot (the same
    // in every workgroup).
    {
      constexpr int kRecLoads =
          ((kPersistMaxGroups + 1) * kRecord + kPersistThreads - 1) / kPersistThreads;
      const int nrec = (G + 1) * kRecord;
      double rv[kRecLoads];
#pragma unroll
      for (int q = 0; q < kRecLoads; ++q) {
        const int idx = threadIdx.x + q * kPersistThreads;
        if (q * kPersistThreads < nrec) rv[q] = GetRecord(slot + (idx < nrec ? idx : 0));
      }
#pragma unroll
      for (int q = 0; q < kRecLoads; ++q) {
        const int idx = threadIdx.x + q * kPersistThreads;
        if (q * kPersistThreads < nrec && idx < nrec) (&srec[0][0])[idx] = rv[q];
      }
    }
    __syncthreads();
    // Every wave reduces the G <= 128 candidates (two per lane) to the same
    // pivot, so that no wave waits for another's.
    double gv = -1.0;
    int grow = INT_MAX;
    for (int b = lane; b < G; b += 64) {
      const double cv = srec[b][0];
      const int crow = int(srec[b][1]);
      if (Better(cv, crow, gv, grow)) {
        gv = cv;
        grow = crow;
      }
    }
    WaveBest(gv, grow);
    // xIAMAX picks the first row when it holds a NaN or the column is zero.
    if (isnan(srec[G][2 + jn]) || !(gv > 0.0)) {
      grow = jn;
      if (gv == 0.0 && gid == 0 && threadIdx.x == 0) atomicMin(g.info, g.row0 + jn + 1);
    }
    if (gid == 0 && threadIdx.x == 0) g.ipiv[jn] = g.row0 + grow + 1;  // 1-based row
    const int pr = grow;
    pivot_rec = (pr == jn) ? G : pr / kPersistRows;
    const double* prow = &srec[pivot_rec][2];
    if (pr != jn) {
#pragma unroll
      for (int q = 0; q < kPersistRpt; ++q) {
        if (rows[q] == jn) {
#pragma unroll
          for (int t = 0; t < kGroupCols; ++t) v[q][t] = prow[c0 + t];
        } else if (rows[q] == pr) {
#pragma unroll
          for (int t = 0; t < kGroupCols; ++t) v[q][t] = srec[G][2 + c0 + t];
        }
      }
    }
  }

#pragma unroll
  for (int q = 0; q < kPersistRpt; ++q) {
    if (rows[q] >= m) continue;
#pragma unroll
    for (int t = 0; t < kGroupCols; ++t) {
      const int c = c0 + t;
      if (c < nb) g.p[rows[q] + int64_t(c) * lda] = v[q][t];
    }
  }

  // Workgroup 0 holds the final unit lower triangle L11 (rows and columns
  // 0 .. mn - 1; rows 0 .. 63 are its waves' q = 0, half = 0 rows) and
  // inverts it for the TRSMs with it: L11 X = I with lane e holding row e
  // of eight right-hand sides per wave, x_e -= L(e, i) x_i in ascending i
  // (the FMA order of InvertUnitLowerKernel's forward substitution).
  if (gid != 0) return;
  __syncthreads();
  double (*st)[kPanel + 1] = reinterpret_cast<double (*)[kPanel + 1]>(&srec[0][0]);
  if (half == 0) {
#pragma unroll
    for (int t = 0; t < kGroupCols; ++t) st[lane][c0 + t] = v[0][t];
  }
  __syncthreads();
  constexpr int kInvRhs = kPanel / (kPersistThreads / 64);  // 8 per wave
  double x[kInvRhs];
  const int cb = wave * kInvRhs;
#pragma unroll
  for (int u = 0; u < kInvRhs; ++u) x[u] = (lane == cb + u) ? 1.0 : 0.0;
  for (int i = 0; i < mn; ++i) {
    const double t = (lane > i && lane < mn) ? st[lane][i] : 0.0;
#pragma unroll
    for (int u = 0; u < kInvRhs; ++u) {
      const long long bits = __double_as_longlong(x[u]);
      const int lo = __builtin_amdgcn_readlane(int(bits & 0xFFFFFFFFll), i);
      const int hi = __builtin_amdgcn_readlane(int(bits >> 32), i);
      const double xi =
          __longlong_as_double((static_cast<long long>(hi) << 32) | static_cast<unsigned int>(lo));
      if (lane > i && lane < mn) x[u] = __builtin_fma(-t, xi, x[u]);
    }
  }
#pragma unroll
  for (int u = 0; u < kInvRhs; ++u) {
    const int col = cb + u;
    g.linv[lane + col * kPanel] = (lane < mn && col < mn) ? x[u] : 0.0;
  }
}

// B (r x n2, ldb) := inv B in place for a kPanel x kPanel inverse (zero
// beyond r): workgroup g stages inv and its 64 columns of B in LDS (every
// load issued before the first wait) and multiplies on the FP64 MFMA
// (v_mfma_f64_16x16x4: wave w owns output rows 16 w .. 16 w + 15 of all four
// column tiles). Each workgroup reads all of its columns before it writes.
__global__ void __launch_bounds__(256)
    InvMulKernel(const double* __restrict__ inv, int r, double* __restrict__ b, int64_t ldb,
                 int n2) {
  __shared__ double si[kPanel][kPanel + 1];
  __shared__ double sb[kPanel][kPanel + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int j0 = blockIdx.x * kPanel;
  constexpr int kPer = kPanel * kPanel / 256;
  double ti[kPer], tb[kPer];
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int idx = tid + 256 * u, i = idx % kPanel, c = idx / kPanel;
    ti[u] = inv[i + c * kPanel];
    tb[u] = (i < r && j0 + c < n2) ? b[i + int64_t(j0 + c) * ldb] : 0.0;
  }
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const int idx = tid + 256 * u, i = idx % kPanel, c = idx / kPanel;
    si[i][c] = ti[u];
    sb[i][c] = tb[u];
  }
  __syncthreads();
  using M = gemm_detail::Mfma<double>;
  M::Acc acc[4] = {};
  const int li = lane & 15, lk = lane >> 4;
  for (int k0 = 0; k0 < kPanel; k0 += 4) {
    const double a = si[16 * wave + li][k0 + lk];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) acc[ct] = M::Run(a, sb[k0 + lk][16 * ct + li], acc[ct]);
  }
#pragma unroll
  for (int ct = 0; ct < 4; ++ct) {
    const int col = j0 + 16 * ct + li;
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int row = 16 * wave + M::Row(lk, rr);
      if (row < r && col < n2) b[row + int64_t(col) * ldb] = acc[ct][rr];
    }
  }
}

struct Work {
  double* base = nullptr;  // the whole matrix
  int ncols = 0;
  double* linv = nullptr;  // inverses of the panels' unit lower triangles
  Candidate* cand = nullptr;
  unsigned* counter = nullptr;
  int* info = nullptr;
  float* f32 = nullptr;
  size_t f32_elems = 0;
  // The cooperative panel kernel (when every workgroup fits on the device).
  int persist_max_groups = 0;
  double* prec = nullptr;      // [2][groups + 1][kRecord]
  unsigned* sync = nullptr;    // [0]: arrivals, [64]: failed
  unsigned arrivals = 0;       // the arrival count after the launches so far
};

// The panel factorization as one launch per column step.
void PanelSteps(double* a, int64_t lda, int m, int nb, int row0, int* ipiv, Work& w,
                hipStream_t s) {
  const double sfmin = std::numeric_limits<double>::min();
  const int mn = std::min(m, nb);
  double* full = w.base + row0;
  for (int j = -1; j < mn; ++j) {
    const int rows = m - (j + 1);
    if (j + 1 >= mn && rows <= 0) break;
    const unsigned blocks = unsigned(std::max(1, (rows + kBlockRows - 1) / kBlockRows));
    hipLaunchKernelGGL(PanelStepKernel, dim3(blocks), dim3(kStepRows), 0, s, a, lda, m, nb, j,
                       full, w.ncols, row0, ipiv, w.cand, w.counter, w.info, sfmin);
  }
}

// Panel a (m x nb) whose first row is row `row0` of the matrix.
void Panel(double* a, int64_t lda, int m, int nb, int row0, int* ipiv, Work& w, hipStream_t s) {
  const double sfmin = std::numeric_limits<double>::min();
  const int mn = std::min(m, nb);
  double* full = w.base + row0;
  const int groups = (m + kPersistRows - 1) / kPersistRows;
  bool persistent = groups <= w.persist_max_groups;
  if (persistent) {
    PersistArgs g{a,           lda,    m,     nb,     row0,
                  ipiv,        w.info, sfmin, w.prec, w.sync,
                  w.sync + 64, w.linv + int64_t(row0 / kPanel) * kPanel * kPanel,
                  w.arrivals};
    void* args[] = {&g};
    // A cooperative launch the runtime refuses (another process holding the
    // device's cooperative queue, for one) leaves this panel and the later
    // ones to the column-step launches.
    const hipError_t e =
        hipLaunchCooperativeKernel(reinterpret_cast<const void*>(PanelPersistentKernel),
                                   dim3(groups), dim3(kPersistThreads), args, 0, s);
    if (e == hipSuccess) {
      w.arrivals += unsigned(mn) * unsigned(groups);  // one arrival per workgroup and column
    } else {
      (void)hipGetLastError();
      w.persist_max_groups = 0;
      persistent = false;
    }
  }
  if (!persistent) {
    // Step j eliminates column j (j >= 0) and pivots column j + 1 (if any).
    PanelSteps(a, lda, m, nb, row0, ipiv, w, s);
  }
  // The panel's interchanges on every other column of the matrix.
  const int pc0 = int((a - full) / lda);  // the panel's first column
  const int others = w.ncols - nb;
  if (others > 0)
    hipLaunchKernelGGL(PanelSwapKernel, dim3((others + 63) / 64), dim3(64), 0, s, full, lda,
                       w.ncols, pc0, nb, ipiv, mn, row0, m);
  // The inverse of the panel's unit lower triangle (final from here on: later
  // interchanges only move rows below it), for the TRSMs with it; the
  // cooperative panel kernel has already written it.
  if (!persistent)
    hipLaunchKernelGGL(InvertUnitLowerKernel, dim3(1), dim3(kPanel), 0, s, a, lda, mn,
                       w.linv + int64_t(row0 / kPanel) * kPanel * kPanel);
  HWY_HIP_CHECK(hipGetLastError());
}

// B (n1 x n2, ldb) := L^-1 B for the unit lower triangle L (n1 x n1) of the
// factorization starting at matrix row g0 (a multiple of kPanel): the
// recursion of the TRSM, with the diagonal blocks applied in place by
// InvMulKernel with the inverses the panels left (rounding as
// X = inv(L_jj) B_j, the usual form of a GPU LU's triangular solves).
void TrsmUnitLower(const double* l, int64_t lda, int n1, double* b, int64_t ldb, int n2, int g0,
                   Work& w, hipStream_t s) {
  if (n1 <= 0 || n2 <= 0) return;
  if (n1 <= kPanel) {
    const double* inv = w.linv + int64_t(g0 / kPanel) * kPanel * kPanel;
    hipLaunchKernelGGL(InvMulKernel, dim3(unsigned((n2 + kPanel - 1) / kPanel)), dim3(256), 0, s,
                       inv, n1, b, ldb, n2);
    HWY_HIP_CHECK(hipGetLastError());
    return;
  }
  const int r1 = std::max(kPanel, (n1 / 2) / kPanel * kPanel);
  TrsmUnitLower(l, lda, r1, b, ldb, n2, g0, w, s);
  GemmDevice<double>(false, false, n1 - r1, n2, r1, -1.0, l + r1, lda, b, ldb, 1.0, b + r1, ldb,
                     s);
  TrsmUnitLower(l + r1 + int64_t(r1) * lda, lda, n1 - r1, b + r1, ldb, n2, g0 + r1, w, s);
}

void Update(int64_t m, int64_t n, int64_t k, const double* a21, const double* a12, double* a22,
            int64_t lda, int update, Work& w, hipStream_t s) {
  if (m <= 0 || n <= 0 || k <= 0) return;
  if (update == 1 && k >= kEmulatedMinK) {
    ozaki::EmulatedDgemm(false, false, m, n, k, -1.0, a21, lda, a12, lda, 1.0, a22, lda, 0, s,
                         nullptr);
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
    // A panel of at most kPanel columns; a wide block (m <= kPanel < n)
    // then has U12 = L11^-1 A12 left for its remaining columns.
    const int nbp = std::min(n, kPanel);
    Panel(a, lda, m, nbp, row0, ipiv, w, s);
    if (n > nbp) TrsmUnitLower(a, lda, m, a + int64_t(nbp) * lda, lda, n - nbp, row0, w, s);
    return;
  }
  int n1 = std::max(kPanel, (mn / 2) / kPanel * kPanel);
  const int n2 = n - n1;
  Recurse(a, lda, m, n1, row0, ipiv, update, w, s);
  // (The left half's row interchanges are already applied to every column.)
  TrsmUnitLower(a, lda, n1, a + int64_t(n1) * lda, lda, n2, row0, w, s);
  Update(m - n1, n2, n1, a + n1, a + int64_t(n1) * lda, a + n1 + int64_t(n1) * lda, lda, update,
         w, s);
  // (Pivots are written as rows of the whole matrix, so none needs an offset.)
  Recurse(a + n1 + int64_t(n1) * lda, lda, m - n1, n2, row0 + n1, ipiv + n1, update, w, s);
}

// How many workgroups of PanelPersistentKernel can be resident at once on
// the current device (0: use one launch per column step).
int PersistentPanelCapacity() {
  const char* e = getenv("HWYBLAS_AMD_LU_PANEL");
  if (e && std::string(e) == "steps") return 0;
  int dev = 0, cus = 0, coop = 0, per_cu = 0;
  HWY_HIP_CHECK(hipGetDevice(&dev));
  HWY_HIP_CHECK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
  HWY_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  HWY_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void*>(PanelPersistentKernel), kPersistThreads, 0));
  return coop ? std::min(per_cu * cus, kPersistMaxGroups) : 0;
}

}  // namespace

// Returns info (0, or the first zero pivot, 1-based, or kGetrfInternalError).
// Synchronizes.
int Getrf(int m, int n, double* a, int64_t lda, int* ipiv, int update, hipStream_t s) {
  if (m == 0 || n == 0) return 0;
  Work w;
  w.base = a;
  w.ncols = n;
  const int max_blocks = (m + kBlockRows - 1) / kBlockRows + 1;
  HWY_HIP_CHECK(hipMalloc(&w.cand, size_t(max_blocks) * sizeof(Candidate)));
  const int panels = (std::min(m, n) + kPanel - 1) / kPanel;
  HWY_HIP_CHECK(hipMalloc(&w.linv, size_t(panels) * kPanel * kPanel * sizeof(double)));
  HWY_HIP_CHECK(hipMalloc(&w.counter, 256));
  HWY_HIP_CHECK(hipMalloc(&w.info, 256));
  hipLaunchKernelGGL(FillUnsigned, dim3(1), dim3(64), 0, s, w.counter, 1, 0u);
  hipLaunchKernelGGL(FillInt, dim3(1), dim3(64), 0, s, w.info, 1, INT_MAX);
  // The cooperative panel kernel serves every panel whose workgroups are all
  // resident at once (HWYBLAS_AMD_LU_PANEL=steps: one launch per column).
  const int groups = (m + kPersistRows - 1) / kPersistRows;
  const int cap = PersistentPanelCapacity();
  if (cap > 0) {
    w.persist_max_groups = std::min(groups, cap);
    const int g = w.persist_max_groups;
    HWY_HIP_CHECK(hipMalloc(&w.prec, size_t(2) * (g + 1) * kRecord * sizeof(double)));
    HWY_HIP_CHECK(hipMalloc(&w.sync, 512));
    hipLaunchKernelGGL(FillUnsigned, dim3(1), dim3(128), 0, s, w.sync, 128, 0u);
    // Test hook: HWYBLAS_AMD_LU_FAIL_BARRIER=1 raises the failure flag before
    // the first panel, so that every bounded wait gives up at its first poll
    // and every workgroup returns: the path a timed-out barrier takes, without
    // waiting for one.
    const char* fail = getenv("HWYBLAS_AMD_LU_FAIL_BARRIER");
    if (fail && *fail == '1')
      hipLaunchKernelGGL(FillUnsigned, dim3(1), dim3(64), 0, s, w.sync + 64, 1, 1u);
  }
  const int kmin = std::min(m, n);
  hipLaunchKernelGGL(IdentityPivotsKernel, dim3((kmin + 255) / 256), dim3(256), 0, s, ipiv, kmin);
  Recurse(a, lda, m, n, 0, ipiv, update, w, s);
  int info = INT_MAX;
  unsigned failed = 0;
  HWY_HIP_CHECK(hipMemcpyAsync(&info, w.info, sizeof(int), hipMemcpyDeviceToHost, s));
  if (w.sync)
    HWY_HIP_CHECK(hipMemcpyAsync(&failed, w.sync + 64, sizeof(unsigned), hipMemcpyDeviceToHost,
                                 s));
  HWY_HIP_CHECK(hipStreamSynchronize(s));
  for (void* p : {static_cast<void*>(w.cand), static_cast<void*>(w.counter),
                  static_cast<void*>(w.info), static_cast<void*>(w.f32),
                  static_cast<void*>(w.linv),
                  static_cast<void*>(w.prec), static_cast<void*>(w.sync)})
    if (p) HWY_HIP_CHECK(hipFree(p));
  if (failed) {
    fprintf(stderr, "hwyblas_amd: dgetrf's panel kernel timed out at a grid barrier\n");
    return kGetrfInternalError;
  }
  return info == INT_MAX ? 0 : info;
}

// src[i] of the composed row interchanges ipiv[0..n) (1-based, forward).
std::vector<int> PivotPermutation(int n, const int* ipiv_host) {
  std::vector<int> src(n);
  for (int i = 0; i < n; ++i) src[i] = i;
  for (int i = 0; i < n; ++i) std::swap(src[i], src[ipiv_host[i] - 1]);
  return src;
}

// Applies P^T (forward interchanges) or P (reverse) to the n x ncols B,
// with src the forward permutation (device) and scratch n * ncols doubles.
void Permute(bool forward, int n, int ncols, double* b, int64_t ldb, const int* src_dev,
             const int* inv_dev, double* scratch, hipStream_t s) {
  hipLaunchKernelGGL(GatherRows, dim3(GridFor(n)), dim3(256), 0, s, b, ldb, n, ncols,
                     forward ? src_dev : inv_dev, scratch);
  hipLaunchKernelGGL(CopyBack, dim3(GridFor(n)), dim3(256), 0, s, scratch, n, ncols, b, ldb);
  HWY_HIP_CHECK(hipGetLastError());
}

// Inverses of the kPanel x kPanel diagonal blocks of the factors (block b:
// rows and columns [kPanel b, kPanel b + r)), one workgroup per block and
// one thread per column j of the inverse, every block in one launch: for L
// (unit lower) x_i = delta_ij - sum_{q < i} L(i, q) x_q, i ascending; for U
// (upper) x_i = (delta_ij - sum_{q > i} U(i, q) x_q) / U(i, i), i
// descending. out: per block kPanel x kPanel column-major, zero beyond r.
template <bool UPPER>
__global__ void __launch_bounds__(kPanel)
    InvertDiagBlocksKernel(const double* __restrict__ a, int64_t lda, int n,
                           double* __restrict__ out) {
  __shared__ double t[kPanel][kPanel + 1];
  const int j = threadIdx.x, i0 = blockIdx.x * kPanel, r = min(kPanel, n - i0);
  const double* d = a + i0 + int64_t(i0) * lda;
  {
    double colv[kPanel];
#pragma unroll
    for (int i = 0; i < kPanel; ++i)
      colv[i] = (j < r && i < r && (UPPER ? i <= j : i > j)) ? d[i + int64_t(j) * lda] : 0.0;
#pragma unroll
    for (int i = 0; i < kPanel; ++i) t[i][j] = colv[i];
  }
  __syncthreads();
  double x[kPanel];
  if (UPPER) {
#pragma unroll
    for (int i = kPanel - 1; i >= 0; --i) {
      double v = (i == j) ? 1.0 : 0.0;
#pragma unroll
      for (int q = i + 1; q < kPanel; ++q) v = __builtin_fma(-t[i][q], x[q], v);
      x[i] = i < r ? v / t[i][i] : 0.0;
    }
  } else {
#pragma unroll
    for (int i = 0; i < kPanel; ++i) {
      double v = (i == j) ? 1.0 : 0.0;
#pragma unroll
      for (int q = 0; q < i; ++q) v = __builtin_fma(-t[i][q], x[q], v);
      x[i] = v;
    }
  }
  double* o = out + int64_t(blockIdx.x) * kPanel * kPanel;
#pragma unroll
  for (int i = 0; i < kPanel; ++i) o[i + j * kPanel] = (i < r && j < r) ? x[i] : 0.0;
}

constexpr int kStepRhs = 16;  // right-hand sides per workgroup of InvStepKernel

// One block step of a blocked triangular solve with the block's inverse:
// x_i = M b_i for the diagonal block rows [i0, i0 + r) (M = minv, or its
// transpose when MTRANS), written to x by workgroup 0 of each
// right-hand-side chunk, then b(s, :) -= T(s, i0 .. i0 + r - 1) x_i for this
// workgroup's kPanel rows s of [u0, u1), in place. T(s, c) = a[s + c ld]
// (COLS: T is the stored matrix) or a[s ld + c] (T is its transpose). b's
// rows [i0, i0 + r) are only read and the updated rows are disjoint from
// them, so no workgroup reads what another writes. Every global load is
// issued before the first use (one memory round trip; out-of-range
// elements read a valid address and are replaced by zero). Both products
// run on the FP64 MFMA (v_mfma_f64_16x16x4: wave w owns rows 16 w .. + 15
// of the 16 right-hand sides), the update as D = b + (-T) x with b in the
// accumulator; M and T are zero beyond r, so every product spans kPanel.
template <bool COLS, bool MTRANS>
__global__ void __launch_bounds__(256)
    InvStepKernel(const double* __restrict__ a, int64_t ld, const double* __restrict__ minv,
                  int i0, int r, int u0, int u1, double* __restrict__ b, int64_t ldb,
                  double* __restrict__ x, int64_t ldx, int nrhs) {
  static_assert(kStepRhs == 16 && kPanel == 64, "one 16 x 16 MFMA tile per wave");
  constexpr int kPer = kPanel * kPanel / 256;       // M and T elements per thread
  constexpr int kPerB = kPanel * kStepRhs / 256;   // b_i elements per thread
  using MF = gemm_detail::Mfma<double>;
  __shared__ double sm[kPanel][kPanel + 1];  // M, then the slab of -T
  __shared__ double sb[kPanel][kStepRhs + 1];
  __shared__ double sx[kPanel][kStepRhs + 1];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int j0 = blockIdx.y * kStepRhs, nj = min(kStepRhs, nrhs - j0);
  const int s0 = u0 + blockIdx.x * kPanel;
  const int sn = max(0, min(kPanel, u1 - s0));
  // Element (p, w) of a kPanel x kPanel tile for the q-th load of a thread:
  // p = idx % kPanel runs along consecutive threads.
  double mv[kPer], tv[kPer], bv[kPerB];
  MF::Acc acc;
#pragma unroll
  for (int q = 0; q < kPer; ++q) mv[q] = minv[tid + q * 256];
#pragma unroll
  for (int q = 0; q < kPerB; ++q) {
    const int idx = tid + q * 256, k = idx % kPanel, jj = idx / kPanel;
    const bool ok = k < r && jj < nj;
    const double v = b[ok ? (i0 + k) + int64_t(j0 + jj) * ldb : int64_t(i0)];
    bv[q] = ok ? v : 0.0;
  }
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    // COLS: p indexes rows s (contiguous), w columns; else the reverse.
    const int idx = tid + q * 256, p = idx % kPanel, w = idx / kPanel;
    const int ss = COLS ? p : w, k = COLS ? w : p;
    const bool ok = ss < sn && k < r;
    const int64_t off = COLS ? int64_t(s0 + ss) + int64_t(i0 + k) * ld
                             : int64_t(s0 + ss) * ld + int64_t(i0 + k);
    const double v = a[ok ? off : 0];
    tv[q] = ok ? -v : 0.0;
  }
#pragma unroll
  for (int rr = 0; rr < 4; ++rr) {
    const int row = 16 * wave + MF::Row(lk, rr);
    const bool ok = row < sn && li < nj;
    const double v = b[ok ? (s0 + row) + int64_t(j0 + li) * ldb : int64_t(i0)];
    acc[rr] = ok ? v : 0.0;
  }

#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int idx = tid + q * 256, p = idx % kPanel, w = idx / kPanel;  // M(p, w)
    if (MTRANS) sm[w][p] = mv[q];
    else sm[p][w] = mv[q];
  }
#pragma unroll
  for (int q = 0; q < kPerB; ++q) {
    const int idx = tid + q * 256;
    sb[idx % kPanel][idx / kPanel] = bv[q];
  }
  __syncthreads();
  {
    MF::Acc xa = {};
#pragma unroll
    for (int k0 = 0; k0 < kPanel; k0 += 4)
      xa = MF::Run(sm[16 * wave + li][k0 + lk], sb[k0 + lk][li], xa);
#pragma unroll
    for (int rr = 0; rr < 4; ++rr) {
      const int row = 16 * wave + MF::Row(lk, rr);
      sx[row][li] = xa[rr];
      if (blockIdx.x == 0 && row < r && li < nj) x[(i0 + row) + int64_t(j0 + li) * ldx] = xa[rr];
    }
  }
  if (sn == 0) return;
  __syncthreads();
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int idx = tid + q * 256, p = idx % kPanel, w = idx / kPanel;
    if (COLS) sm[p][w] = tv[q];
    else sm[w][p] = tv[q];
  }
  __syncthreads();
#pragma unroll
  for (int k0 = 0; k0 < kPanel; k0 += 4)
    acc = MF::Run(sm[16 * wave + li][k0 + lk], sx[k0 + lk][li], acc);
#pragma unroll
  for (int rr = 0; rr < 4; ++rr) {
    const int row = 16 * wave + MF::Row(lk, rr);
    if (row < sn && li < nj) b[(s0 + row) + int64_t(j0 + li) * ldb] = acc[rr];
  }
}

// A triangular solve of few right-hand sides by kPanel-row blocks with the
// blocks' inverses (minv, kPanel x kPanel each, transposed when mtrans):
// forward (T lower: blocks ascending, rows below updated) or backward (T
// upper: blocks descending, rows above updated). b is consumed (its unsolved
// rows updated in place); the solution goes to x. One launch per block.
// trans: T is the transpose of the stored triangle, and M the transpose of
// the stored inverses (both are then read transposed).
void SolveWithBlockInverses(bool forward, int n, int nrhs, const double* a, int64_t lda,
                            bool trans, const double* minv, double* b, int64_t ldb, double* x,
                            int64_t ldx, hipStream_t s) {
  const int nblk = (n + kPanel - 1) / kPanel;
  const unsigned ychunks = unsigned((nrhs + kStepRhs - 1) / kStepRhs);
  for (int t = 0; t < nblk; ++t) {
    const int blk = forward ? t : nblk - 1 - t;
    const int i0 = blk * kPanel, r = std::min(kPanel, n - i0);
    const int u0 = forward ? i0 + r : 0, u1 = forward ? n : i0;
    const dim3 grid(unsigned(std::max(1, (u1 - u0 + kPanel - 1) / kPanel)), ychunks);
    const double* m = minv + int64_t(blk) * kPanel * kPanel;
    if (trans)
      hipLaunchKernelGGL((InvStepKernel<false, true>), grid, dim3(256), 0, s, a, lda, m, i0, r,
                         u0, u1, b, ldb, x, ldx, nrhs);
    else
      hipLaunchKernelGGL((InvStepKernel<true, false>), grid, dim3(256), 0, s, a, lda, m, i0, r,
                         u0, u1, b, ldb, x, ldx, nrhs);
  }
  HWY_HIP_CHECK(hipGetLastError());
}

// The row permutation of a factorization, on the device, and (once asked
// for) the inverses of its factors' diagonal blocks.
struct Perm {
  int* src = nullptr;
  int* inv = nullptr;
  double* scratch = nullptr;
  int n = 0, ncols = 0;
  mutable double* block_inv = nullptr;  // [2][blocks][kPanel x kPanel]: L's, then U's
  mutable const double* block_inv_of = nullptr;

  // The diagonal-block inverses of the factors in a (computed for the
  // first a asked about; a Perm serves one factorization).
  const double* BlockInverses(const double* a, int64_t lda, hipStream_t s) const {
    const int nblk = (n + kPanel - 1) / kPanel;
    if (block_inv_of != a) {
      if (!block_inv)
        HWY_HIP_CHECK(hipMalloc(&block_inv, size_t(2) * nblk * kPanel * kPanel * sizeof(double)));
      hipLaunchKernelGGL(InvertDiagBlocksKernel<false>, dim3(nblk), dim3(kPanel), 0, s, a, lda,
                         n, block_inv);
      hipLaunchKernelGGL(InvertDiagBlocksKernel<true>, dim3(nblk), dim3(kPanel), 0, s, a, lda, n,
                         block_inv + int64_t(nblk) * kPanel * kPanel);
      HWY_HIP_CHECK(hipGetLastError());
      block_inv_of = a;
    }
    return block_inv;
  }

  Perm(int n_, int ncols_, const int* ipiv_dev, hipStream_t s) : n(n_), ncols(ncols_) {
    std::vector<int> piv(n);
    HWY_HIP_CHECK(hipMemcpyAsync(piv.data(), ipiv_dev, size_t(n) * sizeof(int),
                                 hipMemcpyDeviceToHost, s));
    HWY_HIP_CHECK(hipStreamSynchronize(s));
    const std::vector<int> src_h = PivotPermutation(n, piv.data());
    std::vector<int> inv_h(n);
    for (int i = 0; i < n; ++i) inv_h[src_h[i]] = i;
    HWY_HIP_CHECK(hipMalloc(&src, size_t(n) * sizeof(int)));
    HWY_HIP_CHECK(hipMalloc(&inv, size_t(n) * sizeof(int)));
    HWY_HIP_CHECK(hipMalloc(&scratch, size_t(n) * std::max(ncols, 1) * sizeof(double)));
    HWY_HIP_CHECK(hipMemcpy(src, src_h.data(), size_t(n) * sizeof(int), hipMemcpyHostToDevice));
    HWY_HIP_CHECK(hipMemcpy(inv, inv_h.data(), size_t(n) * sizeof(int), hipMemcpyHostToDevice));
  }
  ~Perm() {
    (void)hipFree(src);
    (void)hipFree(inv);
    (void)hipFree(scratch);
    if (block_inv) (void)hipFree(block_inv);
  }
};

// The two triangular solves of getrs for few right-hand sides, blocked
// right-looking by kPanel rows: a diagonal-block solve (substitution, as
// the TRSM's) and one update of all remaining rows, a GEMM of depth kPanel
// whose many row tiles keep the whole GPU busy (the recursive TRSM's deep
// updates of depth n/2 have one tile column and stream A through few
// workgroups).
void SolveLowerUnitBlocked(int n, int nrhs, const double* a, int64_t lda, double* b, int64_t ldb,
                           hipStream_t s) {
  for (int i0 = 0; i0 < n; i0 += kPanel) {
    const int r = std::min(kPanel, n - i0);
    TrsmDevice<double>(true, true, false, true, r, nrhs, a + i0 + int64_t(i0) * lda, lda, b + i0,
                       ldb, s);
    if (i0 + r < n)
      GemmDevice<double>(false, false, n - i0 - r, nrhs, r, -1.0,
                         a + (i0 + r) + int64_t(i0) * lda, lda, b + i0, ldb, 1.0, b + i0 + r,
                         ldb, s);
  }
}

void SolveUpperBlocked(int n, int nrhs, const double* a, int64_t lda, double* b, int64_t ldb,
                       hipStream_t s) {
  const int nblk = (n + kPanel - 1) / kPanel;
  for (int blk = nblk - 1; blk >= 0; --blk) {
    const int i0 = blk * kPanel, r = std::min(kPanel, n - i0);
    TrsmDevice<double>(true, false, false, false, r, nrhs, a + i0 + int64_t(i0) * lda, lda,
                       b + i0, ldb, s);
    if (i0 > 0)
      GemmDevice<double>(false, false, i0, nrhs, r, -1.0, a + int64_t(i0) * lda, lda, b + i0, ldb,
                         1.0, b, ldb, s);
  }
}

// op(U) = U^T is lower triangular: blocks go forward, and the update of
// the rows below a block reads the block's rows of U right of the
// diagonal, transposed.
void SolveUpperTransBlocked(int n, int nrhs, const double* a, int64_t lda, double* b,
                            int64_t ldb, hipStream_t s) {
  for (int i0 = 0; i0 < n; i0 += kPanel) {
    const int r = std::min(kPanel, n - i0);
    TrsmDevice<double>(true, false, true, false, r, nrhs, a + i0 + int64_t(i0) * lda, lda,
                       b + i0, ldb, s);
    if (i0 + r < n)
      GemmDevice<double>(true, false, n - i0 - r, nrhs, r, -1.0,
                         a + i0 + int64_t(i0 + r) * lda, lda, b + i0, ldb, 1.0, b + i0 + r,
                         ldb, s);
  }
}

// op(L) = L^T is unit upper triangular: blocks go backward, and the update
// of the rows above a block reads the block's rows of L left of the
// diagonal, transposed.
void SolveLowerUnitTransBlocked(int n, int nrhs, const double* a, int64_t lda, double* b,
                                int64_t ldb, hipStream_t s) {
  const int nblk = (n + kPanel - 1) / kPanel;
  for (int blk = nblk - 1; blk >= 0; --blk) {
    const int i0 = blk * kPanel, r = std::min(kPanel, n - i0);
    TrsmDevice<double>(true, true, true, true, r, nrhs, a + i0 + int64_t(i0) * lda, lda, b + i0,
                       ldb, s);
    if (i0 > 0)
      GemmDevice<double>(true, false, i0, nrhs, r, -1.0, a + i0, lda, b + i0, ldb, 1.0, b, ldb,
                         s);
  }
}

constexpr int kBlockedSolveMaxRhs = 64;

// xGETRS (src/lapack/getrf.cc:287-339) with the permutation of ipiv.
// For nrhs <= kBlockedSolveMaxRhs the triangular solves go by kPanel-row
// blocks with the diagonal blocks' inverses, one launch per block
// (HWYBLAS_AMD_GETRS=trsm: a TRSM per diagonal block plus a GEMM update).
bool GetrsWithTrsm() {
  const char* e = getenv("HWYBLAS_AMD_GETRS");
  return e && std::string(e) == "trsm";
}

void GetrsPerm(bool trans, int n, int nrhs, const double* a, int64_t lda, const Perm& perm,
               double* b, int64_t ldb, hipStream_t s) {
  if (n == 0 || nrhs == 0) return;
  if (nrhs <= kBlockedSolveMaxRhs && nrhs <= perm.ncols && !GetrsWithTrsm()) {
    const double* inv = perm.BlockInverses(a, lda, s);
    const double* linv = inv;
    const double* uinv = inv + int64_t((n + kPanel - 1) / kPanel) * kPanel * kPanel;
    double* y = perm.scratch;  // n x nrhs, free between the permutations
    if (!trans) {
      Permute(true, n, nrhs, b, ldb, perm.src, perm.inv, perm.scratch, s);
      SolveWithBlockInverses(true, n, nrhs, a, lda, false, linv, b, ldb, y, n, s);
      SolveWithBlockInverses(false, n, nrhs, a, lda, false, uinv, y, n, b, ldb, s);
    } else {
      SolveWithBlockInverses(true, n, nrhs, a, lda, true, uinv, b, ldb, y, n, s);
      SolveWithBlockInverses(false, n, nrhs, a, lda, true, linv, y, n, b, ldb, s);
      Permute(false, n, nrhs, b, ldb, perm.src, perm.inv, perm.scratch, s);
    }
    return;
  }
  if (!trans) {
    Permute(true, n, nrhs, b, ldb, perm.src, perm.inv, perm.scratch, s);
    if (nrhs <= kBlockedSolveMaxRhs) {
      SolveLowerUnitBlocked(n, nrhs, a, lda, b, ldb, s);
      SolveUpperBlocked(n, nrhs, a, lda, b, ldb, s);
      return;
    }
    TrsmDevice<double>(true, true, false, true, n, nrhs, a, lda, b, ldb, s);
    TrsmDevice<double>(true, false, false, false, n, nrhs, a, lda, b, ldb, s);
  } else {
    if (nrhs <= kBlockedSolveMaxRhs) {
      SolveUpperTransBlocked(n, nrhs, a, lda, b, ldb, s);
      SolveLowerUnitTransBlocked(n, nrhs, a, lda, b, ldb, s);
    } else {
      TrsmDevice<double>(true, false, true, false, n, nrhs, a, lda, b, ldb, s);
      TrsmDevice<double>(true, true, true, true, n, nrhs, a, lda, b, ldb, s);
    }
    Permute(false, n, nrhs, b, ldb, perm.src, perm.inv, perm.scratch, s);
  }
}

// Synchronizes (the permutation is composed on the host).
void Getrs(bool trans, int n, int nrhs, const double* a, int64_t lda, const int* ipiv, double* b,
           int64_t ldb, hipStream_t s) {
  if (n == 0 || nrhs == 0) return;
  Perm perm(n, nrhs, ipiv, s);
  GetrsPerm(trans, n, nrhs, a, lda, perm, b, ldb, s);
  HWY_HIP_CHECK(hipStreamSynchronize(s));
}

// ---------------------------------------------------------------------------
// Refinement

constexpr int kMaxRhs = 16;
constexpr int kResChunk = 512;  // columns of A per workgroup of the residual

// Partial sums over the columns [k0, k0 + kResChunk) of A: rp = sum a x,
// wp = sum |a| |x| for every row (one thread each) and right-hand side;
// with ap, also the row's sum of |a| over the chunk (for ||A||inf).
__global__ void __launch_bounds__(256)
    ResidualPartialKernel(const double* __restrict__ a, int64_t lda, int n,
                          const double* __restrict__ x, int64_t ldx, int nrhs,
                          double* __restrict__ rp, double* __restrict__ wp,
                          double* __restrict__ ap) {
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
  double sa = 0.0;
  for (int kk = 0; kk < kn; ++kk) {
    const double v = ai[int64_t(kk) * lda];
    const double av = fabs(v);
    sa += av;
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
  if (ap) ap[int64_t(blockIdx.y) * n + i] = sa;
}

// The same sums for op(A) = A^T, in ResidualPartialKernel's layout with one
// chunk: rp(i) = sum_k a(k, i) x(k), wp(i) = sum_k |a(k, i)| |x(k)| for every
// right-hand side, and ap(i) = sum_k |a(k, i)| (row sums of |A^T|). One wave
// per column i of A (contiguous), its lanes striding k; the lanes' sums are
// combined by a fixed butterfly.
__global__ void __launch_bounds__(256)
    ResidualTransKernel(const double* __restrict__ a, int64_t lda, int n,
                        const double* __restrict__ x, int64_t ldx, int nrhs,
                        double* __restrict__ rp, double* __restrict__ wp,
                        double* __restrict__ ap) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= n) return;  // uniform per wave
  double r[kMaxRhs], w[kMaxRhs];
#pragma unroll
  for (int j = 0; j < kMaxRhs; ++j) r[j] = w[j] = 0.0;
  double sa = 0.0;
  const double* ai = a + int64_t(i) * lda;
  for (int k = lane; k < n; k += 64) {
    const double v = ai[k];
    const double av = fabs(v);
    sa += av;
#pragma unroll
    for (int j = 0; j < kMaxRhs; ++j) {
      const double xk = j < nrhs ? x[k + int64_t(j) * ldx] : 0.0;
      r[j] = __builtin_fma(v, xk, r[j]);
      w[j] = __builtin_fma(av, fabs(xk), w[j]);
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    sa += __shfl_xor(sa, off, 64);
#pragma unroll
    for (int j = 0; j < kMaxRhs; ++j) {
      r[j] += __shfl_xor(r[j], off, 64);
      w[j] += __shfl_xor(w[j], off, 64);
    }
  }
  if (lane == 0) {
    for (int j = 0; j < nrhs; ++j) {
      rp[int64_t(j) * n + i] = r[j];
      wp[int64_t(j) * n + i] = w[j];
    }
    if (ap) ap[i] = sa;
  }
}

// R = B - sum rp; D = |B| + sum wp; per workgroup and right-hand side the
// largest berr term and the largest |R| (columns of `stats`: 2 * nrhs).
// With w_out, also the weights of the forward error bound,
// W = |R| + nzeps D, plus safe1 where D <= safe2 (leading dimension n).
__global__ void __launch_bounds__(256)
    ResidualFinishKernel(const double* __restrict__ rp, const double* __restrict__ wp,
                         int nchunks, const double* __restrict__ b, int64_t ldb, int n, int nrhs,
                         double* __restrict__ r_out, int64_t ldr, double safe1, double safe2,
                         double* __restrict__ stats, double* __restrict__ w_out,
                         double nzeps) {
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
      if (w_out)
        w_out[i + int64_t(j) * n] = fabs(rij) + nzeps * wij + (wij > safe2 ? 0.0 : safe1);
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

// The right-hand side each packed column of the forward-error products
// belongs to.
struct ColumnList {
  int col[kMaxRhs];
};

// z(:, c) *= w(:, list.col[c]) for the packed columns c (grid.y).
__global__ void __launch_bounds__(256)
    ScaleByWeightsKernel(double* __restrict__ z, int n, const double* __restrict__ w,
                         ColumnList list) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int c = blockIdx.y;
  z[i + int64_t(c) * n] *= w[i + int64_t(list.col[c]) * n];
}

// Row sums of |A| from the residual pass's per-chunk sums (chunks in order).
__global__ void RowAbsSumKernel(const double* __restrict__ ap, int nchunks, int n,
                                double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double s = 0.0;
  for (int c = 0; c < nchunks; ++c) s += ap[int64_t(c) * n + i];
  out[i] = s;
}

// Hager's estimate of ||M||_1 with Higham's refinements, the steps of
// xLACN2 (src/lapack/trsen.cc:2235-2366), driven from the host for one
// vector: Start() and Next() name the product wanted next in x(), which
// the caller replaces by M x (kTimesM) or M^T x (kTimesMt) before calling
// Next() again, until kDone. Every estimate is |M w|_1 / |w|_1 for some w.
class NormEstimator {
 public:
  enum Request { kDone = 0, kTimesM = 1, kTimesMt = 2 };

  explicit NormEstimator(int n) : n_(n), x_(n), sign_(n) {}
  double* x() { return x_.data(); }
  double estimate() const { return est_; }

  Request Start() {
    std::fill(x_.begin(), x_.end(), 1.0 / double(n_));
    stage_ = kAfterFirst;
    return kTimesM;
  }

  Request Next() {
    switch (stage_) {
      case kAfterFirst:  // x = M (1/n)
        if (n_ == 1) {
          est_ = std::fabs(x_[0]);
          return kDone;
        }
        est_ = Norm1();
        return AskSigns(kAfterFirstT);
      case kAfterFirstT:  // x = M^T sign
        j_ = FirstLargest();
        iter_ = 2;
        return AskUnit();
      case kAfterUnit: {  // x = M e_j
        const double prev = est_;
        est_ = Norm1();
        bool repeated = true;
        for (int i = 0; i < n_ && repeated; ++i) repeated = Sign(x_[i]) == sign_[i];
        if (repeated || est_ <= prev) return AskAlternating();
        return AskSigns(kAfterLoopT);
      }
      case kAfterLoopT: {  // x = M^T sign
        const int last = j_;
        j_ = FirstLargest();
        if (x_[last] != std::fabs(x_[j_]) && iter_ < kMaxIter) {
          ++iter_;
          return AskUnit();
        }
        return AskAlternating();
      }
      case kAfterAlternating: {  // x = M b
        const double t = 2.0 * (Norm1() / double(3 * n_));
        if (t > est_) est_ = t;
        return kDone;
      }
    }
    return kDone;
  }

 private:
  enum Stage { kAfterFirst, kAfterFirstT, kAfterUnit, kAfterLoopT, kAfterAlternating };
  static constexpr int kMaxIter = 5;

  static int Sign(double t) { return t >= 0.0 ? 1 : -1; }
  double Norm1() const {
    double s = 0.0;
    for (double t : x_) s += std::fabs(t);
    return s;
  }
  // First index of the largest magnitude; 0 when x[0] is NaN, other NaNs
  // skipped (as xIAMAX).
  int FirstLargest() const {
    if (std::isnan(x_[0])) return 0;
    int best = 0;
    double bv = std::fabs(x_[0]);
    for (int i = 1; i < n_; ++i)
      if (std::fabs(x_[i]) > bv) {
        bv = std::fabs(x_[i]);
        best = i;
      }
    return best;
  }
  Request AskSigns(Stage next) {
    for (int i = 0; i < n_; ++i) {
      sign_[i] = Sign(x_[i]);
      x_[i] = double(sign_[i]);
    }
    stage_ = next;
    return kTimesMt;
  }
  Request AskUnit() {
    std::fill(x_.begin(), x_.end(), 0.0);
    x_[j_] = 1.0;
    stage_ = kAfterUnit;
    return kTimesM;
  }
  // b_i = (-1)^i (1 + i / (n - 1)).
  Request AskAlternating() {
    for (int i = 0; i < n_; ++i) x_[i] = (i % 2 ? -1.0 : 1.0) * (1.0 + double(i) / double(n_ - 1));
    stage_ = kAfterAlternating;
    return kTimesM;
  }

  int n_;
  std::vector<double> x_;
  std::vector<int> sign_;
  Stage stage_ = kAfterFirst;
  double est_ = 0.0;
  int j_ = 0, iter_ = 0;
};

// gerfs's forward error bound (src/lapack/gesvx.cc:465-518) for every
// right-hand side at once: est_j = ||inv(A) diag(W_j)||inf, estimated as
// ||M_j||_1 with M_j = diag(W_j) inv(A)^T. The estimators of all columns
// advance together; each round packs the vectors that want M x (getrs with
// A^T, then times W) and those that want M^T x (times W, then getrs with
// A) into z, one solve per kind. w: the weights (n x nrhs, leading
// dimension n); z: n x nrhs scratch.
void ForwardErrorEstimates(bool trans, int n, int nrhs, const double* af, int64_t ldaf,
                           const Perm& perm, const double* w, double* z, hipStream_t s,
                           double* est) {
  std::vector<NormEstimator> ests;
  std::vector<NormEstimator::Request> req(nrhs);
  for (int j = 0; j < nrhs; ++j) {
    ests.emplace_back(n);
    req[j] = ests[j].Start();
  }
  std::vector<double> packed(size_t(n) * nrhs);
  const int rblocks = (n + 255) / 256;
  for (;;) {
    ColumnList first{}, second{};
    int c1 = 0, c2 = 0;
    for (int j = 0; j < nrhs; ++j) {
      if (req[j] == NormEstimator::kTimesM) first.col[c1++] = j;
      if (req[j] == NormEstimator::kTimesMt) second.col[c2++] = j;
    }
    if (c1 + c2 == 0) break;
    for (int c = 0; c < c1 + c2; ++c) {
      const int j = c < c1 ? first.col[c] : second.col[c - c1];
      std::copy(ests[j].x(), ests[j].x() + n, packed.data() + size_t(c) * n);
    }
    HWY_HIP_CHECK(hipMemcpyAsync(z, packed.data(), size_t(n) * (c1 + c2) * 8,
                                 hipMemcpyHostToDevice, s));
    double* z2 = z + size_t(c1) * n;
    // M = diag(W) op(A)^-T: M x solves with op(A)^T first, M^T x with op(A).
    if (c1 > 0) {
      GetrsPerm(!trans, n, c1, af, ldaf, perm, z, n, s);
      hipLaunchKernelGGL(ScaleByWeightsKernel, dim3(rblocks, c1), dim3(256), 0, s, z, n, w,
                         first);
    }
    if (c2 > 0) {
      hipLaunchKernelGGL(ScaleByWeightsKernel, dim3(rblocks, c2), dim3(256), 0, s, z2, n, w,
                         second);
      GetrsPerm(trans, n, c2, af, ldaf, perm, z2, n, s);
    }
    HWY_HIP_CHECK(hipGetLastError());
    HWY_HIP_CHECK(hipMemcpyAsync(packed.data(), z, size_t(n) * (c1 + c2) * 8,
                                 hipMemcpyDeviceToHost, s));
    HWY_HIP_CHECK(hipStreamSynchronize(s));
    for (int c = 0; c < c1 + c2; ++c) {
      const int j = c < c1 ? first.col[c] : second.col[c - c1];
      std::copy(packed.data() + size_t(c) * n, packed.data() + size_t(c + 1) * n, ests[j].x());
      req[j] = ests[j].Next();
    }
  }
  for (int j = 0; j < nrhs; ++j) est[j] = ests[j].estimate();
}

// gerfs's refinement (src/lapack/gesvx.cc:333-521) of X (n x nrhs,
// nrhs <= kMaxRhs) for op(A) X = B with the factors af and their
// permutation: per column, while berr > eps, berr at most half the last one
// and fewer than itmax corrections, X += op(A)^-1 (B - op(A) X), all columns
// of a round in one residual pass over A and one solve; then, with ferr, the
// forward error bound of the final X. berr, ferr, nbe (the normwise backward
// error ||B - op(A) X||inf / (||op(A)||inf ||X||inf)) and iterations (the
// corrections applied) are host arrays; ferr, nbe and iterations may be null.
void Refine(bool trans, int n, int nrhs, const double* a, int64_t lda, const double* af,
            int64_t ldaf, const Perm& perm, const double* b, int64_t ldb, double* x,
            int64_t ldx, int itmax, double* ferr, double* berr, double* nbe, int* iterations,
            hipStream_t s) {
  const int nchunks = trans ? 1 : (n + kResChunk - 1) / kResChunk;
  const int rblocks = (n + 255) / 256;
  double *rp, *wp, *r, *stats, *rowsum, *rowpart, *wts = nullptr, *z = nullptr;
  int* active_d;
  if (ferr) {
    HWY_HIP_CHECK(hipMalloc(&wts, size_t(n) * nrhs * 8));
    HWY_HIP_CHECK(hipMalloc(&z, size_t(n) * nrhs * 8));
  }
  HWY_HIP_CHECK(hipMalloc(&rp, size_t(nchunks) * kMaxRhs * n * 8));
  HWY_HIP_CHECK(hipMalloc(&wp, size_t(nchunks) * kMaxRhs * n * 8));
  HWY_HIP_CHECK(hipMalloc(&r, size_t(n) * nrhs * 8));
  HWY_HIP_CHECK(hipMalloc(&stats, size_t(rblocks) * 2 * nrhs * 8));
  HWY_HIP_CHECK(hipMalloc(&rowsum, size_t(n) * 8));
  HWY_HIP_CHECK(hipMalloc(&rowpart, size_t(nchunks) * n * 8));
  HWY_HIP_CHECK(hipMalloc(&active_d, kMaxRhs * sizeof(int)));
  const double eps = std::numeric_limits<double>::epsilon() * 0.5;  // LAPACK's dlamch('E')
  const double safmin = std::numeric_limits<double>::min();
  const double safe1 = double(n + 1) * safmin, safe2 = safe1 / eps;
  std::vector<double> host_stats(size_t(rblocks) * 2 * nrhs);
  std::vector<double> lstres(nrhs, 3.0), rinf(nrhs, 0.0);
  std::vector<int> count(nrhs, 1), active(nrhs, 1), done(nrhs, 0);
  for (bool first = true;; first = false) {
    if (trans)
      hipLaunchKernelGGL(ResidualTransKernel, dim3((n + 3) / 4), dim3(256), 0, s, a, lda, n, x,
                         ldx, nrhs, rp, wp, first ? rowpart : nullptr);
    else
      hipLaunchKernelGGL(ResidualPartialKernel, dim3(rblocks, nchunks), dim3(256), 0, s, a,
                         lda, n, x, ldx, nrhs, rp, wp, first ? rowpart : nullptr);
    hipLaunchKernelGGL(ResidualFinishKernel, dim3(rblocks), dim3(256), 0, s, rp, wp, nchunks, b,
                       ldb, n, nrhs, r, int64_t(n), safe1, safe2, stats, wts,
                       double(n + 1) * eps);
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
    GetrsPerm(trans, n, nrhs, af, ldaf, perm, r, n, s);
    hipLaunchKernelGGL(AddMaskedKernel, dim3(rblocks), dim3(256), 0, s, x, ldx, r, int64_t(n), n,
                       nrhs, active_d);
    HWY_HIP_CHECK(hipGetLastError());
  }
  // The last residual pass was of the final X: its weights bound the error.
  std::vector<double> est(nrhs, 0.0);
  if (ferr) ForwardErrorEstimates(trans, n, nrhs, af, ldaf, perm, wts, z, s, est.data());
  // Normwise backward error of the final X.
  hipLaunchKernelGGL(RowAbsSumKernel, dim3(rblocks), dim3(256), 0, s, rowpart, nchunks, n,
                     rowsum);
  std::vector<double> rows(n), xs(size_t(n) * nrhs);
  HWY_HIP_CHECK(hipMemcpyAsync(rows.data(), rowsum, size_t(n) * 8, hipMemcpyDeviceToHost, s));
  HWY_HIP_CHECK(hwy_amd::Memcpy2DAsync(xs.data(), size_t(n) * 8, x, size_t(ldx) * 8, size_t(n) * 8,
                                       nrhs, hipMemcpyDeviceToHost, s));
  HWY_HIP_CHECK(hipStreamSynchronize(s));
  const double anorm = *std::max_element(rows.begin(), rows.end());
  for (int j = 0; j < nrhs; ++j) {
    double xn = 0.0;
    for (int i = 0; i < n; ++i) xn = std::max(xn, std::fabs(xs[size_t(j) * n + i]));
    if (nbe) nbe[j] = (anorm > 0 && xn > 0) ? rinf[j] / (anorm * xn) : 0.0;
    if (ferr) ferr[j] = xn != 0.0 ? est[j] / xn : est[j];
    if (iterations) iterations[j] = count[j] - 1;
  }
  for (void* p : {static_cast<void*>(rp), static_cast<void*>(wp), static_cast<void*>(r),
                  static_cast<void*>(stats), static_cast<void*>(rowsum),
                  static_cast<void*>(rowpart), static_cast<void*>(active_d),
                  static_cast<void*>(wts), static_cast<void*>(z)})
    if (p) HWY_HIP_CHECK(hipFree(p));
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
    Reject("DGETRF", -info);
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
    Reject("DGETRS", -info);
    return info;
  }
  if (n == 0 || nrhs == 0) return 0;
  lu::Getrs(op == Op::kTrans, n, nrhs, a, lda, ipiv, b, ldb, StreamOr(stream));
  return 0;
}

// Solves A X = B: af := LU of A (trailing updates in the precision
// `update`), X := getrs(B), then FP64 refinement against A. ferr (forward
// error bound, when not null), berr, nbe (normwise backward error
// ||B - A X||inf / (||A||inf ||X||inf)) and iterations (corrections
// applied) per right-hand side, on the host.
// Returns getrf's info (> 0: singular, nothing solved) or < 0 on an
// invalid argument.
HWYBLAS_AMD_API int hwyblas_dgesv_ir_dev(int n, int nrhs, const double* a, int lda, double* af,
                                         int ldaf, int* ipiv, const double* b, int ldb, double* x,
                                         int ldx, int update, int itmax, double* ferr,
                                         double* berr, double* nbe, int* iterations,
                                         void* stream) {
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
    Reject("DGESVIR", -info);
    return info;
  }
  if (n == 0 || nrhs == 0) {
    for (int j = 0; j < nrhs; ++j) {
      if (ferr) ferr[j] = 0.0;
      if (berr) berr[j] = 0.0;
      if (nbe) nbe[j] = 0.0;
      if (iterations) iterations[j] = 0;
    }
    return 0;
  }
  hipStream_t s = StreamOr(stream);
  HWY_HIP_CHECK(hwy_amd::Memcpy2DAsync(af, size_t(ldaf) * 8, a, size_t(lda) * 8, size_t(n) * 8, n,
                                       hipMemcpyDeviceToDevice, s));
  info = lu::Getrf(n, n, af, ldaf, ipiv, update, s);
  if (info != 0) return info;
  HWY_HIP_CHECK(hwy_amd::Memcpy2DAsync(x, size_t(ldx) * 8, b, size_t(ldb) * 8, size_t(n) * 8, nrhs,
                                       hipMemcpyDeviceToDevice, s));
  const lu::Perm perm(n, nrhs, ipiv, s);
  lu::GetrsPerm(false, n, nrhs, af, ldaf, perm, x, ldx, s);
  lu::Refine(false, n, nrhs, a, lda, af, ldaf, perm, b, ldb, x, ldx, itmax, ferr, berr, nbe,
             iterations, s);
  return 0;
}

// xGERFS (src/lapack/gesvx.cc:333-521) on device operands: refines X for
// op(A) X = B with the factors of hwyblas_dgetrf_dev (af, ipiv: device,
// 1-based), itmax 5, right-hand sides in groups of 16; ferr and berr are
// host arrays of nrhs. Returns 0 or -i for invalid argument i (reported as
// DGERFS, positions as xGERFS's). Synchronizes.
HWYBLAS_AMD_API int hwyblas_dgerfs_dev(char trans, int n, int nrhs, const double* a, int lda,
                                       const double* af, int ldaf, const int* ipiv,
                                       const double* b, int ldb, double* x, int ldx,
                                       double* ferr, double* berr, void* stream) {
  Op op;
  int info = 0;
  if (!DecodeOp(trans, &op)) info = -1;
  else if (n < 0) info = -2;
  else if (nrhs < 0) info = -3;
  else if (lda < std::max(1, n)) info = -5;
  else if (ldaf < std::max(1, n)) info = -7;
  else if (ldb < std::max(1, n)) info = -10;
  else if (ldx < std::max(1, n)) info = -12;
  if (info != 0) {
    Reject("DGERFS", -info);
    return info;
  }
  if (n == 0 || nrhs == 0) {
    for (int j = 0; j < nrhs; ++j) ferr[j] = berr[j] = 0.0;
    return 0;
  }
  hipStream_t s = StreamOr(stream);
  const int group = std::min(nrhs, lu::kMaxRhs);
  const lu::Perm perm(n, group, ipiv, s);
  for (int j0 = 0; j0 < nrhs; j0 += lu::kMaxRhs) {
    const int cols = std::min(lu::kMaxRhs, nrhs - j0);
    lu::Refine(op == Op::kTrans, n, cols, a, lda, af, ldaf, perm, b + int64_t(j0) * ldb, ldb,
               x + int64_t(j0) * ldx, ldx, 5, ferr + j0, berr + j0, nullptr, nullptr, s);
  }
  return 0;
}

}  // extern "C"
