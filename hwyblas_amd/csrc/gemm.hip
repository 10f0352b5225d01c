// Plain GEMM on the MI355X matrix cores: C = alpha op(A) op(B) + beta C in
// FP64 (v_mfma_f64_16x16x4_f64) or FP32 (v_mfma_f32_16x16x4_f32).
//
// hwyBLAS's Goto/BLIS structure (src/gemm-inl.h:153-232: pack panels of A and
// B, run an MR x NR register tile over a KC block) maps onto the GPU as:
//  - a workgroup owns a BM x BN macro-tile of C and walks the whole k range
//    (no KC blocks: the accumulators stay in registers, so beta is applied
//    once, in the epilogue);
//  - "packing" is the global -> LDS stage: each BK slice of op(A) and op(B)
//    is loaded with coalesced reads along whichever dimension is contiguous
//    in memory (the transposition is absorbed by the load, as hwyBLAS's
//    packing absorbs it, pack-inl.h:24-38) and written to LDS in a layout
//    the MFMA operand reads can take conflict-light; two LDS buffers let the
//    next slice load while the current one is multiplied;
//  - the register tile is a wave's WM x WN block of 16 x 16 MFMA tiles;
//  - the epilogue is hwyBLAS's Store (src/gemm_kernel-inl.h:181-224):
//    r = alpha * acc; beta == 1 adds C; other non-zero beta is fused as
//    fma(C, beta, r); beta == 0 never reads C. Edge tiles zero-fill
//    out-of-range operands and mask their stores, so rows >= m of C (the
//    padding up to ldc) are never touched.
// The source is compiled with -ffp-contract=off, so the epilogue rounds
// exactly as written.
#include "common.h"

namespace hwy_amd {
namespace {

typedef double dx4 __attribute__((ext_vector_type(4)));
typedef float fx4 __attribute__((ext_vector_type(4)));

// One 16 x 16 x 4 MFMA step. Operand layout (both types): lane l supplies
// A(l % 16, l / 16) and B(l / 16, l % 16). Result element r of lane l is
// column l % 16 and row Row(l / 16, r): rows interleave for FP64
// (l / 16 + 4 r) but are blocked for FP32 (4 (l / 16) + r), as measured on
// gfx950 with one-hot operands.
template <typename T>
struct Mfma;
template <>
struct Mfma<double> {
  using Acc = dx4;
  static __device__ __forceinline__ Acc Run(double a, double b, Acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int Row(int lk, int r) { return lk + 4 * r; }
};
template <>
struct Mfma<float> {
  using Acc = fx4;
  static __device__ __forceinline__ Acc Run(float a, float b, Acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int Row(int lk, int r) { return 4 * lk + r; }
};

template <typename T>
struct GemmParams {
  const T* a;
  const T* b;
  T* c;
  int64_t m, n, k, lda, ldb, ldc;
  T alpha, beta;
  int tiles_m, tiles_n;
};

constexpr int kXcds = 8;   // MI355X: 8 XCDs, each with its own L2
constexpr int kGroupM = 8;  // tile rows per raster group

// Maps the workgroup id to a (tile row, tile column) so that the workgroups
// one XCD receives (ids congruent mod 8, dispatched round-robin) form a
// contiguous run of the raster order, and the raster order walks groups of
// kGroupM tile rows column by column: tiles that share A row panels and B
// column panels then run close together in time on the same L2.
__device__ __forceinline__ void TileCoords(int tiles_m, int tiles_n, int* tm,
                                           int* tn) {
  const int nblk = gridDim.x, bid = blockIdx.x;
  const int xcd = bid % kXcds, idx = bid / kXcds;
  const int per = nblk / kXcds, rem = nblk % kXcds;
  const int lid = xcd < rem ? xcd * (per + 1) + idx
                            : rem * (per + 1) + (xcd - rem) * per + idx;
  const int group = kGroupM * tiles_n;
  const int first_m = (lid / group) * kGroupM;
  const int gm = min(tiles_m - first_m, kGroupM);
  const int in_group = lid % group;
  *tm = first_m + in_group % gm;
  *tn = in_group / gm;
}

// Macro-tile kernel. TA/TB: op(A)/op(B) transposed. The block is
// WAVES_M x WAVES_N waves; each wave computes (BM / WAVES_M) x (BN / WAVES_N)
// as 16 x 16 MFMA tiles.
template <typename T, bool TA, bool TB, int BM, int BN, int BK, int WAVES_M,
          int WAVES_N>
__global__ void __launch_bounds__(64 * WAVES_M * WAVES_N)
    GemmKernel(GemmParams<T> p) {
  using M = Mfma<T>;
  using Acc = typename M::Acc;
  constexpr int NT = 64 * WAVES_M * WAVES_N;
  constexpr int WM = BM / WAVES_M, WN = BN / WAVES_N;
  constexpr int TM = WM / 16, TN = WN / 16;
  static_assert(WM % 16 == 0 && WN % 16 == 0 && BK % 4 == 0, "tile shape");

  // LDS layouts. An operand whose memory is contiguous along k is kept
  // row-major with k fastest ([row][BK + 1], the odd stride spreads the MFMA
  // reads over the banks); one contiguous along rows/columns is kept k-major
  // ([k][BM + PAD]).
  constexpr int kPadMN = 4;
  constexpr int A_LD = TA ? BK + 1 : BM + kPadMN;
  constexpr int A_SZ = TA ? BM * A_LD : BK * A_LD;
  constexpr int B_LD = TB ? BN + kPadMN : BK + 1;
  constexpr int B_SZ = TB ? BK * B_LD : BN * B_LD;
  __shared__ T As[2][A_SZ];
  __shared__ T Bs[2][B_SZ];

  // Global-load assignment: consecutive threads read consecutive addresses.
  // A: !TA -> contiguous in i (rows of op(A)); TA -> contiguous in p.
  constexpr int A_FAST = TA ? BK : BM;          // extent of the contiguous dim
  constexpr int A_SLOW_STEP = NT / A_FAST;      // slow-dim stride per r
  constexpr int LA = (BM * BK) / NT;            // elements per thread
  constexpr int B_FAST = TB ? BN : BK;
  constexpr int B_SLOW_STEP = NT / B_FAST;
  constexpr int LB = (BN * BK) / NT;
  static_assert(NT % A_FAST == 0 && NT % B_FAST == 0, "load shape");
  static_assert(LA * NT == BM * BK && LB * NT == BN * BK, "load shape");

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave % WAVES_M) * WM, wn0 = (wave / WAVES_M) * WN;

  int tm, tn;
  TileCoords(p.tiles_m, p.tiles_n, &tm, &tn);
  const int64_t m0 = int64_t(tm) * BM, n0 = int64_t(tn) * BN;

  // Per-thread coordinates of the elements it loads (fast, slow + step * r).
  const int a_fast = tid % A_FAST, a_slow = tid / A_FAST;
  const int b_fast = tid % B_FAST, b_slow = tid / B_FAST;
  // A: (i, p) = TA ? (slow, fast) : (fast, slow).
  const int64_t a_i0 = m0 + (TA ? a_slow : a_fast);
  const int a_p0 = TA ? a_fast : a_slow;
  // B: (p, j) = TB ? (slow, fast) : (fast, slow).
  const int64_t b_j0 = n0 + (TB ? b_fast : b_slow);
  const int b_p0 = TB ? b_slow : b_fast;

  T ra[LA], rb[LB];

  auto load_a = [&](int64_t k0) {
#pragma unroll
    for (int r = 0; r < LA; ++r) {
      const int64_t i = TA ? a_i0 + int64_t(A_SLOW_STEP) * r : a_i0;
      const int64_t pp = k0 + (TA ? a_p0 : a_p0 + A_SLOW_STEP * r);
      const bool ok = (i < p.m) && (pp < p.k);
      const int64_t off = TA ? pp + i * p.lda : i + pp * p.lda;
      ra[r] = ok ? p.a[off] : T(0);
    }
  };
  auto load_b = [&](int64_t k0) {
#pragma unroll
    for (int r = 0; r < LB; ++r) {
      const int64_t j = TB ? b_j0 : b_j0 + int64_t(B_SLOW_STEP) * r;
      const int64_t pp = k0 + (TB ? b_p0 + B_SLOW_STEP * r : b_p0);
      const bool ok = (j < p.n) && (pp < p.k);
      const int64_t off = TB ? j + pp * p.ldb : pp + j * p.ldb;
      rb[r] = ok ? p.b[off] : T(0);
    }
  };
  auto store_lds = [&](int buf) {
#pragma unroll
    for (int r = 0; r < LA; ++r) {
      const int i = TA ? a_slow + A_SLOW_STEP * r : a_fast;
      const int pp = TA ? a_fast : a_slow + A_SLOW_STEP * r;
      As[buf][TA ? i * A_LD + pp : pp * A_LD + i] = ra[r];
    }
#pragma unroll
    for (int r = 0; r < LB; ++r) {
      const int j = TB ? b_fast : b_slow + B_SLOW_STEP * r;
      const int pp = TB ? b_slow + B_SLOW_STEP * r : b_fast;
      Bs[buf][TB ? pp * B_LD + j : j * B_LD + pp] = rb[r];
    }
  };

  Acc acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = Acc{};

  const int ktiles = int((p.k + BK - 1) / BK);
  load_a(0);
  load_b(0);
  store_lds(0);
  __syncthreads();

  const int li = lane & 15, lk = lane >> 4;
  for (int kt = 0; kt < ktiles; ++kt) {
    const int cur = kt & 1;
    const bool more = kt + 1 < ktiles;
    if (more) {
      load_a(int64_t(kt + 1) * BK);
      load_b(int64_t(kt + 1) * BK);
    }
    const T* as = As[cur];
    const T* bs = Bs[cur];
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T af[TM], bf[TN];
#pragma unroll
      for (int ti = 0; ti < TM; ++ti) {
        const int i = wm0 + ti * 16 + li, pp = kk + lk;
        af[ti] = as[TA ? i * A_LD + pp : pp * A_LD + i];
      }
#pragma unroll
      for (int tj = 0; tj < TN; ++tj) {
        const int j = wn0 + tj * 16 + li, pp = kk + lk;
        bf[tj] = bs[TB ? pp * B_LD + j : j * B_LD + pp];
      }
#pragma unroll
      for (int ti = 0; ti < TM; ++ti)
#pragma unroll
        for (int tj = 0; tj < TN; ++tj) acc[ti][tj] = M::Run(af[ti], bf[tj], acc[ti][tj]);
    }
    if (more) store_lds(cur ^ 1);
    __syncthreads();
  }

  // Epilogue (hwyBLAS Store semantics).
  const T alpha = p.alpha, beta = p.beta;
  const bool beta_zero = beta == T(0), beta_one = beta == T(1);
#pragma unroll
  for (int ti = 0; ti < TM; ++ti) {
#pragma unroll
    for (int tj = 0; tj < TN; ++tj) {
      const int64_t col = n0 + wn0 + tj * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = m0 + wm0 + ti * 16 + M::Row(lk, r);
        if (row < p.m && col < p.n) {
          T* cij = p.c + row + col * p.ldc;
          T v = acc[ti][tj][r] * alpha;
          if (!beta_zero) v = beta_one ? v + *cij : __builtin_fma(*cij, beta, v);
          *cij = v;
        }
      }
    }
  }
}

template <typename T>
__global__ void ScaleKernel(int64_t m, int64_t n, T beta, T* c, int64_t ldc) {
  const int64_t total = m * n;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
       idx += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = idx % m, j = idx / m;
    T* x = c + i + j * ldc;
    *x = (beta == T(0)) ? T(0) : *x * beta;
  }
}

template <typename T, bool TA, bool TB, int BM, int BN, int BK, int WGM, int WGN>
void Launch(const GemmParams<T>& prm, hipStream_t stream) {
  GemmParams<T> q = prm;
  q.tiles_m = int((q.m + BM - 1) / BM);
  q.tiles_n = int((q.n + BN - 1) / BN);
  const int64_t blocks = int64_t(q.tiles_m) * q.tiles_n;
  hipLaunchKernelGGL((GemmKernel<T, TA, TB, BM, BN, BK, WGM, WGN>), dim3(unsigned(blocks)),
                     dim3(64 * WGM * WGN), 0, stream, q);
  HWY_HIP_CHECK(hipGetLastError());
}

template <typename T, bool TA, bool TB>
void LaunchTrans(const GemmParams<T>& prm, hipStream_t stream) {
  // Large problems: 128 x 128 macro-tiles of 4 waves (64 x 64 each). Fewer
  // tiles than about two per CU: 64 x 64 macro-tiles (32 x 32 per wave).
  const int64_t big_tiles = ((prm.m + 127) / 128) * ((prm.n + 127) / 128);
  if (big_tiles >= 512) {
    Launch<T, TA, TB, 128, 128, 16, 2, 2>(prm, stream);
  } else {
    Launch<T, TA, TB, 64, 64, 16, 2, 2>(prm, stream);
  }
}

}  // namespace

template <typename T>
void GemmDevice(bool ta, bool tb, int64_t m, int64_t n, int64_t k, T alpha,
                const T* a, int64_t lda, const T* b, int64_t ldb, T beta, T* c,
                int64_t ldc, hipStream_t stream) {
  GemmParams<T> prm{a, b, c, m, n, k, lda, ldb, ldc, alpha, beta, 0, 0};
  if (!ta && !tb) LaunchTrans<T, false, false>(prm, stream);
  else if (!ta && tb) LaunchTrans<T, false, true>(prm, stream);
  else if (ta && !tb) LaunchTrans<T, true, false>(prm, stream);
  else LaunchTrans<T, true, true>(prm, stream);
}

template <typename T>
void ScaleDevice(int64_t m, int64_t n, T beta, T* c, int64_t ldc,
                 hipStream_t stream) {
  if (m <= 0 || n <= 0) return;
  const int64_t total = m * n;
  const int64_t blocks = std::min<int64_t>((total + 255) / 256, 65536);
  hipLaunchKernelGGL(ScaleKernel<T>, dim3(unsigned(blocks)), dim3(256), 0, stream,
                     m, n, beta, c, ldc);
  HWY_HIP_CHECK(hipGetLastError());
}

template void GemmDevice<double>(bool, bool, int64_t, int64_t, int64_t, double,
                                 const double*, int64_t, const double*, int64_t,
                                 double, double*, int64_t, hipStream_t);
template void GemmDevice<float>(bool, bool, int64_t, int64_t, int64_t, float,
                                const float*, int64_t, const float*, int64_t,
                                float, float*, int64_t, hipStream_t);
template void ScaleDevice<double>(int64_t, int64_t, double, double*, int64_t,
                                  hipStream_t);
template void ScaleDevice<float>(int64_t, int64_t, float, float*, int64_t,
                                 hipStream_t);

}  // namespace hwy_amd
