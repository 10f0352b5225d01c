// The plain MFMA GEMM kernel template (FP64 / FP32), shared by the
// per-precision instantiation files.
//
// hwyBLAS's Goto/BLIS structure (src/gemm-inl.h:153-232: pack panels of A and
// B, run an MR x NR register tile over a KC block) maps onto the GPU as:
//  - a workgroup owns a BM x BN macro-tile of C and walks the whole k range
//    (no KC blocks: the accumulators stay in registers, so beta is applied
//    once, in the epilogue);
//  - "packing" is the global -> LDS stage: each BK slice of op(A) and op(B)
//    is loaded with coalesced reads along whichever dimension is contiguous
//    in memory (the transposition is absorbed by the load, as hwyBLAS's
//    packing absorbs it, src/pack-inl.h:24-38) and written to LDS in a layout
//    the MFMA operand reads take with few bank conflicts; two LDS buffers let
//    the next slice load while the current one is multiplied;
//  - the register tile is a wave's WM x WN block of 16 x 16 MFMA tiles;
//  - the epilogue is hwyBLAS's Store (src/gemm_kernel-inl.h:181-224):
//    r = alpha * acc; beta == 1 adds C; other non-zero beta is fused as
//    fma(C, beta, r); beta == 0 never reads C. Edge tiles zero-fill
//    out-of-range operands and mask their stores, so rows >= m of C (the
//    padding up to ldc) are never touched.
// Sources are compiled with -ffp-contract=off, so the epilogue rounds
// exactly as written.
#ifndef HWYBLAS_AMD_CSRC_GEMM_KERNEL_H_
#define HWYBLAS_AMD_CSRC_GEMM_KERNEL_H_

#include <type_traits>

#include "common.h"
#include "gemm_configs.h"

namespace hwy_amd {
namespace gemm_detail {

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
  // The part of C computed and written: 0 all of it, 1 the lower triangle
  // (row >= col), 2 the upper triangle (row <= col). Tiles entirely outside
  // the part do no work.
  int part = 0;
  // Set by Launch for a triangle of a square C with square tiles: the grid
  // holds only the T (T + 1) / 2 tiles that meet the triangle.
  int tri_grid = 0;
  // A triangular factor's zero k-range, skipped per tile (only the k slices
  // meeting [kb, ke) run): 0 none; 1 kb = the tile's first row; 2 ke = its
  // last row + 1; 3 ke = its last column + 1; 4 kb = its first column.
  int krange = 0;
  // Non-null: the kernel does its work only if *gate == 0 (a decision made
  // on the device earlier on the stream).
  const int* gate = nullptr;
};

constexpr int kXcds = 8;    // MI355X: 8 XCDs, each with its own L2
constexpr int kGroupM = 8;  // tile rows per raster group

// Maps the workgroup id to a (tile row, tile column) so that the workgroups
// one XCD receives (ids congruent mod 8, dispatched round-robin) form a
// contiguous run of the raster order, and the raster order walks groups of
// kGroupM tile rows column by column: tiles that share A row panels and B
// column panels then run close together in time on the same L2.
// The workgroup's position in a run that gives each XCD (ids congruent mod
// 8, dispatched round-robin) a contiguous share.
__device__ __forceinline__ int XcdRunIndex() {
  const int nblk = gridDim.x, bid = blockIdx.x;
  const int xcd = bid % kXcds, idx = bid / kXcds;
  const int per = nblk / kXcds, rem = nblk % kXcds;
  return xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
}

__device__ __forceinline__ void TileCoords(int tiles_m, int tiles_n, int* tm,
                                           int* tn) {
  const int lid = XcdRunIndex();
  const int group = kGroupM * tiles_n;
  const int first_m = (lid / group) * kGroupM;
  const int gm = min(tiles_m - first_m, kGroupM);
  const int in_group = lid % group;
  *tm = first_m + in_group % gm;
  *tn = in_group / gm;
}

// LDS bytes one macro-tile needs (two buffers of each operand's slice).
// An operand whose memory is contiguous along k is kept row-major with k
// fastest ([row][BK + 1], the odd stride spreads the MFMA reads over the
// banks); one contiguous along rows/columns is kept k-major ([k][BM + PAD]).
template <typename T, bool TA, bool TB, int BM, int BN, int BK>
struct TileLds {
  static constexpr int kPadMN = 4;
  static constexpr int A_LD = TA ? BK + 1 : BM + kPadMN;
  static constexpr int A_SZ = TA ? BM * A_LD : BK * A_LD;
  static constexpr int B_LD = TB ? BN + kPadMN : BK + 1;
  static constexpr int B_SZ = TB ? BK * B_LD : BN * B_LD;
  static constexpr int kElems = 2 * (A_SZ + B_SZ);
};

// One workgroup (WAVES_M x WAVES_N waves) computes the BM x BN macro-tile of
// C at (m0, n0) over the whole k range; each wave computes
// (BM / WAVES_M) x (BN / WAVES_N) as 16 x 16 MFMA tiles. `lds` holds
// TileLds<...>::kElems elements. Every thread of the workgroup must call it.
template <typename T, bool TA, bool TB, int BM, int BN, int BK, int WAVES_M, int WAVES_N>
__device__ __forceinline__ void MacroTile(const GemmParams<T>& p, int64_t m0, int64_t n0,
                                          T* lds) {
  using M = Mfma<T>;
  using Acc = typename M::Acc;
  using Lds = TileLds<T, TA, TB, BM, BN, BK>;
  constexpr int NT = 64 * WAVES_M * WAVES_N;
  constexpr int WM = BM / WAVES_M, WN = BN / WAVES_N;
  constexpr int TM = WM / 16, TN = WN / 16;
  static_assert(WM % 16 == 0 && WN % 16 == 0 && BK % 4 == 0, "tile shape");
  constexpr int A_LD = Lds::A_LD, A_SZ = Lds::A_SZ;
  constexpr int B_LD = Lds::B_LD, B_SZ = Lds::B_SZ;
  // Buffer b of op(A)'s slice starts at lds + b * A_SZ, of op(B)'s at
  // lds + 2 * A_SZ + b * B_SZ.
  T* const as_base = lds;
  T* const bs_base = lds + 2 * A_SZ;

  // Global-load assignment: consecutive threads read consecutive addresses.
  constexpr int A_FAST = TA ? BK : BM;      // extent of the contiguous dim
  constexpr int A_SLOW_STEP = NT / A_FAST;  // slow-dim stride per element
  constexpr int LA = (BM * BK) / NT;        // elements per thread
  constexpr int B_FAST = TB ? BN : BK;
  constexpr int B_SLOW_STEP = NT / B_FAST;
  constexpr int LB = (BN * BK) / NT;
  static_assert(NT % A_FAST == 0 && NT % B_FAST == 0, "load shape");
  static_assert(LA * NT == BM * BK && LB * NT == BN * BK, "load shape");

  const int tid = threadIdx.x;
  const int lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave % WAVES_M) * WM, wn0 = (wave / WAVES_M) * WN;

  // Per-thread coordinates of the elements it loads: (fast, slow + step r).
  const int a_fast = tid % A_FAST, a_slow = tid / A_FAST;
  const int b_fast = tid % B_FAST, b_slow = tid / B_FAST;
  // A: (i, p) = TA ? (slow, fast) : (fast, slow); B: (p, j) = TB ? (slow,
  // fast) : (fast, slow).
  const int64_t a_i0 = m0 + (TA ? a_slow : a_fast);
  const int a_p0 = TA ? a_fast : a_slow;
  const int64_t b_j0 = n0 + (TB ? b_fast : b_slow);
  const int b_p0 = TB ? b_slow : b_fast;
  // Element r of this thread's A slice at k offset k0 is
  // a_base[k0 * (TA ? 1 : lda) + r * a_rstep].
  const T* a_base = p.a + (TA ? a_p0 + a_i0 * p.lda : a_i0 + int64_t(a_p0) * p.lda);
  const int64_t a_rstep = TA ? int64_t(A_SLOW_STEP) * p.lda : int64_t(A_SLOW_STEP) * p.lda;
  const int64_t a_kstep = TA ? 1 : p.lda;
  const T* b_base = p.b + (TB ? b_j0 + int64_t(b_p0) * p.ldb : b_p0 + b_j0 * p.ldb);
  const int64_t b_rstep = int64_t(B_SLOW_STEP) * p.ldb;
  const int64_t b_kstep = TB ? p.ldb : 1;
  // The tile lies inside op(A)'s rows and op(B)'s columns: only the k tail
  // needs checks.
  const bool inside = (m0 + BM <= p.m) && (n0 + BN <= p.n);

  T ra[LA], rb[LB];

  auto load = [&](int64_t k0, auto checked) {
    constexpr bool kCheck = decltype(checked)::value;
#pragma unroll
    for (int r = 0; r < LA; ++r) {
      const T* src = a_base + k0 * a_kstep + r * a_rstep;
      if (kCheck) {
        const int64_t i = TA ? a_i0 + int64_t(A_SLOW_STEP) * r : a_i0;
        const int64_t pp = k0 + (TA ? a_p0 : a_p0 + A_SLOW_STEP * r);
        ra[r] = (i < p.m && pp < p.k) ? *src : T(0);
      } else {
        ra[r] = *src;
      }
    }
#pragma unroll
    for (int r = 0; r < LB; ++r) {
      const T* src = b_base + k0 * b_kstep + r * b_rstep;
      if (kCheck) {
        const int64_t j = TB ? b_j0 : b_j0 + int64_t(B_SLOW_STEP) * r;
        const int64_t pp = k0 + (TB ? b_p0 + B_SLOW_STEP * r : b_p0);
        rb[r] = (j < p.n && pp < p.k) ? *src : T(0);
      } else {
        rb[r] = *src;
      }
    }
  };
  auto load_tile = [&](int64_t k0) {
    if (inside && k0 + BK <= p.k) load(k0, std::false_type{});
    else load(k0, std::true_type{});
  };
  auto store_lds = [&](int buf) {
#pragma unroll
    for (int r = 0; r < LA; ++r) {
      const int i = TA ? a_slow + A_SLOW_STEP * r : a_fast;
      const int pp = TA ? a_fast : a_slow + A_SLOW_STEP * r;
      as_base[buf * A_SZ + (TA ? i * A_LD + pp : pp * A_LD + i)] = ra[r];
    }
#pragma unroll
    for (int r = 0; r < LB; ++r) {
      const int j = TB ? b_fast : b_slow + B_SLOW_STEP * r;
      const int pp = TB ? b_slow + B_SLOW_STEP * r : b_fast;
      bs_base[buf * B_SZ + (TB ? pp * B_LD + j : j * B_LD + pp)] = rb[r];
    }
  };

  Acc acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = Acc{};

  int kt0 = 0, kt1 = int((p.k + BK - 1) / BK);
  if (p.krange) {
    int64_t kb = 0, ke = p.k;
    if (p.krange == 1) kb = m0;
    else if (p.krange == 2) ke = m0 + BM;
    else if (p.krange == 3) ke = n0 + BN;
    else kb = n0;
    kt0 = int(min(kb, p.k) / BK);
    kt1 = int((min(ke, p.k) + BK - 1) / BK);
  }
  load_tile(int64_t(kt0) * BK);
  store_lds(0);
  __syncthreads();

  const int li = lane & 15, lk = lane >> 4;
  for (int kt = kt0; kt < kt1; ++kt) {
    const int cur = (kt - kt0) & 1;
    const bool more = kt + 1 < kt1;
    if (more) load_tile(int64_t(kt + 1) * BK);
    const T* as = as_base + cur * A_SZ;
    const T* bs = bs_base + cur * B_SZ;
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
        const bool in_part = p.part == 0 || (p.part == 1 ? row >= col : row <= col);
        if (row < p.m && col < p.n && in_part) {
          T* cij = p.c + row + col * p.ldc;
          T v = acc[ti][tj][r] * alpha;
          if (!beta_zero) v = beta_one ? v + *cij : __builtin_fma(*cij, beta, v);
          *cij = v;
        }
      }
    }
  }
}


// Macro-tile kernel: one workgroup per tile, tiles mapped XCD-aware.
template <typename T, bool TA, bool TB, int BM, int BN, int BK, int WAVES_M,
          int WAVES_N, int OCC>
__global__ void __launch_bounds__(64 * WAVES_M * WAVES_N)
    __attribute__((amdgpu_waves_per_eu(OCC))) GemmKernel(GemmParams<T> p) {
  __shared__ T lds[TileLds<T, TA, TB, BM, BN, BK>::kElems];
  if (p.gate && *p.gate != 0) return;  // uniform
  int tm, tn;
  if (p.tri_grid) {
    // Tile t of the triangle, row by row: r (r + 1) / 2 <= t < (r + 1) (r + 2) / 2.
    const int t = XcdRunIndex();
    int r = int((sqrtf(8.0f * float(t) + 1.0f) - 1.0f) * 0.5f);
    while ((r + 1) * (r + 2) / 2 <= t) ++r;
    while (r * (r + 1) / 2 > t) --r;
    const int c = t - r * (r + 1) / 2;
    tm = p.part == 1 ? r : c;
    tn = p.part == 1 ? c : r;
  } else if (p.krange == 1 || p.krange == 2) {
    // Work that varies by tile row: the raster bands run over tile columns,
    // so every XCD's share spans all rows.
    TileCoords(p.tiles_n, p.tiles_m, &tn, &tm);
  } else {
    TileCoords(p.tiles_m, p.tiles_n, &tm, &tn);
  }
  const int64_t m0 = int64_t(tm) * BM, n0 = int64_t(tn) * BN;
  // A tile entirely on the other side of the diagonal (uniform exit).
  if ((p.part == 1 && m0 + BM - 1 < n0) || (p.part == 2 && m0 > n0 + BN - 1)) return;
  MacroTile<T, TA, TB, BM, BN, BK, WAVES_M, WAVES_N>(p, m0, n0, lds);
}

template <typename T, bool TA, bool TB, int BM, int BN, int BK, int WGM, int WGN, int OCC>
void Launch(const GemmParams<T>& prm, hipStream_t stream) {
  GemmParams<T> q = prm;
  q.tiles_m = int((q.m + BM - 1) / BM);
  q.tiles_n = int((q.n + BN - 1) / BN);
  int64_t blocks = int64_t(q.tiles_m) * q.tiles_n;
  q.tri_grid = 0;
  if (q.part != 0 && BM == BN && q.m == q.n) {
    q.tri_grid = 1;
    blocks = int64_t(q.tiles_m) * (q.tiles_m + 1) / 2;
  }
  hipLaunchKernelGGL((GemmKernel<T, TA, TB, BM, BN, BK, WGM, WGN, OCC>), dim3(unsigned(blocks)),
                     dim3(64 * WGM * WGN), 0, stream, q);
  HWY_HIP_CHECK(hipGetLastError());
}

// Runs configuration `cfg` (an index into kGemmConfigs).
template <typename T, bool TA, bool TB>
void LaunchConfig(int cfg, const GemmParams<T>& prm, hipStream_t stream) {
  switch (cfg) {
#define HWY_GEMM_CASE(ID, BM, BN, BK, WM, WN, OCC) \
  case ID:                                         \
    Launch<T, TA, TB, BM, BN, BK, WM, WN, OCC>(prm, stream); break;
    HWY_GEMM_CONFIGS(HWY_GEMM_CASE)
#undef HWY_GEMM_CASE
    default:
      Launch<T, TA, TB, 64, 64, 16, 2, 2, 1>(prm, stream);
  }
}

template <typename T>
void LaunchAny(int cfg, bool ta, bool tb, const GemmParams<T>& prm, hipStream_t stream) {
  if (!ta && !tb) LaunchConfig<T, false, false>(cfg, prm, stream);
  else if (!ta && tb) LaunchConfig<T, false, true>(cfg, prm, stream);
  else if (ta && !tb) LaunchConfig<T, true, false>(cfg, prm, stream);
  else LaunchConfig<T, true, true>(cfg, prm, stream);
}

}  // namespace gemm_detail
}  // namespace hwy_amd

#endif  // HWYBLAS_AMD_CSRC_GEMM_KERNEL_H_
