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
#include "dma.h"
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
  // Split k (register-staged kernel only): grid row y computes the k range
  // [y ksplit, (y + 1) ksplit) into C + y c_split (alpha, beta and part
  // apply to each; 0: no split).
  int64_t ksplit = 0, c_split = 0;
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

// The workgroup's output tile (tm, tn): the triangle's tiles row by row when
// the grid holds only those (tri_grid), else the XCD-aware raster, its bands
// over tile columns when the work varies by tile row (krange 1, 2).
template <typename T>
__device__ __forceinline__ void TileOf(const GemmParams<T>& p, int* tm, int* tn) {
  if (p.tri_grid) {
    // Tile t of the triangle, row by row: r (r + 1) / 2 <= t < (r + 1) (r + 2) / 2.
    const int t = XcdRunIndex();
    int r = int((sqrtf(8.0f * float(t) + 1.0f) - 1.0f) * 0.5f);
    while ((r + 1) * (r + 2) / 2 <= t) ++r;
    while (r * (r + 1) / 2 > t) --r;
    const int c = t - r * (r + 1) / 2;
    *tm = p.part == 1 ? r : c;
    *tn = p.part == 1 ? c : r;
  } else if (p.krange == 1 || p.krange == 2) {
    // Work that varies by tile row: the raster bands run over tile columns,
    // so every XCD's share spans all rows.
    TileCoords(p.tiles_n, p.tiles_m, tn, tm);
  } else {
    TileCoords(p.tiles_m, p.tiles_n, tm, tn);
  }
}

// The k slices [kt0, kt1) of BK a tile at (m0, n0) runs: all of k, or those
// meeting a triangular factor's nonzero k-range (GemmParams::krange).
template <typename T, int BM, int BN, int BK>
__device__ __forceinline__ void KSlices(const GemmParams<T>& p, int64_t m0, int64_t n0, int* kt0,
                                        int* kt1) {
  *kt0 = 0;
  *kt1 = int((p.k + BK - 1) / BK);
  if (p.krange) {
    int64_t kb = 0, ke = p.k;
    if (p.krange == 1) kb = m0;
    else if (p.krange == 2) ke = m0 + BM;
    else if (p.krange == 3) ke = n0 + BN;
    else kb = n0;
    *kt0 = int(min(kb, p.k) / BK);
    *kt1 = int((min(ke, p.k) + BK - 1) / BK);
  }
}

// LDS bytes one macro-tile needs (two buffers of each operand's slice).
// An operand whose memory is contiguous along k is kept row-major with k
// fastest ([row][K_LD]); one contiguous along rows/columns is kept k-major
// ([k][MN + PAD]). An MFMA operand read takes 16 consecutive rows/columns
// (lane % 16) at 4 consecutive k (lane / 16), so the strides are chosen for
// those 64 addresses to fall in distinct banks (64 banks of 4 bytes):
//  - FP32, k-contiguous: K_LD = BK + 4 with K_LD / 4 odd, so row r sits
//    4 (r K_LD / 4 mod 16) banks on and the four k of a row fill the gap;
//  - FP32, k-major: MN + 16, so each k row starts 16 banks after the last;
//  - FP64 (8-byte values, a read served a half-wave at a time): BK + 1 and
//    MN + 4, measured free of conflicts.
template <typename T, bool TA, bool TB, int BM, int BN, int BK>
struct TileLds {
  static constexpr bool kF32 = sizeof(T) == 4;
  static constexpr int kKLd = kF32 ? BK + 4 : BK + 1;
  static constexpr int kPadMN = kF32 ? 16 : 4;
  static_assert(!kF32 || (kKLd / 4) % 2 == 1, "k-contiguous stride");
  static constexpr int A_LD = TA ? kKLd : BM + kPadMN;
  static constexpr int A_SZ = TA ? BM * A_LD : BK * A_LD;
  static constexpr int B_LD = TB ? BN + kPadMN : kKLd;
  static constexpr int B_SZ = TB ? BK * B_LD : BN * B_LD;
  static constexpr int kElems = 2 * (A_SZ + B_SZ);
};

// The epilogue (hwyBLAS Store semantics) of a wave's TM x TN block of 16 x 16
// MFMA tiles whose first row and column of C are (r0, c0): r = alpha * acc;
// beta == 1 adds C; other non-zero beta is fused as fma(C, beta, r); beta == 0
// never reads C. Entries outside m x n or the requested triangle are skipped.
// One row of MFMA tiles at a time, all of its C loads issued before the first
// is used (a skipped entry loads C[0], which exists whenever there is a tile,
// and ignores it): a row costs one memory round trip instead of one per value.
// The compiler cannot move a load of C above a store to C by itself (it cannot
// prove the addresses differ), and a per-value branch on the beta form leaves
// it waiting for every store before the next value: the body is straight-line,
// beta == 0 in a body of its own, the two forms with C as a select.
template <bool READ_C, typename T, int TM, int TN>
__device__ __forceinline__ void StoreTileForm(const GemmParams<T>& p,
                                              const typename Mfma<T>::Acc (&acc)[TM][TN],
                                              int64_t r0, int64_t c0, int lane) {
  using M = Mfma<T>;
  const int li = lane & 15, lk = lane >> 4;
  const T alpha = p.alpha, beta = p.beta;
  const bool beta_one = beta == T(1);
#pragma unroll
  for (int ti = 0; ti < TM; ++ti) {
    bool ok[TN][4];
    T cv[TN][4];
#pragma unroll
    for (int tj = 0; tj < TN; ++tj) {
      const int64_t col = c0 + tj * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = r0 + ti * 16 + M::Row(lk, r);
        const bool in_part = p.part == 0 || (p.part == 1 ? row >= col : row <= col);
        ok[tj][r] = row < p.m && col < p.n && in_part;
        if (READ_C) cv[tj][r] = p.c[ok[tj][r] ? row + col * p.ldc : 0];
      }
    }
#pragma unroll
    for (int tj = 0; tj < TN; ++tj) {
      const int64_t col = c0 + tj * 16 + li;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int64_t row = r0 + ti * 16 + M::Row(lk, r);
        T v = acc[ti][tj][r] * alpha;
        if (READ_C) {
          const T sum = v + cv[tj][r], fused = __builtin_fma(cv[tj][r], beta, v);
          v = beta_one ? sum : fused;
        }
        if (ok[tj][r]) p.c[row + col * p.ldc] = v;
      }
    }
  }
}

template <typename T, int TM, int TN>
__device__ __forceinline__ void StoreTile(const GemmParams<T>& p,
                                          const typename Mfma<T>::Acc (&acc)[TM][TN], int64_t r0,
                                          int64_t c0, int lane) {
  if (p.beta == T(0)) StoreTileForm<false>(p, acc, r0, c0, lane);
  else StoreTileForm<true>(p, acc, r0, c0, lane);
}

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

  int kt0, kt1;
  KSlices<T, BM, BN, BK>(p, m0, n0, &kt0, &kt1);
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

  StoreTile<T, TM, TN>(p, acc, m0 + wm0, n0 + wn0, lane);
}

// Macro-tile kernel: one workgroup per tile, tiles mapped XCD-aware.
template <typename T, bool TA, bool TB, int BM, int BN, int BK, int WAVES_M,
          int WAVES_N, int OCC>
__global__ void __launch_bounds__(64 * WAVES_M * WAVES_N)
    __attribute__((amdgpu_waves_per_eu(OCC))) GemmKernel(GemmParams<T> p) {
  __shared__ T lds[TileLds<T, TA, TB, BM, BN, BK>::kElems];
  if (p.gate && *p.gate != 0) return;  // uniform
  int tm, tn;
  TileOf(p, &tm, &tn);
  const int64_t m0 = int64_t(tm) * BM, n0 = int64_t(tn) * BN;
  // A tile entirely on the other side of the diagonal (uniform exit).
  if ((p.part == 1 && m0 + BM - 1 < n0) || (p.part == 2 && m0 > n0 + BN - 1)) return;
  if (p.ksplit) {
    GemmParams<T> q = p;
    const int64_t k0 = int64_t(blockIdx.y) * p.ksplit;
    q.k = min(p.ksplit, p.k - k0);
    q.a += TA ? k0 : k0 * p.lda;
    q.b += TB ? k0 * p.ldb : k0;
    q.c += int64_t(blockIdx.y) * p.c_split;
    q.ksplit = 0;
    MacroTile<T, TA, TB, BM, BN, BK, WAVES_M, WAVES_N>(q, m0, n0, lds);
    return;
  }
  MacroTile<T, TA, TB, BM, BN, BK, WAVES_M, WAVES_N>(p, m0, n0, lds);
}

// ---------------------------------------------------------------------------
// The DMA-fed kernel. One operand's BK-deep slice in LDS as 16-byte chunks:
// KC (contiguous along k in memory: op(A) with A transposed, op(B) with B
// not transposed) holds MN rows of BK values, the other layout BK rows of MN
// values. A global_load_lds_dwordx4 writes 64 consecutive chunks, so the
// layout is fixed by which global chunk each lane fetches: chunk c of a row
// is stored at slot c ^ Swz(row), chosen so that the MFMA operand reads
// (16 consecutive rows or columns, lane % 16, at 4 consecutive k, lane / 16)
// meet distinct banks:
//  - KC: a read takes one chunk from each of 16 rows; RPB = 16 / (chunks per
//    row) rows share a 256-byte bank line, and Swz = (row / RPB) mod chunks
//    per row spreads the 16 over all 16 chunk slots of the line;
//  - otherwise: a read takes 16 consecutive values (16 sizeof(T) bytes) from
//    each of 4 k rows; the rows served together (2 for FP64, whose 8-byte
//    reads go a half-wave at a time, 4 for FP32) are moved apart by
//    Swz = (row mod rows together) * sizeof(T) chunks.
template <typename T, bool KC, int MN, int BK>
struct DmaOperand {
  static constexpr int E = 16 / int(sizeof(T));  // values per chunk
  static constexpr int kRowVals = KC ? BK : MN;
  static constexpr int kRows = KC ? MN : BK;
  static constexpr int kCpr = kRowVals / E;  // chunks per row
  static constexpr int kChunks = kRows * kCpr;
  static_assert(kRowVals % E == 0, "rows of whole chunks");
  static_assert(KC ? (kCpr >= 2 && kCpr <= 16 && 16 % kCpr == 0) : kCpr >= 16, "bank layout");
  static __device__ __forceinline__ int Swz(int row) {
    if (KC) return (row / (16 / kCpr)) % kCpr;
    constexpr int together = 256 / (16 * int(sizeof(T)));
    return (row % together) * int(sizeof(T));
  }
  // Index (in values) of op(X)'s element at (mn index x, k index y).
  static __device__ __forceinline__ int At(int x, int y) {
    const int row = KC ? x : y, v = KC ? y : x;
    return row * kRowVals + (((v / E) ^ Swz(row)) * E) + v % E;
  }
  // Offset (in values, from the tile's first row/column at k = 0) of the
  // global chunk that belongs in LDS slot `slot`.
  static __device__ __forceinline__ int64_t Source(int slot, int64_t ld) {
    const int row = slot / kCpr, c = (slot % kCpr) ^ Swz(row);
    return int64_t(row) * ld + c * E;
  }
};

typedef int i32x4_gemm __attribute__((ext_vector_type(4)));

// Full BM x BN tiles only (m, n multiples of the tile, k of BK; the launcher
// guarantees it), the whole k range per workgroup, the epilogue StoreTile.
// Each iteration: wait for this thread's DMAs of slice kt, barrier (every
// wave's DMAs of slice kt have landed and every wave is done reading the
// slot of slice kt - 1), multiply slice kt, refilling that slot with slice
// kt + STAGES - 1 (A's chunks after the first MFMA step, B's after the
// second).
template <typename T, bool TA, bool TB, int BM, int BN, int BK, int WAVES_M, int WAVES_N,
          int OCC, int STAGES>
__global__ void __launch_bounds__(64 * WAVES_M * WAVES_N)
    __attribute__((amdgpu_waves_per_eu(OCC))) GemmDmaKernel(GemmParams<T> prm) {
  using M = Mfma<T>;
  using Acc = typename M::Acc;
  using OA = DmaOperand<T, TA, BM, BK>;
  using OB = DmaOperand<T, !TB, BN, BK>;
  constexpr int NW = WAVES_M * WAVES_N;
  constexpr int WM = BM / WAVES_M, WN = BN / WAVES_N;
  constexpr int TM = WM / 16, TN = WN / 16;
  constexpr int A_DMA = OA::kChunks / (64 * NW), B_DMA = OB::kChunks / (64 * NW);
  static_assert(A_DMA * 64 * NW == OA::kChunks && B_DMA * 64 * NW == OB::kChunks,
                "whole DMA instructions per wave");
  constexpr int PER_SLICE = A_DMA + B_DMA;
  static_assert(PER_SLICE * 3 < 64 && STAGES >= 2 && STAGES <= 4, "vmcnt table");
  static_assert(WM % 16 == 0 && WN % 16 == 0 && BK % 8 == 0, "tile shape");
  constexpr int kStage = OA::kChunks + OB::kChunks;  // chunks per stage
  __shared__ i32x4_gemm lds[STAGES * kStage];
  if (prm.gate && *prm.gate != 0) return;  // uniform
  // Split k: grid row y computes the k range [y ksplit, (y + 1) ksplit) into
  // C + y c_split (the launcher guarantees whole slices in every range).
  GemmParams<T> p = prm;
  if (p.ksplit) {
    const int64_t k0 = int64_t(blockIdx.y) * p.ksplit;
    p.k = min(p.ksplit, p.k - k0);
    p.a += TA ? k0 : k0 * p.lda;
    p.b += TB ? k0 * p.ldb : k0;
    p.c += int64_t(blockIdx.y) * p.c_split;
  }

  int tm, tn;
  TileOf(p, &tm, &tn);
  const int64_t m0 = int64_t(tm) * BM, n0 = int64_t(tn) * BN;
  // A tile entirely on the other side of the diagonal (uniform exit).
  if ((p.part == 1 && m0 + BM - 1 < n0) || (p.part == 2 && m0 > n0 + BN - 1)) return;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave % WAVES_M) * WM, wn0 = (wave / WAVES_M) * WN;
  // op(A)'s tile: rows m0.. (KC: A's columns m0.., k along them).
  const T* A = p.a + (TA ? m0 * p.lda : m0);
  const T* B = p.b + (TB ? n0 : n0 * p.ldb);
  const int64_t a_kstep = TA ? 1 : p.lda, b_kstep = TB ? p.ldb : 1;
  int64_t a_src[A_DMA], b_src[B_DMA];
#pragma unroll
  for (int q = 0; q < A_DMA; ++q) a_src[q] = OA::Source((q * NW + wave) * 64 + lane, p.lda);
#pragma unroll
  for (int q = 0; q < B_DMA; ++q) b_src[q] = OB::Source((q * NW + wave) * 64 + lane, p.ldb);
  auto issue_a = [&](int stage, int64_t k0) {
#pragma unroll
    for (int q = 0; q < A_DMA; ++q)
      DmaToLds(A + a_src[q] + k0 * a_kstep, &lds[stage * kStage + (q * NW + wave) * 64]);
  };
  auto issue_b = [&](int stage, int64_t k0) {
#pragma unroll
    for (int q = 0; q < B_DMA; ++q)
      DmaToLds(B + b_src[q] + k0 * b_kstep,
               &lds[stage * kStage + OA::kChunks + (q * NW + wave) * 64]);
  };

  Acc acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = Acc{};

  // Slices kt0 + i, i < ktiles; ring position i % STAGES.
  int kt0, kt1;
  KSlices<T, BM, BN, BK>(p, m0, n0, &kt0, &kt1);
  const int ktiles = kt1 - kt0;
  const int64_t kbase = int64_t(kt0) * BK;
#pragma unroll
  for (int st = 0; st < STAGES - 1; ++st) {
    if (st < ktiles) {
      issue_a(st, kbase + int64_t(st) * BK);
      issue_b(st, kbase + int64_t(st) * BK);
    }
  }
  const int li = lane & 15, lk = lane >> 4;
  for (int kt = 0; kt < ktiles; ++kt) {
    WaitSlices<PER_SLICE>(min(STAGES - 2, ktiles - 1 - kt));
    BarrierNoFence();
    const int next = kt + STAGES - 1, cur = kt % STAGES;
    const T* as = reinterpret_cast<const T*>(&lds[cur * kStage]);
    const T* bs = reinterpret_cast<const T*>(&lds[cur * kStage + OA::kChunks]);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T af[TM], bf[TN];
#pragma unroll
      for (int ti = 0; ti < TM; ++ti) af[ti] = as[OA::At(wm0 + ti * 16 + li, kk + lk)];
#pragma unroll
      for (int tj = 0; tj < TN; ++tj) bf[tj] = bs[OB::At(wn0 + tj * 16 + li, kk + lk)];
#pragma unroll
      for (int ti = 0; ti < TM; ++ti)
#pragma unroll
        for (int tj = 0; tj < TN; ++tj) acc[ti][tj] = M::Run(af[ti], bf[tj], acc[ti][tj]);
      if (next < ktiles) {
        if (kk == 0) issue_a(next % STAGES, kbase + int64_t(next) * BK);
        if (kk == 4) issue_b(next % STAGES, kbase + int64_t(next) * BK);
      }
    }
  }
  StoreTile<T, TM, TN>(p, acc, m0 + wm0, n0 + wn0, lane);
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
  const unsigned splits = q.ksplit ? unsigned((q.k + q.ksplit - 1) / q.ksplit) : 1u;
  hipLaunchKernelGGL((GemmKernel<T, TA, TB, BM, BN, BK, WGM, WGN, OCC>),
                     dim3(unsigned(blocks), splits), dim3(64 * WGM * WGN), 0, stream, q);
  HWY_HIP_CHECK(hipGetLastError());
}

template <typename T, bool TA, bool TB>
void LaunchConfig(int cfg, const GemmParams<T>& prm, hipStream_t stream);

// A DMA configuration: its kernel on the full tiles when the call allows it
// (k a multiple of BK, 16-byte aligned operands, leading dimensions of whole
// chunks; a triangle of C or a k-range only without strips), the built-in
// plain configuration on the bottom row strip and the right column strip;
// otherwise the plain configuration on all of C.
// Whether a DMA configuration is built for type T: its LDS ring fits the
// 160 KB of a CU and its accumulators leave room for the rest of the
// registers at its occupancy (otherwise the configuration index runs the
// plain rule for that type).
template <typename T, int BM, int BN, int BK, int WGM, int WGN, int OCC, int STAGES>
constexpr bool DmaBuilt() {
  return int64_t(STAGES) * (BM + BN) * BK * int(sizeof(T)) <= 160 * 1024 &&
         (BM / WGM / 16) * (BN / WGN / 16) * int(sizeof(T)) <= 512 / OCC - 64;
}

template <typename T, bool TA, bool TB, int BM, int BN, int BK, int WGM, int WGN, int OCC,
          int STAGES>
void LaunchDma(const GemmParams<T>& prm, hipStream_t stream) {
  constexpr int E = 16 / int(sizeof(T));
  const int64_t mi = prm.m / BM * BM, ni = prm.n / BN * BN;
  // A triangle of C or a k-range that follows the tile position (defined on
  // the whole C) is taken only when there are no strips; a triangle needs a
  // square C on square tiles.
  const bool plain = prm.part == 0 && prm.krange == 0;
  const bool whole = mi == prm.m && ni == prm.n &&
                     (prm.part == 0 || (prm.m == prm.n && BM == BN));
  // Split k (GemmParams::ksplit) only without strips, in whole slices.
  const bool split_ok =
      prm.ksplit == 0 || (plain && mi == prm.m && ni == prm.n && prm.ksplit % BK == 0);
  const bool ok = (plain || whole) && split_ok && mi > 0 && ni > 0 && prm.k > 0 &&
                  prm.k % BK == 0 &&
                  prm.lda % E == 0 && prm.ldb % E == 0 &&
                  reinterpret_cast<uintptr_t>(prm.a) % 16 == 0 &&
                  reinterpret_cast<uintptr_t>(prm.b) % 16 == 0 &&
                  (mi / BM) * (ni / BN) < (int64_t(1) << 31);
  if (!ok) {
    LaunchConfig<T, TA, TB>(ChooseRegularGemmConfig(sizeof(T), prm.m, prm.n, prm.k), prm, stream);
    return;
  }
  if constexpr (!DmaBuilt<T, BM, BN, BK, WGM, WGN, OCC, STAGES>()) {
    LaunchConfig<T, TA, TB>(ChooseRegularGemmConfig(sizeof(T), prm.m, prm.n, prm.k), prm, stream);
  } else {
    GemmParams<T> q = prm;
    q.m = mi;
    q.n = ni;
    q.tiles_m = int(mi / BM);
    q.tiles_n = int(ni / BN);
    q.tri_grid = q.part != 0 ? 1 : 0;
    const int64_t blocks = q.tri_grid ? int64_t(q.tiles_m) * (q.tiles_m + 1) / 2
                                      : int64_t(q.tiles_m) * q.tiles_n;
    const unsigned splits = q.ksplit ? unsigned((q.k + q.ksplit - 1) / q.ksplit) : 1u;
    hipLaunchKernelGGL((GemmDmaKernel<T, TA, TB, BM, BN, BK, WGM, WGN, OCC, STAGES>),
                       dim3(unsigned(blocks), splits), dim3(64 * WGM * WGN), 0, stream, q);
    HWY_HIP_CHECK(hipGetLastError());
    if (mi < prm.m) {  // rows mi .. m - 1, every column
      GemmParams<T> r = prm;
      r.m = prm.m - mi;
      r.a = prm.a + (TA ? mi * prm.lda : mi);
      r.c = prm.c + mi;
      LaunchConfig<T, TA, TB>(ChooseRegularGemmConfig(sizeof(T), r.m, r.n, r.k), r, stream);
    }
    if (ni < prm.n) {  // rows 0 .. mi - 1, columns ni .. n - 1
      GemmParams<T> r = prm;
      r.m = mi;
      r.n = prm.n - ni;
      r.b = prm.b + (TB ? ni : ni * prm.ldb);
      r.c = prm.c + ni * prm.ldc;
      LaunchConfig<T, TA, TB>(ChooseRegularGemmConfig(sizeof(T), r.m, r.n, r.k), r, stream);
    }
  }
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
#define HWY_GEMM_DMA_CASE(ID, BM, BN, BK, WM, WN, OCC, ST) \
  case ID:                                                 \
    LaunchDma<T, TA, TB, BM, BN, BK, WM, WN, OCC, ST>(prm, stream); break;
    HWY_GEMM_DMA_CONFIGS(HWY_GEMM_DMA_CASE)
#undef HWY_GEMM_DMA_CASE
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
