// The 64-right-hand-side triangular solve one wave runs (shared by the
// grouped TRSM and the diagonal blocks of the single-call TRSM).
//
// A wave solves L X = B for up to 64 right-hand sides, where L is the solve
// matrix in solve order (lower triangular). Blocks of 16 rows are solved in
// the wave's LDS, each element as (b - sum_q l_iq x_q) * (1 / l_ii)
// (hwyBLAS's dot form with the reciprocal, src/trsm-inl.h:107-174, :69);
// each solved block then updates every later block, B_t -= L(t, b) X_b, on
// the MFMA units with X_b read from LDS. Left and right sides, both
// triangles and both transpositions map to one lower solve through index
// maps; only the referenced triangle is read, a unit diagonal never.
#ifndef HWYBLAS_AMD_CSRC_TRSM_CHUNK_H_
#define HWYBLAS_AMD_CSRC_TRSM_CHUNK_H_

#include "common.h"
#include "gemm_kernel.h"

namespace hwy_amd {
namespace trsm_chunk {

using gemm_detail::Mfma;

template <typename T>
struct TrsmDesc {
  const T* a;       // op(A)(r, c) = a[r * rs + c * cs]
  T* b;
  int64_t rs, cs, ldb;
  int m, n;
  int left;         // op(A) X = alpha B (1) or X op(A) = alpha B (0)
  int lower_solve;  // the solve matrix L(i, q) is lower triangular
  int unit;
  T alpha;
};

constexpr int kTrsmTile = 16;  // rows per block: one MFMA tile
constexpr int kTrsmRhs = 64;   // right-hand sides per chunk: one per lane

// L(i, q): the solve matrix, op(A) for the left side and op(A)^T for the
// right side; op(A)(i, q) = A(i, q) or A(q, i).
template <typename T>
__device__ __forceinline__ T SolveElem(const TrsmDesc<T>& d, int i, int q) {
  const int r = d.left ? i : q, c = d.left ? q : i;   // op(A)(r, c)
  return d.a[int64_t(r) * d.rs + int64_t(c) * d.cs];
}

// One wave's LDS: the current block of right-hand sides (solve rows x
// right-hand sides), the block's triangle and its inverse diagonal.
template <typename T>
struct TrsmWaveLds {
  T x[kTrsmTile][kTrsmRhs + 1];
  T l[kTrsmTile][kTrsmTile + 1];
  T inv[kTrsmTile];
};

// Orders this wave's LDS accesses across its lanes (no workgroup barrier:
// the waves of a workgroup work on unrelated chunks).
__device__ __forceinline__ void WaveSync() {
  __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
  __builtin_amdgcn_wave_barrier();
}

// Solves for the right-hand sides v0 .. v0 + 63 of problem d. In solve order
// e (ascending rows when L is lower, descending otherwise) L is lower
// triangular. Blocks of 16 rows are solved in turn: the block's triangle is
// staged in LDS and each lane solves its right-hand side there,
// x_i = (b_i - sum_q l_iq x_q) * (1 / l_ii) (hwyBLAS's dot form with the
// reciprocal, src/trsm-inl.h:107-174); then every later block is updated,
// B_t -= L(t, b) X_b, on the MFMA units with X_b read from LDS.
template <typename T, bool LEFT>
__device__ void TrsmChunkSide(const TrsmDesc<T>& d, int v0, TrsmWaveLds<T>& s) {
  using M = Mfma<T>;
  using Acc = typename M::Acc;
  const int lane = threadIdx.x & 63, li = lane & 15, lk = lane >> 4;
  const int len = LEFT ? d.m : d.n;  // system size
  const int nv = LEFT ? d.n : d.m;   // right-hand sides
  const int nvc = min(kTrsmRhs, nv - v0);
  const bool asc = d.lower_solve;
  auto phys = [&](int e) { return asc ? e : len - 1 - e; };
  // Element e of right-hand side v0 + c.
  auto bptr = [&](int e, int c) -> T* {
    const int64_t row = phys(e), v = v0 + c;
    return LEFT ? d.b + row + v * d.ldb : d.b + v + row * d.ldb;
  };
  auto lval = [&](int e, int q) { return SolveElem(d, phys(e), phys(q)); };

  // Memory order: for the left side the elements of one right-hand side are
  // contiguous (lanes walk rows), for the right side the right-hand sides
  // are (lanes walk right-hand sides); loads and stores follow it.
  constexpr bool left = LEFT;

  // B := alpha B; alpha == 0 leaves zeros without reading A
  // (src/trsm.cc:172-175).
  if (d.alpha != T(1)) {
    if constexpr (left) {
      for (int c = 0; c < nvc; ++c)
        for (int e = lane; e < len; e += 64) {
          T* x = bptr(e, c);
          *x = d.alpha == T(0) ? T(0) : *x * d.alpha;
        }
    } else if (lane < nvc) {
      for (int e = 0; e < len; ++e) {
        T* x = bptr(e, lane);
        *x = d.alpha == T(0) ? T(0) : *x * d.alpha;
      }
    }
    if (d.alpha == T(0)) return;
  }
  const int nb = (len + kTrsmTile - 1) / kTrsmTile;
  for (int b = 0; b < nb; ++b) {
    const int e0 = b * kTrsmTile, cnt = min(kTrsmTile, len - e0);
    // The diagonal block in LDS, then the block of right-hand sides.
    for (int idx = lane; idx < kTrsmTile * kTrsmTile; idx += 64) {
      const int i = idx / kTrsmTile, q = idx % kTrsmTile;
      s.l[i][q] = (q < i && i < cnt) ? lval(e0 + i, e0 + q) : T(0);
    }
    if (lane < kTrsmTile)
      s.inv[lane] = (lane < cnt && !d.unit) ? T(1) / lval(e0 + lane, e0 + lane) : T(1);
    if constexpr (left) {
#pragma unroll 4
      for (int it = 0; it < kTrsmRhs / 4; ++it) {
        const int i = li, c = 4 * it + lk;
        s.x[i][c] = (i < cnt && c < nvc) ? *bptr(e0 + i, c) : T(0);
      }
    } else {
#pragma unroll 4
      for (int i = 0; i < kTrsmTile; ++i)
        s.x[i][lane] = (i < cnt && lane < nvc) ? *bptr(e0 + i, lane) : T(0);
    }
    WaveSync();
    // Each lane solves its own column of s.x (rows past cnt stay 0).
    for (int i = 0; i < cnt; ++i) {
      T t = s.x[i][lane];
      for (int q = 0; q < i; ++q) t = __builtin_fma(-s.l[i][q], s.x[q][lane], t);
      s.x[i][lane] = t * s.inv[i];
    }
    WaveSync();
    if constexpr (left) {
#pragma unroll 4
      for (int it = 0; it < kTrsmRhs / 4; ++it) {
        const int i = li, c = 4 * it + lk;
        if (i < cnt && c < nvc) *bptr(e0 + i, c) = s.x[i][c];
      }
    } else if (lane < nvc) {
      for (int i = 0; i < cnt; ++i) *bptr(e0 + i, lane) = s.x[i][lane];
    }
    // B_t -= L(t, b) X_b for every later block t. The MFMA tile is
    // (right-hand sides x rows) for the left side, (rows x right-hand sides)
    // for the right side, so that its loads and stores are contiguous.
    for (int t = b + 1; t < nb; ++t) {
      const int r0 = t * kTrsmTile;
      Acc acc[kTrsmRhs / 16];
#pragma unroll
      for (int jt = 0; jt < kTrsmRhs / 16; ++jt) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = left ? r0 + li : r0 + M::Row(lk, r);
          const int col = left ? 16 * jt + M::Row(lk, r) : 16 * jt + li;
          acc[jt][r] = (row < len && col < nvc) ? *bptr(row, col) : T(0);
        }
      }
#pragma unroll
      for (int u = 0; u < kTrsmTile / 4; ++u) {
        const int i = r0 + li, q = 4 * u + lk;
        const T a = (i < len && q < cnt) ? -lval(i, e0 + q) : T(0);
        if constexpr (left) {
#pragma unroll
          for (int jt = 0; jt < kTrsmRhs / 16; ++jt)
            acc[jt] = M::Run(s.x[q][16 * jt + li], a, acc[jt]);
        } else {
#pragma unroll
          for (int jt = 0; jt < kTrsmRhs / 16; ++jt)
            acc[jt] = M::Run(a, s.x[q][16 * jt + li], acc[jt]);
        }
      }
#pragma unroll
      for (int jt = 0; jt < kTrsmRhs / 16; ++jt) {
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = left ? r0 + li : r0 + M::Row(lk, r);
          const int col = left ? 16 * jt + M::Row(lk, r) : 16 * jt + li;
          if (row < len && col < nvc) *bptr(row, col) = acc[jt][r];
        }
      }
    }
    WaveSync();  // the next block overwrites the LDS
  }
}

template <typename T>
__device__ __forceinline__ void TrsmChunk(const TrsmDesc<T>& d, int v0, TrsmWaveLds<T>& s) {
  if (d.left) TrsmChunkSide<T, true>(d, v0, s);
  else TrsmChunkSide<T, false>(d, v0, s);
}

}  // namespace trsm_chunk
}  // namespace hwy_amd

#endif  // HWYBLAS_AMD_CSRC_TRSM_CHUNK_H_
