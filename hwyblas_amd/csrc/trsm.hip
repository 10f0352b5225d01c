// Triangular solve op(A) X = B (left) or X op(A) = B (right) on the GPU.
//
// Structure (hwyBLAS src/trsm.cc:87-135): the triangle is halved at
// multiples of the base size until a diagonal block has at most kNB rows;
// every off-diagonal update B2 -= T21 X1 (and its upper/right variants) is a
// GEMM with alpha = -1, beta = 1 on the MFMA kernel, so almost all flops run
// there. Diagonal blocks (order <= kNB) are solved by DiagNarrowKernel: the
// block's rows across a wave's lanes, a few right-hand sides per wave in
// registers, x_j = (b_j - sum_q t(j, q) x_q) * (1 / t(j, j)) with the sum as
// FMAs in ascending q (the dot form of hwyBLAS's base kernel,
// src/trsm-inl.h:107-174, reciprocal at :69). Only the referenced triangle
// is read; a unit diagonal is never read.
#include <stdlib.h>

#include <algorithm>

#include "common.h"
#include "trsm_chunk.h"

namespace hwy_amd {
namespace {

constexpr int kNB = 64;

template <typename T>
struct TriView {
  const T* data;   // element (i, j) of the triangle op(A): data[i*rs + j*cs]
  int64_t rs, cs;
  bool lower;      // of op(A)
  bool unit;
};

template <typename T>
TriView<T> SubView(const TriView<T>& t, int64_t i0, int64_t j0) {
  TriView<T> s = t;
  s.data += i0 * t.rs + j0 * t.cs;
  return s;
}

// A diagonal block (order r <= kNB): a workgroup of eight waves takes
// kNarrowRhs right-hand sides (columns of B for the left side, rows for the
// right side), kWaveRhs per wave; lane e holds solve row e of its wave's
// right-hand sides in registers, and the block's triangle sits in LDS. Step i scales row i by
// 1 / t(i, i) and broadcasts it, and every later row subtracts t(e, i) x_i
// with one FMA per right-hand side; so x_i = (b_i - sum_{q<i} t(i, q) x_q)
// (1 / t(i, i)) with the sum accumulated by FMAs in ascending q, the dot
// form of hwyBLAS's base kernel (src/trsm-inl.h:107-174, reciprocal at
// :69). B is read and written once; every load is issued before the first
// is waited for.
constexpr int kWaveRhs = 2;
constexpr int kNarrowWaves = 8;
constexpr int kNarrowRhs = kWaveRhs * kNarrowWaves;

// Lane i's value in every lane (i wave-uniform): v_readlane per 32 bits.
template <typename T>
__device__ __forceinline__ T Broadcast(T v, int i) {
  if constexpr (sizeof(T) == 8) {
    const long long bits = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane(int(bits & 0xFFFFFFFFll), i);
    const int hi = __builtin_amdgcn_readlane(int(bits >> 32), i);
    return __longlong_as_double((static_cast<long long>(hi) << 32) |
                                static_cast<unsigned int>(lo));
  } else {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), i));
  }
}

template <typename T, bool LEFT>
__global__ void __launch_bounds__(64 * kNarrowWaves)
    DiagNarrowKernel(trsm_chunk::TrsmDesc<T> d) {
  __shared__ T st[kNB][kNB + 1];  // st[e][q] = L(e, q), q < e, in solve order
  __shared__ T sinv[kNB];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int r = LEFT ? d.m : d.n;      // the system's order
  const int nrhs = LEFT ? d.n : d.m;   // right-hand sides: columns (left) or rows of B
  const int c0 = int(blockIdx.x) * kNarrowRhs + wave * kWaveRhs;
  const int nvc = min(kWaveRhs, nrhs - c0);  // may be <= 0 for the last waves
  const bool asc = d.lower_solve;
  auto phys = [&](int e) { return asc ? e : r - 1 - e; };
  T x[kWaveRhs];
  const int64_t e_phys = lane < r ? phys(lane) : 0;
  // Element (solve row e, right-hand side c) of B.
  auto at = [&](int c) -> T* {
    return LEFT ? d.b + e_phys + int64_t(c0 + c) * d.ldb : d.b + (c0 + c) + e_phys * d.ldb;
  };
#pragma unroll
  for (int c = 0; c < kWaveRhs; ++c) x[c] = (lane < r && c < nvc) ? *at(c) : T(0);
  // Wave w loads the triangle's columns q = w, w + 8, ... of its lane's row.
  constexpr int kCols = kNB / kNarrowWaves;
  T trow[kCols];
#pragma unroll
  for (int u = 0; u < kCols; ++u) {
    const int q = wave + kNarrowWaves * u;
    trow[u] = (lane < r && q < lane) ? trsm_chunk::SolveElem(d, phys(lane), phys(q)) : T(0);
  }
  if (wave == 0) {
    const T diag =
        (lane < r && !d.unit) ? trsm_chunk::SolveElem(d, phys(lane), phys(lane)) : T(1);
    sinv[lane] = T(1) / diag;
  }
#pragma unroll
  for (int u = 0; u < kCols; ++u) st[lane][wave + kNarrowWaves * u] = trow[u];
  __syncthreads();
  if (nvc <= 0) return;
  // The lane's row of the triangle and its own reciprocal in registers, so
  // the solve's dependent chain is only scale -> v_readlane -> FMA. Steps at
  // or past r touch zeros only (padding rows and columns are 0, their
  // diagonal 1).
  T tcol[kNB];
#pragma unroll
  for (int i = 0; i < kNB; ++i) tcol[i] = st[lane][i];
  const T myinv = sinv[lane];
#pragma unroll
  for (int i = 0; i < kNB; ++i) {
#pragma unroll
    for (int c = 0; c < kWaveRhs; ++c) {
      if (lane == i) x[c] = x[c] * myinv;
      const T xi = Broadcast(x[c], i);
      if (lane > i) x[c] = __builtin_fma(-tcol[i], xi, x[c]);
    }
  }
  if (lane < r) {
#pragma unroll
    for (int c = 0; c < kWaveRhs; ++c)
      if (c < nvc) *at(c) = x[c];
  }
}

// The recursion's leaves fused: a diagonal block of order 128 or 256 is
// solved in one launch that runs the recursion's own steps on kNarrowRhs
// right-hand sides per workgroup, the block of B kept in LDS: each solve of
// a 64-row diagonal block as DiagNarrowKernel does it, and each
// off-diagonal update as the GEMM kernel does it (v_mfma 16x16x4 steps over
// k ascending from a zero accumulator, the same operand roles, then
// (-acc) + B), so the bits are the unfused recursion's.
constexpr int kFusedMax = 256;
constexpr int kMaxFusedSteps = 8;

struct FusedStep {
  int update;  // 0: solve rows [r0, r0 + len); 1: rows [r0, r0 + len) -= T(., [k0, k0 + klen)) X
  int r0, len, k0, klen;
};
struct FusedPlan {
  int nsteps;
  FusedStep s[kMaxFusedSteps];
};

// The steps Recurse takes on a diagonal block of order r at solve index r0;
// `lower`: the solve runs in ascending index order (a lower triangle on the
// left side, an upper one on the right side).
constexpr void PlanFused(bool lower, int r0, int r, FusedPlan& plan) {
  if (r <= kNB) {
    plan.s[plan.nsteps++] = FusedStep{0, r0, r, 0, 0};
    return;
  }
  const int half = (r / 2) / kNB * kNB;
  const int r1 = half > kNB ? half : kNB, r2 = r - r1;
  if (lower) {
    PlanFused(lower, r0, r1, plan);
    plan.s[plan.nsteps++] = FusedStep{1, r0 + r1, r2, r0, r1};
    PlanFused(lower, r0 + r1, r2, plan);
  } else {
    PlanFused(lower, r0 + r1, r2, plan);
    plan.s[plan.nsteps++] = FusedStep{1, r0, r1, r0 + r1, r2};
    PlanFused(lower, r0, r1, plan);
  }
}
constexpr FusedPlan MakeFusedPlan(bool lower, int r) {
  FusedPlan plan{};
  PlanFused(lower, 0, r, plan);
  return plan;
}
// Indexed by 2 (order == 256) + lower.
__constant__ FusedPlan kFusedPlans[4] = {MakeFusedPlan(false, 128), MakeFusedPlan(true, 128),
                                         MakeFusedPlan(false, 256), MakeFusedPlan(true, 256)};
static_assert(MakeFusedPlan(true, 256).nsteps == 7 && MakeFusedPlan(false, 128).nsteps == 3,
              "fused plans");

// At most 128 VGPRs (four waves per SIMD): two workgroups per CU.
template <typename T, bool LEFT>
__global__ void __launch_bounds__(64 * kNarrowWaves, 4)
    FusedBlockKernel(trsm_chunk::TrsmDesc<T> d, int R) {
  using M = gemm_detail::Mfma<T>;
  using Acc = typename M::Acc;
  // sb[e][c]: solve index e (a row of B for the left side, a column for the
  // right side) x this workgroup's right-hand side c.
  __shared__ T sb[kFusedMax][kNarrowRhs + 1];
  __shared__ T st[kNB][kNB + 1];
  __shared__ T sinv[kNB];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int li = lane & 15, lk = lane >> 4;
  const int c0 = int(blockIdx.x) * kNarrowRhs;
  const int nvc = min(kNarrowRhs, (LEFT ? d.n : d.m) - c0);
  // Element (solve index e, right-hand side c) of B; the thread order of the
  // copies follows B's contiguous direction.
  auto at = [&](int e, int c) -> T* {
    return LEFT ? d.b + e + int64_t(c0 + c) * d.ldb : d.b + (c0 + c) + int64_t(e) * d.ldb;
  };
  auto split = [&](int idx, int& e, int& c) {
    if (LEFT) {
      e = idx % kFusedMax;
      c = idx / kFusedMax;
    } else {
      c = idx % kNarrowRhs;
      e = idx / kNarrowRhs;
    }
  };
  constexpr int NT = 64 * kNarrowWaves;
  {
    constexpr int kPer = kFusedMax * kNarrowRhs / NT;
    T v[kPer];
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      int e, c;
      split(int(threadIdx.x) + u * NT, e, c);
      v[u] = (e < R && c < nvc) ? *at(e, c) : T(0);
    }
#pragma unroll
    for (int u = 0; u < kPer; ++u) {
      int e, c;
      split(int(threadIdx.x) + u * NT, e, c);
      sb[e][c] = v[u];
    }
  }
  __syncthreads();
  const bool asc = d.lower_solve;
  const FusedPlan& plan = kFusedPlans[(R == 256 ? 2 : 0) + (asc ? 1 : 0)];
  for (int si = 0; si < plan.nsteps; ++si) {
    const FusedStep sp = plan.s[si];
    if (!sp.update) {
      // A diagonal block: rows r0 .. r0 + len - 1, solve order e.
      const int r = sp.len;
      auto phys = [&](int e) { return sp.r0 + (asc ? e : r - 1 - e); };
      constexpr int kCols = kNB / kNarrowWaves;
      T trow[kCols];
#pragma unroll
      for (int u = 0; u < kCols; ++u) {
        const int q = wave + kNarrowWaves * u;
        trow[u] = (lane < r && q < lane) ? trsm_chunk::SolveElem(d, phys(lane), phys(q)) : T(0);
      }
      if (wave == 0) {
        const T diag =
            (lane < r && !d.unit) ? trsm_chunk::SolveElem(d, phys(lane), phys(lane)) : T(1);
        sinv[lane] = T(1) / diag;
      }
#pragma unroll
      for (int u = 0; u < kCols; ++u) st[lane][wave + kNarrowWaves * u] = trow[u];
      __syncthreads();
      const T myinv = sinv[lane];
      const int prow = lane < r ? phys(lane) : 0;
      T x[kWaveRhs];
#pragma unroll
      for (int c = 0; c < kWaveRhs; ++c) x[c] = lane < r ? sb[prow][wave * kWaveRhs + c] : T(0);
      // An opaque copy of the lane index: the 2 x 64 lane masks below are
      // then formed where they are used, not hoisted out of the step loop
      // into (spilled) scalar registers.
      int ln = lane;
      asm volatile("" : "+v"(ln));
      // The lane's row of the triangle read kChunk columns ahead of use, the
      // compiler fence keeping the later chunks' reads (and registers) out
      // of the current chunk's steps.
      constexpr int kChunk = 16;
      T tc[kChunk], tn[kChunk];
#pragma unroll
      for (int j = 0; j < kChunk; ++j) tc[j] = st[lane][j];
#pragma unroll
      for (int ib = 0; ib < kNB; ib += kChunk) {
        if (ib + kChunk < kNB) {
#pragma unroll
          for (int j = 0; j < kChunk; ++j) tn[j] = st[lane][ib + kChunk + j];
        }
        asm volatile("" ::: "memory");
#pragma unroll
        for (int j = 0; j < kChunk; ++j) {
          const int i = ib + j;
#pragma unroll
          for (int c = 0; c < kWaveRhs; ++c) {
            if (ln == i) x[c] = x[c] * myinv;
            const T xi = Broadcast(x[c], i);
            if (ln > i) x[c] = __builtin_fma(-tc[j], xi, x[c]);
          }
        }
#pragma unroll
        for (int j = 0; j < kChunk; ++j) tc[j] = tn[j];
      }
      if (lane < r) {
#pragma unroll
        for (int c = 0; c < kWaveRhs; ++c) sb[prow][wave * kWaveRhs + c] = x[c];
      }
    } else {
      // Solve indices r0 .. r0 + len - 1 less the product of the solved
      // k0 .. k0 + klen - 1 with the triangle, as the GEMM kernel forms it:
      // left B(rows) -= T(rows, ks) X(ks, rhs), right B(rhs, cols) -= X(rhs,
      // ks) T(ks, cols); 16 solve indices per block, block rb on wave
      // rb % kNarrowWaves, all 16 right-hand sides at once.
      // When a wave's triangle loads are coalesced (its 16 rows contiguous),
      // all of them are issued before the first product (one memory round
      // trip per update); strided ones go in batches of 8, which measured
      // faster for them. The products are the same either way.
      constexpr int kMaxSteps = kFusedMax / 2 / 4;  // the longest update has k = 128
      const int nrb = sp.len / 16, nsteps = sp.klen / 4;
      const bool coalesced = (LEFT ? d.rs : d.cs) == 1;
      for (int rb = wave; rb < nrb; rb += kNarrowWaves) {
        const int i0 = sp.r0 + rb * 16;
        Acc acc = Acc{};
        if (coalesced) {
          T ta[kMaxSteps];
#pragma unroll
          for (int u = 0; u < kMaxSteps; ++u)
            ta[u] = u < nsteps ? trsm_chunk::SolveElem(d, i0 + li, sp.k0 + 4 * u + lk) : T(0);
#pragma unroll
          for (int u = 0; u < kMaxSteps; ++u) {
            if (u < nsteps) {
              const T x = sb[sp.k0 + 4 * u + lk][li];
              acc = LEFT ? M::Run(ta[u], x, acc) : M::Run(x, ta[u], acc);
            }
          }
        } else {
#pragma unroll 8
          for (int kk = sp.k0; kk < sp.k0 + sp.klen; kk += 4) {
            const T t = trsm_chunk::SolveElem(d, i0 + li, kk + lk);
            const T x = sb[kk + lk][li];
            acc = LEFT ? M::Run(t, x, acc) : M::Run(x, t, acc);
          }
        }
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          // Left: output row i0 + Row(lk, e) (solve index), column li (rhs);
          // right: output row Row(lk, e) (rhs), column i0 + li (solve index).
          const int eo = LEFT ? i0 + M::Row(lk, e) : i0 + li;
          const int co = LEFT ? li : M::Row(lk, e);
          T v = acc[e] * T(-1);
          v = v + sb[eo][co];
          sb[eo][co] = v;
        }
      }
    }
    __syncthreads();
  }
#pragma unroll 4
  for (int idx = int(threadIdx.x); idx < R * kNarrowRhs; idx += NT) {
    const int e = LEFT ? idx % R : idx / kNarrowRhs, c = LEFT ? idx / R : idx % kNarrowRhs;
    if (c < nvc) *at(e, c) = sb[e][c];
  }
}

// Whether a diagonal block of order r is solved by FusedBlockKernel (off
// with HWYBLAS_AMD_TRSM_FUSED=0).
bool FusedBlock(int64_t r) {
  static const bool enabled = [] {
    const char* e = getenv("HWYBLAS_AMD_TRSM_FUSED");
    return !(e && e[0] == '0');
  }();
  return enabled && (r == 128 || r == 256);
}

template <typename T>
void FusedSolve(bool left, const TriView<T>& t, int r, T* b, int64_t ldb, int64_t other,
                hipStream_t stream) {
  trsm_chunk::TrsmDesc<T> d;
  d.a = t.data;
  d.rs = t.rs;
  d.cs = t.cs;
  d.b = b;
  d.ldb = ldb;
  d.m = left ? r : int(other);
  d.n = left ? int(other) : r;
  d.left = left;
  d.lower_solve = left ? t.lower : !t.lower;
  d.unit = t.unit;
  d.alpha = T(1);
  const dim3 grid(unsigned((other + kNarrowRhs - 1) / kNarrowRhs));
  if (left)
    hipLaunchKernelGGL((FusedBlockKernel<T, true>), grid, dim3(64 * kNarrowWaves), 0, stream, d,
                       r);
  else
    hipLaunchKernelGGL((FusedBlockKernel<T, false>), grid, dim3(64 * kNarrowWaves), 0, stream, d,
                       r);
  HWY_HIP_CHECK(hipGetLastError());
}

template <typename T>
void DiagSolve(bool left, const TriView<T>& t, int r, T* b, int64_t ldb,
               int64_t other, hipStream_t stream) {
  trsm_chunk::TrsmDesc<T> d;
  d.a = t.data;
  d.rs = t.rs;
  d.cs = t.cs;
  d.b = b;
  d.ldb = ldb;
  d.m = left ? r : int(other);
  d.n = left ? int(other) : r;
  d.left = left;
  // Solve orientation: L = T (left) or T^T (right); lower L solves ascending.
  d.lower_solve = left ? t.lower : !t.lower;
  d.unit = t.unit;
  d.alpha = T(1);
  const dim3 grid(unsigned((other + kNarrowRhs - 1) / kNarrowRhs));
  if (left)
    hipLaunchKernelGGL((DiagNarrowKernel<T, true>), grid, dim3(64 * kNarrowWaves), 0, stream, d);
  else
    hipLaunchKernelGGL((DiagNarrowKernel<T, false>), grid, dim3(64 * kNarrowWaves), 0, stream,
                       d);
  HWY_HIP_CHECK(hipGetLastError());
}

// B(m x n, ldb) -= op1 op2, with op1/op2 general strided operands (one of
// rs, cs is 1) and B column-major. Skinny updates with a long k (few
// right-hand sides against a large triangle) split k like the GEMM entries.
template <typename T>
void SubtractProduct(int64_t m, int64_t n, int64_t k, const T* d1, int64_t rs1,
                     int64_t cs1, const T* d2, int64_t rs2, int64_t cs2, T* b,
                     int64_t ldb, hipStream_t stream) {
  if (m == 0 || n == 0 || k == 0) return;
  const bool t1 = rs1 != 1, t2 = rs2 != 1;
  if (GemmSplitPays<T>(m, n, k)) {
    GemmSplitKPublic<T>(t1, t2, m, n, k, T(-1), d1, t1 ? rs1 : cs1, d2, t2 ? rs2 : cs2, T(1),
                        b, ldb, stream);
    return;
  }
  GemmDevice<T>(t1, t2, m, n, k, T(-1), d1, t1 ? rs1 : cs1, d2, t2 ? rs2 : cs2,
                T(1), b, ldb, stream);
}

template <typename T>
void Recurse(bool left, const TriView<T>& t, int64_t r, int64_t other, T* b,
             int64_t ldb, hipStream_t stream) {
  if (r <= kNB) {
    DiagSolve(left, t, int(r), b, ldb, other, stream);
    return;
  }
  if (FusedBlock(r)) {
    FusedSolve(left, t, int(r), b, ldb, other, stream);
    return;
  }
  const int64_t r1 = std::max<int64_t>(kNB, (r / 2) / kNB * kNB);
  const int64_t r2 = r - r1;
  const TriView<T> t11 = t;
  const TriView<T> t22 = SubView(t, r1, r1);
  const TriView<T> tof = t.lower ? SubView(t, r1, 0) : SubView(t, 0, r1);
  if (left) {
    T* b1 = b;
    T* b2 = b + r1;
    if (t.lower) {
      Recurse(left, t11, r1, other, b1, ldb, stream);
      SubtractProduct(r2, other, r1, tof.data, tof.rs, tof.cs, b1, 1, ldb, b2, ldb, stream);
      Recurse(left, t22, r2, other, b2, ldb, stream);
    } else {
      Recurse(left, t22, r2, other, b2, ldb, stream);
      SubtractProduct(r1, other, r2, tof.data, tof.rs, tof.cs, b2, 1, ldb, b1, ldb, stream);
      Recurse(left, t11, r1, other, b1, ldb, stream);
    }
  } else {
    T* b1 = b;
    T* b2 = b + r1 * ldb;
    if (t.lower) {
      Recurse(left, t22, r2, other, b2, ldb, stream);
      SubtractProduct(other, r1, r2, b2, 1, ldb, tof.data, tof.rs, tof.cs, b1, ldb, stream);
      Recurse(left, t11, r1, other, b1, ldb, stream);
    } else {
      Recurse(left, t11, r1, other, b1, ldb, stream);
      SubtractProduct(other, r2, r1, b1, 1, ldb, tof.data, tof.rs, tof.cs, b2, ldb, stream);
      Recurse(left, t22, r2, other, b2, ldb, stream);
    }
  }
}

}  // namespace

template <typename T>
void TrsmDevice(bool left, bool lower, bool trans, bool unit, int64_t m,
                int64_t n, const T* a, int64_t lda, T* b, int64_t ldb,
                hipStream_t stream) {
  if (m <= 0 || n <= 0) return;
  // op(A) as a strided view; a transposition swaps the strides and the
  // stored triangle (hwyBLAS src/api.h:133-140).
  TriView<T> t;
  t.data = a;
  t.rs = trans ? lda : 1;
  t.cs = trans ? 1 : lda;
  t.lower = trans ? !lower : lower;
  t.unit = unit;
  Recurse(left, t, left ? m : n, left ? n : m, b, ldb, stream);
}

template void TrsmDevice<double>(bool, bool, bool, bool, int64_t, int64_t,
                                 const double*, int64_t, double*, int64_t,
                                 hipStream_t);
template void TrsmDevice<float>(bool, bool, bool, bool, int64_t, int64_t,
                                const float*, int64_t, float*, int64_t,
                                hipStream_t);

}  // namespace hwy_amd
