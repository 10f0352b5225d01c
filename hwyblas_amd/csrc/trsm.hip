// Triangular solve op(A) X = B (left) or X op(A) = B (right) on the GPU.
//
// Structure (hwyBLAS src/trsm.cc:87-135): the triangle is halved at
// multiples of the base size until a diagonal block has at most kNB rows;
// every off-diagonal update B2 -= T21 X1 (and its upper/right variants) is a
// GEMM with alpha = -1, beta = 1 on the MFMA kernel, so almost all flops run
// there. Diagonal blocks are solved by DiagSolveKernel:
//  - the block of op(A) is loaded into LDS as a dense kNB x kNB matrix L in
//    "solve orientation" (L = T for the left side, T^T for the right side;
//    only the referenced triangle is read, the unit diagonal is never read);
//  - each thread owns one right-hand side (a column of B for the left side,
//    a row of B for the right side) in registers, staged through LDS so that
//    global reads and writes are coalesced;
//  - x_j = (b_j - sum_q x_q t(q, j)) * (1 / t(j, j)), the sum fused in solve
//    order, as hwyBLAS's base kernel computes it (src/trsm-inl.h:107-174,
//    reciprocal at :69). The update is done column by column (after x_q is
//    final, every later x_j gets its fma), which performs the same fma
//    sequence on each accumulator as the dot form but with independent
//    chains across j.
// Indices are padded to kNB so that register arrays are indexed by
// compile-time constants: an ascending solve keeps the block at [0, r), a
// descending one at [kNB - r, kNB); padding rows are zero with unit diagonal
// and are never stored.
#include <algorithm>

#include "common.h"

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

// Solves L x = b for NT right-hand sides per block: columns of B (LEFT,
// L = T) or rows of B (right side, L = T^T). ASC: ascending solve order
// (L lower). Both sides stage their right-hand sides through LDS, so global
// reads and writes are coalesced and every register index is a constant.
template <typename T, bool LEFT, bool ASC>
__global__ void __launch_bounds__(64)
    DiagSolveKernel(TriView<T> t, int r, T* b, int64_t ldb, int64_t other) {
  constexpr int NT = 64;
  __shared__ T Ls[kNB][kNB + 1];
  __shared__ T inv[kNB];
  __shared__ T Bs[NT][kNB + 1];
  const int tid = threadIdx.x;
  const int off = ASC ? 0 : kNB - r;  // padded index = real index + off

  // L(i, q) = op(A)(i, q) (left) or op(A)(q, i) (right), strictly inside the
  // referenced triangle; everything else is 0.
  for (int idx = tid; idx < kNB * kNB; idx += NT) {
    const int P = idx % kNB, Q = idx / kNB;
    const int i = P - off, q = Q - off;
    T v = T(0);
    const bool in = (i >= 0 && i < r && q >= 0 && q < r) && (ASC ? q < i : q > i);
    if (in) v = LEFT ? t.data[i * t.rs + q * t.cs] : t.data[q * t.rs + i * t.cs];
    Ls[P][Q] = v;
  }
  for (int P = tid; P < kNB; P += NT) {
    const int i = P - off;
    T d = T(1);
    if (!t.unit && i >= 0 && i < r) d = T(1) / t.data[i * t.rs + i * t.cs];
    inv[P] = d;
  }

  // Right-hand side c of this block, element i: B(i, first + c) (left) or
  // B(first + c, i) (right). Consecutive threads read consecutive addresses.
  const int64_t first = int64_t(blockIdx.x) * NT;
  const int cols = int(std::min<int64_t>(NT, other - first));
  for (int idx = tid; idx < r * cols; idx += NT) {
    int i, c;
    if (LEFT) { i = idx % r; c = idx / r; }
    else { c = idx % cols; i = idx / cols; }
    const int64_t g = LEFT ? i + (first + c) * ldb : (first + c) + int64_t(i) * ldb;
    Bs[c][i + off] = b[g];
  }
  __syncthreads();

  T x[kNB];
#pragma unroll
  for (int P = 0; P < kNB; ++P) {
    const int i = P - off;
    x[P] = (i >= 0 && i < r && tid < cols) ? Bs[tid][P] : T(0);
  }

  const bool unit = t.unit;
#pragma unroll
  for (int s = 0; s < kNB; ++s) {
    const int P = ASC ? s : kNB - 1 - s;
    if (!unit) x[P] = x[P] * inv[P];
#pragma unroll
    for (int s2 = s + 1; s2 < kNB; ++s2) {
      const int R = ASC ? s2 : kNB - 1 - s2;
      x[R] = __builtin_fma(-Ls[R][P], x[P], x[R]);
    }
  }

  __syncthreads();  // every thread has read its right-hand side from Bs
#pragma unroll
  for (int P = 0; P < kNB; ++P) Bs[tid][P] = x[P];
  __syncthreads();
  for (int idx = tid; idx < r * cols; idx += NT) {
    int i, c;
    if (LEFT) { i = idx % r; c = idx / r; }
    else { c = idx % cols; i = idx / cols; }
    const int64_t g = LEFT ? i + (first + c) * ldb : (first + c) + int64_t(i) * ldb;
    b[g] = Bs[c][i + off];
  }
}

template <typename T>
void DiagSolve(bool left, const TriView<T>& t, int r, T* b, int64_t ldb,
               int64_t other, hipStream_t stream) {
  // Solve orientation: L = T (left) or T^T (right); lower L solves ascending.
  const bool asc = left ? t.lower : !t.lower;
  const unsigned blocks = unsigned((other + 63) / 64);
  if (left) {
    if (asc)
      hipLaunchKernelGGL((DiagSolveKernel<T, true, true>), dim3(blocks), dim3(64), 0, stream,
                         t, r, b, ldb, other);
    else
      hipLaunchKernelGGL((DiagSolveKernel<T, true, false>), dim3(blocks), dim3(64), 0, stream,
                         t, r, b, ldb, other);
  } else {
    if (asc)
      hipLaunchKernelGGL((DiagSolveKernel<T, false, true>), dim3(blocks), dim3(64), 0, stream,
                         t, r, b, ldb, other);
    else
      hipLaunchKernelGGL((DiagSolveKernel<T, false, false>), dim3(blocks), dim3(64), 0, stream,
                         t, r, b, ldb, other);
  }
  HWY_HIP_CHECK(hipGetLastError());
}

// B(m x n, ldb) -= op1 op2, with op1/op2 general strided operands (one of
// rs, cs is 1) and B column-major.
template <typename T>
void SubtractProduct(int64_t m, int64_t n, int64_t k, const T* d1, int64_t rs1,
                     int64_t cs1, const T* d2, int64_t rs2, int64_t cs2, T* b,
                     int64_t ldb, hipStream_t stream) {
  if (m == 0 || n == 0 || k == 0) return;
  const bool t1 = rs1 != 1, t2 = rs2 != 1;
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
