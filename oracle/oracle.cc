// CPU oracle for the Level-3 hot path. TEST INFRASTRUCTURE ONLY (see
// oracle.h): the product library never links or calls this code.
//
// The restatement follows hwyBLAS's algorithm, file by file:
//   Gemm quick returns / ScaleMatrix ....... src/gemm.cc:48-64, :113-119
//   GeneralOperand / StructuredOperand ..... src/api.h:118-140
//   BlockedGemmImpl (jc/pc/ic/jr/ir loops) .. src/gemm-inl.h:153-232
//   PackOperand (k-major, zero padded) ...... src/pack-inl.h:24-85, :131-172
//   micro-kernel Tile / Store ............... src/gemm_kernel-inl.h:138-224
//   blocking constants (AVX2 target) ........ src/gemm_kernel-inl.h:60-75
//   Trsm (alpha first) / TrsmRecursive ...... src/trsm.cc:87-180
//   TrsmBase / SolveRowChunkReal ............ src/trsm-inl.h:58-174, :277-325
// It is written from that description, not copied: loops are scalar C++
// with explicit fma(), so the arithmetic (products fused into the running
// sum in k order within a KC block, alpha applied after accumulation, beta on
// the first KC block only) is that of an FMA target such as AVX2.
#include "oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <limits>
#include <vector>

#if defined(__AVX2__) && defined(__FMA__)
#include <immintrin.h>
#endif

// The planted mistake (orc_plant_mistake), shared with lu.cc.
int g_mistake = 0;

namespace {

// ---------------------------------------------------------------------------
// Operands: op(X)(i, p) = data[i * rs + p * cs]   (src/gemm.h:28-36)

template <typename T>
struct Opd {
  const T* data;
  ptrdiff_t rs, cs;
  bool lower = false, unit = false;
  const T& at(ptrdiff_t i, ptrdiff_t p) const { return data[i * rs + p * cs]; }
};

bool IsTrans(char c) { return c == 'T' || c == 't' || c == 'C' || c == 'c'; }

template <typename T>
Opd<T> General(char op, const T* x, int ld) {  // src/api.h:118-127
  Opd<T> o;
  o.data = x;
  const bool t = IsTrans(op) != (g_mistake == 2);
  o.rs = t ? ld : 1;
  o.cs = t ? 1 : ld;
  return o;
}

// ---------------------------------------------------------------------------
// Blocking constants of the AVX2 target (4 doubles or 8 floats per vector,
// two vectors per column, NR = 6 columns): src/gemm_kernel-inl.h:60-75.

template <typename T>
struct Blocking;
template <>
struct Blocking<double> {
  static constexpr int MR = 8, NR = 6, KC = 256, MC = 96, NC = 4080;
};
template <>
struct Blocking<float> {
  static constexpr int MR = 16, NR = 6, KC = 384, MC = 144, NC = 4080;
};

template <typename T>
inline T Fma(T a, T b, T c) {
  return std::fma(a, b, c);
}

// Packs rows [r0, r0 + rows) and k range [p0, p0 + kc) of operand x into
// panels of R rows, k-major: panel q holds element (r0 + q*R + i, p0 + p) at
// out[q*R*kc + p*R + i], zero padded to whole panels (src/pack-inl.h:24-85).
template <int R, typename T>
void Pack(const Opd<T>& x, int r0, int rows, int p0, int kc, T* out) {
  for (int q = 0; q * R < rows; ++q) {
    T* panel = out + static_cast<size_t>(q) * R * kc;
    const int rr = std::min(R, rows - q * R);
    for (int p = 0; p < kc; ++p) {
      for (int i = 0; i < R; ++i) {
        panel[p * R + i] = i < rr ? x.at(r0 + q * R + i, p0 + p) : T(0);
      }
    }
  }
}

// C[mr x nr] = alpha * (A_panel B_panel) + beta * C, with the store rules of
// src/gemm_kernel-inl.h:181-224: r = alpha * acc; beta == 1 adds C; any other
// non-zero beta is fused as fma(C, beta, r); beta == 0 never reads C.
// The accumulation loop: acc[j][i] = fma(a_p[i], b_p[j], acc[j][i]) for p in
// k order. With AVX2+FMA it runs on two 256-bit vectors per column, as the
// reference's AVX2 target does (src/gemm_kernel-inl.h:161-176); the scalar
// loop computes the same bits.
template <typename T>
void Accumulate(int kc, const T* a, const T* b,
                T (&acc)[Blocking<T>::NR][Blocking<T>::MR]) {
  constexpr int MR = Blocking<T>::MR, NR = Blocking<T>::NR;
  for (int p = 0; p < kc; ++p) {
    const T* ap = a + p * MR;
    const T* bp = b + p * NR;
    for (int j = 0; j < NR; ++j) {
      const T bj = bp[j];
      for (int i = 0; i < MR; ++i) acc[j][i] = Fma(ap[i], bj, acc[j][i]);
    }
  }
}

#if defined(__AVX2__) && defined(__FMA__)
template <>
void Accumulate<double>(int kc, const double* a, const double* b,
                        double (&acc)[6][8]) {
  __m256d v[6][2];
  for (int j = 0; j < 6; ++j) {
    v[j][0] = _mm256_loadu_pd(&acc[j][0]);
    v[j][1] = _mm256_loadu_pd(&acc[j][4]);
  }
  for (int p = 0; p < kc; ++p) {
    const __m256d a0 = _mm256_loadu_pd(a + p * 8);
    const __m256d a1 = _mm256_loadu_pd(a + p * 8 + 4);
    for (int j = 0; j < 6; ++j) {
      const __m256d bj = _mm256_broadcast_sd(b + p * 6 + j);
      v[j][0] = _mm256_fmadd_pd(a0, bj, v[j][0]);
      v[j][1] = _mm256_fmadd_pd(a1, bj, v[j][1]);
    }
  }
  for (int j = 0; j < 6; ++j) {
    _mm256_storeu_pd(&acc[j][0], v[j][0]);
    _mm256_storeu_pd(&acc[j][4], v[j][1]);
  }
}

template <>
void Accumulate<float>(int kc, const float* a, const float* b,
                       float (&acc)[6][16]) {
  __m256 v[6][2];
  for (int j = 0; j < 6; ++j) {
    v[j][0] = _mm256_loadu_ps(&acc[j][0]);
    v[j][1] = _mm256_loadu_ps(&acc[j][8]);
  }
  for (int p = 0; p < kc; ++p) {
    const __m256 a0 = _mm256_loadu_ps(a + p * 16);
    const __m256 a1 = _mm256_loadu_ps(a + p * 16 + 8);
    for (int j = 0; j < 6; ++j) {
      const __m256 bj = _mm256_broadcast_ss(b + p * 6 + j);
      v[j][0] = _mm256_fmadd_ps(a0, bj, v[j][0]);
      v[j][1] = _mm256_fmadd_ps(a1, bj, v[j][1]);
    }
  }
  for (int j = 0; j < 6; ++j) {
    _mm256_storeu_ps(&acc[j][0], v[j][0]);
    _mm256_storeu_ps(&acc[j][8], v[j][1]);
  }
}
#endif

template <typename T>
void MicroTile(int kc, const T* a, const T* b, T alpha, T beta, T* c,
               ptrdiff_t ldc, int mr, int nr) {
  constexpr int MR = Blocking<T>::MR, NR = Blocking<T>::NR;
  T acc[NR][MR];
  for (int j = 0; j < NR; ++j)
    for (int i = 0; i < MR; ++i) acc[j][i] = T(0);
  const int kend = (g_mistake == 1 && kc > 0) ? kc - 1 : kc;
  Accumulate<T>(kend, a, b, acc);
  const bool beta_zero = (beta == T(0)) || g_mistake == 3;
  const bool beta_one = (beta == T(1));
  for (int j = 0; j < nr; ++j) {
    T* cj = c + j * ldc;
    for (int i = 0; i < mr; ++i) {
      T r = acc[j][i] * alpha;
      if (!beta_zero) r = beta_one ? r + cj[i] : Fma(cj[i], beta, r);
      cj[i] = r;
    }
  }
}

// C = alpha op(A) op(B) + beta C for m, n, k > 0 (src/gemm-inl.h:153-232,
// packed-B path; the direct-B variant reads the same values in the same
// order, so it gives the same bits).
template <typename T>
void BlockedGemm(int m, int n, int k, T alpha, const Opd<T>& a,
                 const Opd<T>& b, T beta, T* c, ptrdiff_t ldc) {
  constexpr int MR = Blocking<T>::MR, NR = Blocking<T>::NR;
  constexpr int MC = Blocking<T>::MC / MR * MR, KC = Blocking<T>::KC;
  constexpr int NC = Blocking<T>::NC / NR * NR;
  const int kc_max = std::min(KC, k);
  std::vector<T> pa(static_cast<size_t>(MC) * kc_max);
  std::vector<T> pb(static_cast<size_t>(std::min(NC, (n + NR - 1) / NR * NR)) *
                    kc_max);
  Opd<T> bt = b;  // op(B)^T: rows are the columns of op(B)
  std::swap(bt.rs, bt.cs);

  for (int jc = 0; jc < n; jc += NC) {
    const int nc = std::min(NC, n - jc);
    for (int pc = 0; pc < k; pc += KC) {
      const int kc = std::min(KC, k - pc);
      const T beta_blk = (pc == 0) ? beta : T(1);
      Pack<NR>(bt, jc, nc, pc, kc, pb.data());
      for (int ic = 0; ic < m; ic += MC) {
        const int mc = std::min(MC, m - ic);
        Pack<MR>(a, ic, mc, pc, kc, pa.data());
        for (int jr = 0; jr < nc; jr += NR) {
          const int nr = std::min(NR, nc - jr);
          for (int ir = 0; ir < mc; ir += MR) {
            const int mr = std::min(MR, mc - ir);
            MicroTile(kc, pa.data() + static_cast<size_t>(ir) * kc,
                      pb.data() + static_cast<size_t>(jr) * kc, alpha,
                      beta_blk, c + (ic + ir) + (jc + jr) * ldc, ldc, mr, nr);
          }
        }
      }
    }
  }
}

// C := beta C over m x n; beta == 0 writes zeros without reading
// (src/gemm.cc:48-64).
template <typename T>
void Scale(int m, int n, T beta, T* c, ptrdiff_t ldc) {
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < m; ++i) {
      T& x = c[i + j * ldc];
      x = (beta == T(0)) ? T(0) : x * beta;
    }
  }
}

// Gemm<T> after validation: quick returns, then the blocked driver
// (src/gemm.cc:113-131).
template <typename T>
void Gemm(char ta, char tb, int m, int n, int k, T alpha, const T* a, int lda,
          const T* b, int ldb, T beta, T* c, int ldc) {
  if (m == 0 || n == 0) return;
  if ((k == 0 || alpha == T(0)) && beta == T(1)) return;
  if (k == 0 || alpha == T(0)) {
    Scale(m, n, beta, c, ldc);
    return;
  }
  BlockedGemm(m, n, k, alpha, General(ta, a, lda), General(tb, b, ldb), beta,
              c, ldc);
}

// ---------------------------------------------------------------------------
// TRSM

constexpr int kTrsmBase = 32;  // src/common.h:136

// Solves X T = B in place for the r x r triangle t (r <= kTrsmBase), for
// `rows` independent rows of X; element (row, col) of B is
// b[row * brs + col * bcs]. Columns are solved two per pass in dot form, and
// each is multiplied by the reciprocal of its diagonal
// (src/trsm-inl.h:58-71, :107-174).
template <typename T>
void BaseSolveRight(const Opd<T>& t, int r, int rows, T* b, ptrdiff_t brs,
                    ptrdiff_t bcs) {
  T inv[kTrsmBase];
  for (int j = 0; j < r; ++j) inv[j] = t.unit ? T(1) : T(1) / t.at(j, j);
  const bool scale = !t.unit && g_mistake != 4;
  auto col = [&](int s) { return t.lower ? r - 1 - s : s; };
  for (int row = 0; row < rows; ++row) {
    T* x = b + row * brs;
    auto X = [&](int j) -> T& { return x[j * bcs]; };
    int step = 0;
    for (; step + 2 <= r; step += 2) {
      const int j0 = col(step), j1 = col(step + 1);
      T acc0 = X(j0), acc1 = X(j1);
      for (int s = 0; s < step; ++s) {
        const int q = col(s);
        const T xq = X(q);
        acc0 = Fma(-xq, t.at(q, j0), acc0);
        acc1 = Fma(-xq, t.at(q, j1), acc1);
      }
      if (scale) acc0 *= inv[j0];
      X(j0) = acc0;
      acc1 = Fma(-acc0, t.at(j0, j1), acc1);
      if (scale) acc1 *= inv[j1];
      X(j1) = acc1;
    }
    if (step < r) {
      const int j = col(step);
      T acc = X(j);
      for (int s = 0; s < step; ++s) {
        const int q = col(s);
        acc = Fma(-X(q), t.at(q, j), acc);
      }
      if (scale) acc *= inv[j];
      X(j) = acc;
    }
  }
}

// Base case: the right side solves rows of B directly; the left side solves
// X^T T^T = B^T, i.e. the same kernel with the triangle transposed and the
// roles of B's strides swapped (src/trsm-inl.h:277-325).
template <typename T>
void TrsmBase(bool left, const Opd<T>& tri, int r, int other, T* b,
              ptrdiff_t ldb) {
  if (!left) {
    BaseSolveRight(tri, r, other, b, 1, ldb);
    return;
  }
  Opd<T> tt = tri;
  std::swap(tt.rs, tt.cs);
  tt.lower = !tri.lower;
  BaseSolveRight(tt, r, other, b, ldb, 1);
}

template <typename T>
Opd<T> Sub(const Opd<T>& x, int i0, int j0) {
  Opd<T> s = x;
  s.data += static_cast<ptrdiff_t>(i0) * x.rs + static_cast<ptrdiff_t>(j0) * x.cs;
  return s;
}

template <typename T>
Opd<T> ColMajor(const T* p, ptrdiff_t ld) {
  Opd<T> o;
  o.data = p;
  o.rs = 1;
  o.cs = ld;
  return o;
}

// B -= op1 op2 through the blocked GEMM with alpha = -1, beta = 1
// (src/trsm.cc:69-83).
template <typename T>
void SubtractProduct(int m, int n, int k, const Opd<T>& op1, const Opd<T>& op2,
                     T* b, ptrdiff_t ldb) {
  if (m == 0 || n == 0 || k == 0) return;
  BlockedGemm(m, n, k, T(-1), op1, op2, T(1), b, ldb);
}

// Halving recursion at multiples of kTrsmBase (src/trsm.cc:87-135).
template <typename T>
void TrsmRecursive(bool left, const Opd<T>& tri, int r, int other, T* b,
                   ptrdiff_t ldb) {
  if (r <= kTrsmBase) {
    TrsmBase(left, tri, r, other, b, ldb);
    return;
  }
  const int r1 = std::max(kTrsmBase, (r / 2) / kTrsmBase * kTrsmBase);
  const int r2 = r - r1;
  const Opd<T> t11 = tri;
  const Opd<T> t22 = Sub(tri, r1, r1);
  Opd<T> off = tri.lower ? Sub(tri, r1, 0) : Sub(tri, 0, r1);
  if (left) {
    T* b1 = b;
    T* b2 = b + r1;
    if (tri.lower) {
      TrsmRecursive(left, t11, r1, other, b1, ldb);
      SubtractProduct(r2, other, r1, off, ColMajor(b1, ldb), b2, ldb);
      TrsmRecursive(left, t22, r2, other, b2, ldb);
    } else {
      TrsmRecursive(left, t22, r2, other, b2, ldb);
      SubtractProduct(r1, other, r2, off, ColMajor(b2, ldb), b1, ldb);
      TrsmRecursive(left, t11, r1, other, b1, ldb);
    }
  } else {
    T* b1 = b;
    T* b2 = b + r1 * ldb;
    if (tri.lower) {
      TrsmRecursive(left, t22, r2, other, b2, ldb);
      SubtractProduct(other, r1, r2, ColMajor(b2, ldb), off, b1, ldb);
      TrsmRecursive(left, t11, r1, other, b1, ldb);
    } else {
      TrsmRecursive(left, t11, r1, other, b1, ldb);
      SubtractProduct(other, r2, r1, ColMajor(b1, ldb), off, b2, ldb);
      TrsmRecursive(left, t22, r2, other, b2, ldb);
    }
  }
}

bool IsLower(char c) { return c == 'L' || c == 'l'; }
bool IsLeft(char c) { return c == 'L' || c == 'l'; }
bool IsUnit(char c) { return c == 'U' || c == 'u'; }

// Trsm<T> after validation (src/trsm.cc:170-179): B := alpha B first, then the
// recursion with the structured operand (a transposition flips `lower`).
template <typename T>
void Trsm(char side, char uplo, char ta, char diag, int m, int n, T alpha,
          const T* a, int lda, T* b, int ldb) {
  if (m == 0 || n == 0) return;
  if (alpha != T(1)) Scale(m, n, alpha, b, ldb);
  if (alpha == T(0)) return;
  const bool left = IsLeft(side);
  Opd<T> tri = General(ta, a, lda);
  tri.lower = IsTrans(ta) ? !IsLower(uplo) : IsLower(uplo);
  tri.unit = IsUnit(diag);
  TrsmRecursive(left, tri, left ? m : n, left ? n : m, b, ldb);
}

// ---------------------------------------------------------------------------
// Double-double arithmetic for the references.

struct DD {
  double hi = 0, lo = 0;
};

inline DD TwoSum(double a, double b) {
  const double s = a + b;
  const double bb = s - a;
  return {s, (a - (s - bb)) + (b - bb)};
}

inline DD TwoProd(double a, double b) {
  const double p = a * b;
  return {p, std::fma(a, b, -p)};
}

inline DD Add(DD x, DD y) {
  DD s = TwoSum(x.hi, y.hi);
  double lo = s.lo + x.lo + y.lo;
  if (g_mistake == 5) lo = 0;
  const DD r = TwoSum(s.hi, lo);
  return r;
}

inline DD MulD(DD x, double a) {
  DD p = TwoProd(x.hi, a);
  p.lo = std::fma(x.lo, a, p.lo);
  return TwoSum(p.hi, p.lo);
}

// Expected value and error scale of one GEMM entry.
template <typename T>
void GemmEntry(const Opd<T>& a, const Opd<T>& b, int i, int j, int k, T alpha,
               T beta, const T* c, ptrdiff_t ldc, double* expected,
               double* bound) {
  // Four interleaved double-double chains (p mod 4), combined at the end:
  // the same accuracy as one chain, four times the instruction parallelism.
  DD part[4];
  double abs_part[4] = {0, 0, 0, 0};
  for (int p = 0; p < k; ++p) {
    const double x = static_cast<double>(a.at(i, p));
    const double y = static_cast<double>(b.at(p, j));
    part[p & 3] = Add(part[p & 3], TwoProd(x, y));
    abs_part[p & 3] += std::fabs(x) * std::fabs(y);
  }
  const DD sum = Add(Add(part[0], part[1]), Add(part[2], part[3]));
  const double abs_sum = (abs_part[0] + abs_part[1]) + (abs_part[2] + abs_part[3]);
  DD r = MulD(sum, static_cast<double>(alpha));
  double scale = std::fabs(static_cast<double>(alpha)) * abs_sum;
  if (beta != T(0)) {
    const double cij = static_cast<double>(c[i + j * ldc]);
    r = Add(r, TwoProd(static_cast<double>(beta), cij));
    scale += std::fabs(static_cast<double>(beta)) * std::fabs(cij);
  }
  *expected = r.hi + r.lo;
  *bound = scale + static_cast<double>(std::numeric_limits<T>::min());
}

template <typename T>
void GemmRef(char ta, char tb, int m, int n, int k, T alpha, const T* a,
             int lda, const T* b, int ldb, T beta, const T* c, int ldc,
             double* expected, double* bound) {
  const Opd<T> oa = General(ta, a, lda), ob = General(tb, b, ldb);
  for (int j = 0; j < n; ++j) {
    for (int i = 0; i < m; ++i) {
      GemmEntry(oa, ob, i, j, k, alpha, beta, c, ldc,
                &expected[i + static_cast<size_t>(j) * m],
                &bound[i + static_cast<size_t>(j) * m]);
    }
  }
}

template <typename T>
void GemmRefEntries(char ta, char tb, int k, T alpha, const T* a, int lda,
                    const T* b, int ldb, T beta, const T* c, int ldc,
                    const int* rows, const int* cols, int count,
                    double* expected, double* bound) {
  const Opd<T> oa = General(ta, a, lda), ob = General(tb, b, ldb);
  for (int t = 0; t < count; ++t) {
    GemmEntry(oa, ob, rows[t], cols[t], k, alpha, beta, c, ldc, &expected[t],
              &bound[t]);
  }
}

// Dense op(A) element with the triangle, unit diagonal and unreferenced
// elements of a triangular matrix applied.
template <typename T>
double TriAt(const T* a, int lda, bool lower_stored, bool trans, bool unit,
             int i, int j) {
  // op(A)(i, j) = A(i, j) or A(j, i).
  const int r = trans ? j : i, c = trans ? i : j;
  if (r == c) return unit ? 1.0 : static_cast<double>(a[r + static_cast<size_t>(c) * lda]);
  const bool in = lower_stored ? (r > c) : (r < c);
  return in ? static_cast<double>(a[r + static_cast<size_t>(c) * lda]) : 0.0;
}

template <typename T>
void TrsmResidual(char side, char uplo, char ta, char diag, int m, int n,
                  T alpha, const T* a, int lda, const T* b0, int ldb,
                  const T* x, int ldx, const int* which, int count,
                  double* ratio) {
  const bool left = IsLeft(side), lower = IsLower(uplo), tr = IsTrans(ta);
  const bool unit = IsUnit(diag);
  const int ka = left ? m : n;
  const double tiny = static_cast<double>(std::numeric_limits<T>::min());
  for (int t = 0; t < count; ++t) {
    const int w = which[t];
    double worst = 0;
    if (left) {  // column w of op(A) X - alpha B0
      for (int i = 0; i < m; ++i) {
        DD s;
        double abs_sum = 0;
        for (int p = 0; p < ka; ++p) {
          const double aij = TriAt(a, lda, lower, tr, unit, i, p);
          if (aij == 0.0) continue;
          const double xv = static_cast<double>(x[p + static_cast<size_t>(w) * ldx]);
          s = Add(s, TwoProd(aij, xv));
          abs_sum += std::fabs(aij) * std::fabs(xv);
        }
        const double rhs = static_cast<double>(alpha) *
                           static_cast<double>(b0[i + static_cast<size_t>(w) * ldb]);
        const DD d = Add(s, TwoProd(-static_cast<double>(alpha),
                                    static_cast<double>(b0[i + static_cast<size_t>(w) * ldb])));
        const double err = std::fabs(d.hi + d.lo);
        const double scale = abs_sum + std::fabs(rhs) + tiny;
        worst = std::max(worst, err / scale);
        if (err != err) worst = INFINITY;
      }
    } else {  // row w of X op(A) - alpha B0
      for (int j = 0; j < n; ++j) {
        DD s;
        double abs_sum = 0;
        for (int p = 0; p < ka; ++p) {
          const double aij = TriAt(a, lda, lower, tr, unit, p, j);
          if (aij == 0.0) continue;
          const double xv = static_cast<double>(x[w + static_cast<size_t>(p) * ldx]);
          s = Add(s, TwoProd(xv, aij));
          abs_sum += std::fabs(aij) * std::fabs(xv);
        }
        const double bv = static_cast<double>(b0[w + static_cast<size_t>(j) * ldb]);
        const DD d = Add(s, TwoProd(-static_cast<double>(alpha), bv));
        const double err = std::fabs(d.hi + d.lo);
        const double scale = abs_sum + std::fabs(static_cast<double>(alpha) * bv) + tiny;
        worst = std::max(worst, err / scale);
        if (err != err) worst = INFINITY;
      }
    }
    ratio[t] = worst;
  }
}

// ---------------------------------------------------------------------------
// Emulated DGEMM (Ozaki scheme II; Ozaki, Uchino, Imamura 2025), the CPU
// restatement of the product's INT8 emulation, step for step.

namespace ozk {

constexpr int kMax = 24, kLimbs = 4;  // 160-bit fractions in 40-bit limbs
constexpr int kModuli[kMax] = {256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217,
                               211, 199, 197, 193, 191, 181, 179, 173, 167, 163, 157, 151};

int64_t ModInv(int64_t a, int64_t m) {
  for (int64_t y = 1; y < m; ++y)
    if ((a * y) % m == 1) return y;
  return 0;
}

struct Crt {
  int nmod;
  double P;
  double frac[kMax][kLimbs];  // floor(2^160 y_l / p_l), 40-bit limbs, low first
};

Crt Make(int nmod) {
  Crt c{};
  c.nmod = nmod;
  c.P = 1;
  for (int l = 0; l < nmod; ++l) c.P *= kModuli[l];
  for (int l = 0; l < nmod; ++l) {
    const int64_t p = kModuli[l];
    int64_t ml = 1;
    for (int q = 0; q < nmod; ++q)
      if (q != l) ml = (ml * kModuli[q]) % p;
    uint64_t rem = uint64_t(ModInv(ml, p));
    for (int t = kLimbs - 1; t >= 0; --t) {
      const uint64_t num = rem << 40;
      c.frac[l][t] = double(num / uint64_t(p));
      rem = num % uint64_t(p);
    }
  }
  return c;
}

int TotalBits(int nmod, int64_t k) {
  double lp = 0;
  for (int l = 0; l < nmod; ++l) lp += std::log2(double(kModuli[l]));
  return int(std::floor(lp - 1.0 - std::log2(double(k)) - 1e-9));
}

// x mod p in [0, p).
int64_t ModU(int64_t x, int64_t p) {
  int64_t r = x % p;
  return r < 0 ? r + p : r;
}

// Symmetric residue of round(v) (v = x 2^s), as the product computes it.
int Residue(double v, int p) {
  if (std::fabs(v) < 0x1p52) {
    const double xi = std::nearbyint(v);
    int r = int(ModU(int64_t(xi), p));
    if (r >= p / 2) r -= p;
    return r;
  }
  int e;
  const double m = std::frexp(v, &e);
  const int64_t mi = int64_t(std::ldexp(m, 53));
  int64_t r = ModU(mi, p);
  for (int E = e - 53; E > 0; --E) r = (r * 2) % p;
  if (r >= p / 2) r -= p;
  return int(r);
}

// The centred integer with residues res[l], as a double: the fixed-point
// sum S = sum_l res[l] F_l (mod 2^160) in four 40-bit limbs accumulated
// exactly in doubles, carries propagated, the top limb wrapped and centred,
// then (S / 2^160) P.
double Reconstruct(const Crt& c, const int* res) {
  double S[kLimbs] = {};
  const int nuse = (g_mistake == 6) ? c.nmod - 1 : c.nmod;
  for (int l = 0; l < nuse; ++l)
    for (int t = 0; t < kLimbs; ++t) S[t] = std::fma(double(res[l]), c.frac[l][t], S[t]);
  for (int t = 0; t + 1 < kLimbs; ++t) {
    const double carry = std::floor(S[t] * 0x1p-40);
    S[t] = std::fma(-carry, 0x1p40, S[t]);
    S[t + 1] += carry;
  }
  S[kLimbs - 1] = std::fma(-std::floor(S[kLimbs - 1] * 0x1p-40), 0x1p40, S[kLimbs - 1]);
  const bool neg = S[kLimbs - 1] >= 0x1p39;  // fraction >= 1/2: negative
  if (neg) {  // 2^160 - S exactly: limb complement plus one
    double carry = 1;
    for (int t = 0; t < kLimbs; ++t) {
      S[t] = (0x1p40 - 1) - S[t] + carry;
      carry = (S[t] >= 0x1p40) ? 1.0 : 0.0;
      S[t] -= carry * 0x1p40;
    }
  }
  double v = S[0] * 0x1p-160;
  v = v + S[1] * 0x1p-120;
  v = v + S[2] * 0x1p-80;
  v = v + S[3] * 0x1p-40;
  v = v * c.P;
  return neg ? -v : v;
}

// Scale exponent of a row with max |x| = f 2^E (f in [0.5, 1)): bits - E.
int ScaleExp(double rowmax, int bits) {
  if (!(rowmax > 0)) return 0;
  int e;
  std::frexp(rowmax, &e);
  return bits - e;
}

void Dgemm(char ta, char tb, int m, int n, int k, double alpha, const double* a, int lda,
           const double* b, int ldb, double beta, double* c, int ldc, int nmod) {
  if (m == 0 || n == 0) return;
  if ((k == 0 || alpha == 0.0) && beta == 1.0) return;
  if (k == 0 || alpha == 0.0) {
    Scale(m, n, beta, c, ldc);
    return;
  }
  const Opd<double> A = General(ta, a, lda), B = General(tb, b, ldb);
  const Crt crt = Make(nmod);
  const int bits = std::max(TotalBits(nmod, k), 2);
  const int ba = bits / 2, bb = bits - ba;
  std::vector<int> sa(m), sb(n);
  for (int i = 0; i < m; ++i) {
    double mx = 0;
    for (int p = 0; p < k; ++p) mx = std::max(mx, std::fabs(A.at(i, p)));
    sa[i] = ScaleExp(mx, ba);
  }
  for (int j = 0; j < n; ++j) {
    double mx = 0;
    for (int p = 0; p < k; ++p) mx = std::max(mx, std::fabs(B.at(p, j)));
    sb[j] = ScaleExp(mx, bb);
  }
  // Residue slices: ra[l][i][p], rb[l][j][p].
  std::vector<int8_t> ra(size_t(nmod) * m * k), rb(size_t(nmod) * n * k);
  for (int l = 0; l < nmod; ++l) {
    const int p_l = kModuli[l];
    for (int i = 0; i < m; ++i)
      for (int p = 0; p < k; ++p) {
        const double v = std::ldexp(A.at(i, p), sa[i]);
        ra[(size_t(l) * m + i) * k + p] = int8_t(v == 0.0 ? 0 : Residue(v, p_l));
      }
    for (int j = 0; j < n; ++j)
      for (int p = 0; p < k; ++p) {
        const double v = std::ldexp(B.at(p, j), sb[j]);
        rb[(size_t(l) * n + j) * k + p] = int8_t(v == 0.0 ? 0 : Residue(v, p_l));
      }
  }
  std::vector<int> res(nmod);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      for (int l = 0; l < nmod; ++l) {
        int64_t s = 0;
        const int8_t* x = &ra[(size_t(l) * m + i) * k];
        const int8_t* y = &rb[(size_t(l) * n + j) * k];
        for (int p = 0; p < k; ++p) s += int64_t(x[p]) * y[p];
        res[l] = int(ModU(s, kModuli[l]));
      }
      const double x = std::ldexp(Reconstruct(crt, res.data()), -(sa[i] + sb[j]));
      double& cij = c[i + size_t(j) * ldc];
      double v = x * alpha;
      if (beta != 0.0) v = (beta == 1.0) ? v + cij : std::fma(cij, beta, v);
      cij = v;
    }
}

}  // namespace ozk

}  // namespace

extern "C" {

uint64_t orc_splitmix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

double orc_uniform01(uint64_t seed, uint64_t index) {
  const uint64_t s = orc_splitmix64(seed ^ 0x5851F42D4C957F2Dull);
  return static_cast<double>(orc_splitmix64(s + index) >> 11) * 0x1.0p-53;
}

void orc_fill_uniform_f64(double* x, int64_t rows, int64_t cols, int64_t ld,
                          uint64_t seed, double lo, double hi) {
  const double w = hi - lo;
  for (int64_t j = 0; j < cols; ++j)
    for (int64_t i = 0; i < rows; ++i)
      x[i + j * ld] = std::fma(orc_uniform01(seed, static_cast<uint64_t>(i + j * rows)), w, lo);
}

// The distribution "u 2^e" of the emulated-DGEMM workloads; identical bit
// for bit to the library's device fill (hwyblas_amd_fill_pow2).
void orc_fill_pow2(double* x, int64_t rows, int64_t cols, int64_t ld, uint64_t seed, int emin,
                   int emax) {
  const uint64_t key = orc_splitmix64(seed ^ 0x5851F42D4C957F2Dull);
  const uint64_t key2 = orc_splitmix64(seed ^ 0x2545F4914F6CDD1Dull);
  const uint64_t span = uint64_t(emax - emin + 1);
  for (int64_t j = 0; j < cols; ++j)
    for (int64_t i = 0; i < rows; ++i) {
      const uint64_t idx = uint64_t(i + j * rows);
      const uint64_t w = orc_splitmix64(key + idx);
      const uint64_t w2 = orc_splitmix64(key2 + idx);
      const double u = 1.0 + double(w >> 12) * 0x1.0p-52;
      const double v = std::ldexp(u, emin + int((w2 >> 32) % span));
      x[i + j * ld] = (w & 1) ? -v : v;
    }
}

void orc_fill_uniform_f32(float* x, int64_t rows, int64_t cols, int64_t ld,
                          uint64_t seed, double lo, double hi) {
  const double w = hi - lo;
  for (int64_t j = 0; j < cols; ++j)
    for (int64_t i = 0; i < rows; ++i)
      x[i + j * ld] = static_cast<float>(
          std::fma(orc_uniform01(seed, static_cast<uint64_t>(i + j * rows)), w, lo));
}

void orc_dgemm(char ta, char tb, int m, int n, int k, double alpha,
               const double* a, int lda, const double* b, int ldb, double beta,
               double* c, int ldc) {
  Gemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}
void orc_sgemm(char ta, char tb, int m, int n, int k, float alpha,
               const float* a, int lda, const float* b, int ldb, float beta,
               float* c, int ldc) {
  Gemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}
void orc_dtrsm(char side, char uplo, char ta, char diag, int m, int n,
               double alpha, const double* a, int lda, double* b, int ldb) {
  Trsm(side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb);
}
void orc_strsm(char side, char uplo, char ta, char diag, int m, int n,
               float alpha, const float* a, int lda, float* b, int ldb) {
  Trsm(side, uplo, ta, diag, m, n, alpha, a, lda, b, ldb);
}

void orc_dgemm_ref(char ta, char tb, int m, int n, int k, double alpha,
                   const double* a, int lda, const double* b, int ldb,
                   double beta, const double* c, int ldc, double* expected,
                   double* bound) {
  GemmRef(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, expected, bound);
}
void orc_sgemm_ref(char ta, char tb, int m, int n, int k, float alpha,
                   const float* a, int lda, const float* b, int ldb, float beta,
                   const float* c, int ldc, double* expected, double* bound) {
  GemmRef(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, expected, bound);
}
void orc_dgemm_ref_entries(char ta, char tb, int k, double alpha,
                           const double* a, int lda, const double* b, int ldb,
                           double beta, const double* c, int ldc,
                           const int* rows, const int* cols, int count,
                           double* expected, double* bound) {
  GemmRefEntries(ta, tb, k, alpha, a, lda, b, ldb, beta, c, ldc, rows, cols,
                 count, expected, bound);
}
void orc_sgemm_ref_entries(char ta, char tb, int k, float alpha,
                           const float* a, int lda, const float* b, int ldb,
                           float beta, const float* c, int ldc,
                           const int* rows, const int* cols, int count,
                           double* expected, double* bound) {
  GemmRefEntries(ta, tb, k, alpha, a, lda, b, ldb, beta, c, ldc, rows, cols,
                 count, expected, bound);
}
void orc_dtrsm_residual(char side, char uplo, char ta, char diag, int m, int n,
                        double alpha, const double* a, int lda,
                        const double* b0, int ldb, const double* x, int ldx,
                        const int* which, int count, double* ratio) {
  TrsmResidual(side, uplo, ta, diag, m, n, alpha, a, lda, b0, ldb, x, ldx,
               which, count, ratio);
}
void orc_strsm_residual(char side, char uplo, char ta, char diag, int m, int n,
                        float alpha, const float* a, int lda, const float* b0,
                        int ldb, const float* x, int ldx, const int* which,
                        int count, double* ratio) {
  TrsmResidual(side, uplo, ta, diag, m, n, alpha, a, lda, b0, ldb, x, ldx,
               which, count, ratio);
}

int orc_ozaki_modulus(int l) { return (l >= 0 && l < ozk::kMax) ? ozk::kModuli[l] : 0; }

int orc_ozaki_total_bits(int nmod, int64_t k) { return ozk::TotalBits(nmod, k); }

double orc_ozaki_crt(int nmod, const int* res) {
  return ozk::Reconstruct(ozk::Make(nmod), res);
}

void orc_dgemm_ozaki(char ta, char tb, int m, int n, int k, double alpha, const double* a,
                     int lda, const double* b, int ldb, double beta, double* c, int ldc,
                     int nmod) {
  ozk::Dgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, nmod);
}

void orc_plant_mistake(int which) { g_mistake = which; }

}  // extern "C"
