// CPU restatement of the product's BF16-slice emulated DGEMM (Ozaki scheme
// I: Ozaki, Ogita, Oishi, Rump, Numer. Algorithms 59, 2012), step for step
// (TEST INFRASTRUCTURE ONLY; see oracle.h). hwyBLAS has no emulated GEMM;
// the call it stands in for is dgemm_ (src/gemm.cc:73-132), whose argument
// rules and quick returns (alpha = 0 or k = 0: C = beta C) it keeps.
//
// Digits are held as integers and the chunk sums in int64, so every step
// but the FP64 accumulation is exact; the FP64 accumulation follows the
// product's order (levels from s + 1 down to 2, t ascending, k chunks of
// 1024 ascending, the last one of a pair shorter), so the result is the
// product's bit for bit.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

#include "oracle.h"

extern int g_mistake;

namespace {

constexpr int kMaxSlices = 16;
constexpr int kChunk = 1024;

// -(E + 1) for a row whose largest magnitude is f 2^E (f in [0.5, 1)), so
// the scaled row is below 1/2 in magnitude; 0 for a zero row.
int ScaleExp(double rowmax) {
  if (!(rowmax > 0)) return 0;
  int e;
  std::frexp(rowmax, &e);
  return -(e + 1);
}

// Balanced digits d_1..d_s of x 2^sc: d_t = nearest integer to 256 r_{t-1}
// (ties to even), r_t = 256 r_{t-1} - d_t, r_0 = x 2^sc.
void Digits(double x, int sc, int slices, int* d) {
  double r = std::ldexp(x, sc);
  for (int t = 0; t < slices; ++t) {
    const double y = r * 256.0;
    const double q = std::nearbyint(y);
    r = y - q;
    d[t] = int(q);
  }
}

}  // namespace

extern "C" {

void orc_bf16_digits(double x, int sc, int slices, int* digits) {
  Digits(x, sc, std::min(std::max(slices, 0), kMaxSlices), digits);
}

void orc_dgemm_ozaki_bf16(char ta, char tb, int m, int n, int k, double alpha, const double* a,
                          int lda, const double* b, int ldb, double beta, double* c, int ldc,
                          int slices) {
  if (m == 0 || n == 0) return;
  if ((k == 0 || alpha == 0.0) && beta == 1.0) return;
  if (k == 0 || alpha == 0.0) {
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < m; ++i) {
        double& x = c[i + int64_t(j) * ldc];
        x = (beta == 0.0) ? 0.0 : beta * x;
      }
    return;
  }
  slices = std::min(std::max(slices, 1), kMaxSlices);
  const bool at = ta == 'T' || ta == 't' || ta == 'C' || ta == 'c';
  const bool bt = tb == 'T' || tb == 't' || tb == 'C' || tb == 'c';
  auto A = [&](int i, int p) { return at ? a[p + int64_t(i) * lda] : a[i + int64_t(p) * lda]; };
  auto B = [&](int p, int j) { return bt ? b[j + int64_t(p) * ldb] : b[p + int64_t(j) * ldb]; };

  std::vector<int> sa(m), sb(n);
  for (int i = 0; i < m; ++i) {
    double mx = 0;
    for (int p = 0; p < k; ++p) mx = std::max(mx, std::fabs(A(i, p)));
    sa[i] = ScaleExp(mx);
  }
  for (int j = 0; j < n; ++j) {
    double mx = 0;
    for (int p = 0; p < k; ++p) mx = std::max(mx, std::fabs(B(p, j)));
    sb[j] = ScaleExp(mx);
  }
  // da[(t * m + i) * k + p], db[(t * n + j) * k + p].
  std::vector<int16_t> da(size_t(slices) * m * k), db(size_t(slices) * n * k);
  int d[kMaxSlices];
  for (int i = 0; i < m; ++i)
    for (int p = 0; p < k; ++p) {
      Digits(A(i, p), sa[i], slices, d);
      for (int t = 0; t < slices; ++t) da[(size_t(t) * m + i) * k + p] = int16_t(d[t]);
    }
  for (int j = 0; j < n; ++j)
    for (int p = 0; p < k; ++p) {
      Digits(B(p, j), sb[j], slices, d);
      for (int t = 0; t < slices; ++t) db[(size_t(t) * n + j) * k + p] = int16_t(d[t]);
    }
  // Pairs (t, u), 1-based, by level from s + 1 down to 2, t ascending.
  // (Planted mistake 8: the top level t + u = s + 1 is left out.)
  struct Pair {
    int t, u, level;
  };
  std::vector<Pair> pairs;
  const int top = (g_mistake == 8) ? slices : slices + 1;
  for (int l = top; l >= 2; --l)
    for (int t = 1; t <= l - 1; ++t) pairs.push_back({t, l - t, l});

  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      double acc = 0.0;
      for (const Pair& pr : pairs) {
        const int16_t* x = &da[(size_t(pr.t - 1) * m + i) * k];
        const int16_t* y = &db[(size_t(pr.u - 1) * n + j) * k];
        const double scale = std::ldexp(1.0, -8 * pr.level);
        for (int p0 = 0; p0 < k; p0 += kChunk) {
          int64_t s = 0;
          for (int p = p0; p < std::min(k, p0 + kChunk); ++p) s += int64_t(x[p]) * y[p];
          acc = std::fma(double(s), scale, acc);
        }
      }
      const double v0 = std::ldexp(acc, -(sa[i] + sb[j]));
      double* cij = c + i + int64_t(j) * ldc;
      double v = v0 * alpha;
      if (beta != 0.0) v = (beta == 1.0) ? v + *cij : std::fma(*cij, beta, v);
      *cij = v;
    }
}

}  // extern "C"
