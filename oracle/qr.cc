// Householder QR (xGEQRF and its pieces), restated on the CPU after
// hwyBLAS's LAPACK port. Test infrastructure only.
//
// - Nrm2: the sum of squares in double, square-rooted when it lies in
//   [2^-500, DBL_MAX]; otherwise a scaled sum (src/level1.cc:318-353 uses
//   Blue's method there; both give the norm to a few ulps).
// - Lapy2 (src/lapack/auxiliary.cc:254-276): sqrt(x^2 + y^2) as
//   w sqrt(1 + (z / w)^2), w = max(|x|, |y|), z = min.
// - Larfg (src/lapack/householder.cc:44-104): beta = -sign(alpha)
//   lapy2(alpha, |x|), tau = (beta - alpha) / beta, x := x / (alpha - beta),
//   alpha := beta; with the rescaling by 1/safmin (safmin = tiny / eps) while
//   |beta| < safmin, at most 20 times, and tau = 0 when x = 0.
// - Larf1f (householder.cc:160-220), left side: w = C(2:lastv, :)^T v(2:) +
//   C(1, :), C(1, :) -= tau w, C(2:lastv, :) -= tau v w^T, with lastv the
//   last non-zero of v and lastc the last non-zero column of C(1:lastv, :).
// - Geqr2 (src/lapack/geqrf.cc:28-55): column by column, Larfg then Larf1f
//   on the columns to the right.
// - Larft, forward and columnwise (householder.cc:336-440): recursive,
//   T11 and T22 for the halves (l = k / 2), T12 = -T11 (V1^T V2) T22.
// - Larfb, left, transposed, forward, columnwise (householder.cc:681-730):
//   W = C^T V; W := W T; C2 -= V2 W^T; W := W V1^T; C1 -= W^T.
// - Geqrf (geqrf.cc:66-150): NB = 32, NX = 128, blocked while i <= k - NX:
//   Geqr2 on the panel, then Larft and Larfb on the columns to its right;
//   Geqr2 on the rest.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>
#include <limits>
#include <vector>

#include "oracle.h"

extern int g_mistake;

namespace {

struct Mat {
  double* p;
  int64_t ld;
  double& operator()(int64_t i, int64_t j) const { return p[i + j * ld]; }  // 0-based
};

double Nrm2(int64_t n, const double* x) {
  if (n < 1) return 0.0;
  double s = 0.0;
  for (int64_t i = 0; i < n; ++i) s = std::fma(x[i], x[i], s);
  if (s >= std::ldexp(1.0, -500) && s <= DBL_MAX) return std::sqrt(s);
  double mx = 0.0;
  for (int64_t i = 0; i < n; ++i) mx = std::max(mx, std::fabs(x[i]));
  if (mx == 0.0 || !(mx <= DBL_MAX)) return mx;
  double t = 0.0;
  for (int64_t i = 0; i < n; ++i) {
    const double y = x[i] / mx;
    t = std::fma(y, y, t);
  }
  return mx * std::sqrt(t);
}

double Lapy2(double x, double y) {
  if (std::isnan(x)) return x;
  if (std::isnan(y)) return y;
  const double w = std::max(std::fabs(x), std::fabs(y));
  const double z = std::min(std::fabs(x), std::fabs(y));
  if (z == 0.0 || w > DBL_MAX) return w;
  return w * std::sqrt(1.0 + (z / w) * (z / w));
}

// n: length of (alpha, x); x has n - 1 entries.
void Larfg(int64_t n, double* alpha, double* x, double* tau) {
  if (n <= 1) {
    *tau = 0.0;
    return;
  }
  double xnorm = Nrm2(n - 1, x);
  if (xnorm == 0.0) {
    *tau = 0.0;
    return;
  }
  double alph = *alpha;
  double beta = -std::copysign(Lapy2(alph, xnorm), alph);
  const double safmin = std::numeric_limits<double>::min() / (0x1p-53);
  const double rsafmn = 1.0 / safmin;
  int knt = 0;
  if (std::fabs(beta) < safmin) {
    do {
      ++knt;
      for (int64_t i = 0; i < n - 1; ++i) x[i] *= rsafmn;
      beta *= rsafmn;
      alph *= rsafmn;
    } while (std::fabs(beta) < safmin && knt < 20);
    xnorm = Nrm2(n - 1, x);
    beta = -std::copysign(Lapy2(alph, xnorm), alph);
  }
  *tau = (beta - alph) / beta;
  const double scal = 1.0 / (alph - beta);
  for (int64_t i = 0; i < n - 1; ++i) x[i] *= scal;
  for (int j = 0; j < knt; ++j) beta *= safmin;
  *alpha = beta;
}

// H C from the left, H = I - tau v v^T, v(0) = 1 (not read), v(1:) = vtail;
// C is m x n.
void Larf1fLeft(int64_t m, int64_t n, const double* vtail, double tau, Mat c) {
  if (tau == 0.0) return;
  int64_t lastv = m;
  while (lastv > 1 && vtail[lastv - 2] == 0.0) --lastv;
  int64_t lastc = n;
  while (lastc > 0) {
    bool zero = true;
    for (int64_t i = 0; i < lastv && zero; ++i) zero = c(i, lastc - 1) == 0.0;
    if (!zero) break;
    --lastc;
  }
  if (lastc == 0) return;
  if (lastv == 1) {
    for (int64_t j = 0; j < lastc; ++j) c(0, j) *= 1.0 - tau;
    return;
  }
  std::vector<double> w(lastc);
  for (int64_t j = 0; j < lastc; ++j) {
    double s = 0.0;
    for (int64_t i = 1; i < lastv; ++i) s += c(i, j) * vtail[i - 1];
    w[j] = s + c(0, j);
  }
  for (int64_t j = 0; j < lastc; ++j) c(0, j) -= tau * w[j];
  for (int64_t j = 0; j < lastc; ++j) {
    const double t = -tau * w[j];
    for (int64_t i = 1; i < lastv; ++i) c(i, j) += vtail[i - 1] * t;
  }
}

void Geqr2(int64_t m, int64_t n, Mat a, double* tau) {
  const int64_t k = std::min(m, n);
  for (int64_t i = 0; i < k; ++i) {
    double* x = &a(std::min(i + 1, m - 1), i);
    Larfg(m - i, &a(i, i), x, &tau[i]);
    if (i + 1 < n) {
      // Planted mistake 15: the reflector applied with tau of the wrong sign.
      const double t = g_mistake == 15 ? -tau[i] : tau[i];
      Larf1fLeft(m - i, n - i - 1, &a(i + 1, i), t, Mat{&a(i, i + 1), a.ld});
    }
  }
}

// V (n x k, unit lower trapezoidal: V(i, i) = 1, zeros above; the stored
// entries on and above the diagonal are not read).
double VAt(const Mat& v, int64_t i, int64_t j) {
  return i == j ? 1.0 : (i > j ? v(i, j) : 0.0);
}

// T (k x k, upper) of the block reflector H(0) ... H(k-1), forward columnwise.
void Larft(int64_t n, int64_t k, Mat v, const double* tau, Mat t) {
  if (n == 0 || k == 0) return;
  if (n == 1 || k == 1) {
    t(0, 0) = tau[0];
    return;
  }
  const int64_t l = k / 2;
  Larft(n, l, v, tau, t);
  Larft(n - l, k - l, Mat{&v(l, l), v.ld}, tau + l, Mat{&t(l, l), t.ld});
  // T12 = V1^T V2 over the rows where V2 is not structurally zero (rows >= l).
  std::vector<double> g(size_t(l) * (k - l));
  for (int64_t c = 0; c < k - l; ++c)
    for (int64_t r = 0; r < l; ++r) {
      double s = 0.0;
      for (int64_t i = l; i < n; ++i) s += VAt(v, i, r) * VAt(v, i, l + c);
      g[r + c * l] = s;
    }
  // T12 = -T11 G T22.
  std::vector<double> h(size_t(l) * (k - l));
  for (int64_t c = 0; c < k - l; ++c)
    for (int64_t r = 0; r < l; ++r) {
      double s = 0.0;
      for (int64_t q = r; q < l; ++q) s += t(r, q) * g[q + c * l];
      h[r + c * l] = -s;
    }
  for (int64_t c = 0; c < k - l; ++c)
    for (int64_t r = 0; r < l; ++r) {
      double s = 0.0;
      for (int64_t q = 0; q <= c; ++q) s += h[r + q * l] * t(l + q, l + c);
      t(r, l + c) = s;
    }
}

// C := H^T C = C - V T^T V^T C, C m x n, V m x k.
void LarfbLeftTrans(int64_t m, int64_t n, int64_t k, Mat v, Mat t, Mat c) {
  if (m <= 0 || n <= 0) return;
  std::vector<double> w(size_t(n) * k), w2(size_t(n) * k);
  for (int64_t j = 0; j < k; ++j)
    for (int64_t i = 0; i < n; ++i) {
      double s = 0.0;
      for (int64_t p = j; p < m; ++p) s += c(p, i) * VAt(v, p, j);
      w[i + j * n] = s;
    }
  for (int64_t j = 0; j < k; ++j)  // W := W T
    for (int64_t i = 0; i < n; ++i) {
      double s = 0.0;
      for (int64_t q = 0; q <= j; ++q) s += w[i + q * n] * t(q, j);
      w2[i + j * n] = s;
    }
  for (int64_t i = 0; i < n; ++i)  // C -= V W^T
    for (int64_t p = 0; p < m; ++p) {
      double s = 0.0;
      for (int64_t j = 0; j < std::min<int64_t>(k, p + 1); ++j) s += VAt(v, p, j) * w2[i + j * n];
      c(p, i) -= s;
    }
}

int Geqrf(int64_t m, int64_t n, Mat a, double* tau) {
  const int64_t k = std::min(m, n);
  if (k == 0) return 0;
  const int64_t nb = 32, nx = 128;
  int64_t i = 0;
  if (nb < k && nx < k) {
    std::vector<double> t(size_t(nb) * nb);
    for (; i < k - nx; i += nb) {
      const int64_t ib = std::min(k - i, nb);
      Geqr2(m - i, ib, Mat{&a(i, i), a.ld}, tau + i);
      if (i + ib < n) {
        std::fill(t.begin(), t.end(), 0.0);
        Larft(m - i, ib, Mat{&a(i, i), a.ld}, tau + i, Mat{t.data(), nb});
        LarfbLeftTrans(m - i, n - i - ib, ib, Mat{&a(i, i), a.ld}, Mat{t.data(), nb},
                       Mat{&a(i, i + ib), a.ld});
      }
    }
  }
  if (i < k) Geqr2(m - i, n - i, Mat{&a(i, i), a.ld}, tau + i);
  return 0;
}

}  // namespace

extern "C" {

int orc_dgeqrf(int m, int n, double* a, int lda, double* tau) {
  if (m < 0) return -1;
  if (n < 0) return -2;
  if (lda < std::max(1, m)) return -4;
  return Geqrf(m, n, Mat{a, lda}, tau);
}

void orc_dgeqr2(int m, int n, double* a, int lda, double* tau) {
  Geqr2(m, n, Mat{a, lda}, tau);
}

void orc_dlarft(int n, int k, const double* v, int ldv, const double* tau, double* t, int ldt) {
  Larft(n, k, Mat{const_cast<double*>(v), ldv}, tau, Mat{t, ldt});
}

}  // extern "C"
