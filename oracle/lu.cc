// CPU restatement of hwyBLAS's LU with partial pivoting, its solve and the
// backward error of gerfs (TEST INFRASTRUCTURE ONLY; see oracle.h).
//
// Follows src/lapack/getrf.cc (xLASWP :28-81, xGETRF2 :112-207, xGETRF
// :216-280, xGETRS :287-339) with hwyBLAS's leaf kernel
// (src/lapack/factor_kernels.cc:107-138, leaf 16 rows or columns,
// factor_kernels.h:27), GETRF's block size 64 (src/lapack/environment.cc:
// 86-88), the componentwise backward error of xGERFS
// (src/lapack/gesvx.cc:395-425), the whole of xGERFS with its forward
// error bound (src/lapack/gesvx.cc:333-521) and the norm estimator xLACN2
// it drives (src/lapack/trsen.cc:2235-2366). GEMM and TRSM are the
// oracle's own restatements (oracle.cc).
#include <algorithm>
#include <cmath>
#include <limits>
#include <utility>
#include <vector>

#include "oracle.h"

extern int g_mistake;

namespace {

constexpr int kGetrfLeaf = 16;
constexpr int kGetrfNb = 64;

// Index of the first entry of largest magnitude; 0 when the first entry is
// NaN; other NaNs are skipped (src/level1-inl.h:353-410, kMax = true).
int IamaxFirst(int n, const double* x) {
  const double first = std::fabs(x[0]);
  if (std::isnan(first)) return 0;
  int best = 0;
  double bv = first;
  for (int i = 1; i < n; ++i) {
    const double v = std::fabs(x[i]);
    if (v > bv) {
      bv = v;
      best = i;
    }
  }
  return best;
}

// src/lapack/factor_kernels.cc:107-138 (0-based, ipiv 1-based).
int GetrfLeaf(int m, int n, double* a, int lda, int* ipiv) {
  auto at = [&](int i, int j) -> double& { return a[i + int64_t(j) * lda]; };
  const double sfmin = std::numeric_limits<double>::min();
  int info = 0;
  for (int j = 0; j < std::min(m, n); ++j) {
    const int p = j + IamaxFirst(m - j, &at(j, j));
    ipiv[j] = p + 1;
    const double pivot = at(p, j);
    if (pivot != 0.0) {
      if (p != j)  // (planted mistake 7: the columns left of j are not swapped)
        for (int k = g_mistake == 7 ? j : 0; k < n; ++k) std::swap(at(j, k), at(p, k));
      if (std::fabs(pivot) >= sfmin) {
        const double r = 1.0 / pivot;
        for (int i = j + 1; i < m; ++i) at(i, j) = at(i, j) * r;
      } else {
        for (int i = j + 1; i < m; ++i) at(i, j) = at(i, j) / pivot;
      }
    } else if (info == 0) {
      info = j + 1;
    }
    for (int k = j + 1; k < n; ++k) {
      const double alpha = -at(j, k);
      for (int i = j + 1; i < m; ++i) at(i, k) = std::fma(alpha, at(i, j), at(i, k));
    }
  }
  return info;
}

// src/lapack/getrf.cc:28-81 with 1-based k1, k2 (the 32-column blocking
// there does not change the result).
void Laswp(int n, double* a, int lda, int k1, int k2, const int* ipiv, int incx) {
  if (incx == 0) return;
  auto at = [&](int i, int j) -> double& { return a[(i - 1) + int64_t(j - 1) * lda]; };
  if (incx > 0) {
    for (int i = k1, ix = k1; i <= k2; ++i, ix += incx) {
      const int ip = ipiv[ix - 1];
      if (ip != i)
        for (int k = 1; k <= n; ++k) std::swap(at(i, k), at(ip, k));
    }
  } else {
    for (int i = k2, ix = k1 + (k1 - k2) * incx; i >= k1; --i, ix += incx) {
      const int ip = ipiv[ix - 1];
      if (ip != i)
        for (int k = 1; k <= n; ++k) std::swap(at(i, k), at(ip, k));
    }
  }
}

// src/lapack/getrf.cc:112-207.
int Getrf2(int m, int n, double* a, int lda, int* ipiv) {
  if (m == 0 || n == 0) return 0;
  if (std::min(m, n) <= kGetrfLeaf) return GetrfLeaf(m, n, a, lda, ipiv);
  int info = 0;
  const int n1 = std::min(m, n) / 2;
  const int n2 = n - n1;
  int iinfo = Getrf2(m, n1, a, lda, ipiv);
  if (info == 0 && iinfo > 0) info = iinfo;
  double* a12 = a + int64_t(n1) * lda;
  Laswp(n2, a12, lda, 1, n1, ipiv, 1);
  orc_dtrsm('L', 'L', 'N', 'U', n1, n2, 1.0, a, lda, a12, lda);
  orc_dgemm('N', 'N', m - n1, n2, n1, -1.0, a + n1, lda, a12, lda, 1.0, a12 + n1, lda);
  iinfo = Getrf2(m - n1, n2, a12 + n1, lda, ipiv + n1);
  if (info == 0 && iinfo > 0) info = iinfo + n1;
  for (int i = n1; i < std::min(m, n); ++i) ipiv[i] += n1;
  Laswp(n1, a, lda, n1 + 1, std::min(m, n), ipiv, 1);
  return info;
}

}  // namespace

extern "C" {

void orc_dlaswp(int n, double* a, int lda, int k1, int k2, const int* ipiv, int incx) {
  Laswp(n, a, lda, k1, k2, ipiv, incx);
}

// src/lapack/getrf.cc:216-280.
int orc_dgetrf(int m, int n, double* a, int lda, int* ipiv) {
  if (m == 0 || n == 0) return 0;
  const int nb = kGetrfNb;
  if (nb >= std::min(m, n)) return Getrf2(m, n, a, lda, ipiv);
  int info = 0;
  auto ptr = [&](int i, int j) { return a + (i - 1) + int64_t(j - 1) * lda; };
  for (int j = 1; j <= std::min(m, n); j += nb) {
    const int jb = std::min(std::min(m, n) - j + 1, nb);
    const int iinfo = Getrf2(m - j + 1, jb, ptr(j, j), lda, ipiv + (j - 1));
    if (info == 0 && iinfo > 0) info = iinfo + j - 1;
    for (int i = j; i <= std::min(m, j + jb - 1); ++i) ipiv[i - 1] += j - 1;
    Laswp(j - 1, a, lda, j, j + jb - 1, ipiv, 1);
    if (j + jb <= n) {
      Laswp(n - j - jb + 1, ptr(1, j + jb), lda, j, j + jb - 1, ipiv, 1);
      orc_dtrsm('L', 'L', 'N', 'U', jb, n - j - jb + 1, 1.0, ptr(j, j), lda, ptr(j, j + jb), lda);
      if (j + jb <= m)
        orc_dgemm('N', 'N', m - j - jb + 1, n - j - jb + 1, jb, -1.0, ptr(j + jb, j), lda,
                  ptr(j, j + jb), lda, 1.0, ptr(j + jb, j + jb), lda);
    }
  }
  return info;
}

// src/lapack/getrf.cc:287-339.
void orc_dgetrs(char trans, int n, int nrhs, const double* a, int lda, const int* ipiv,
                double* b, int ldb) {
  if (n == 0 || nrhs == 0) return;
  if (trans == 'N' || trans == 'n') {
    Laswp(nrhs, b, ldb, 1, n, ipiv, 1);
    orc_dtrsm('L', 'L', 'N', 'U', n, nrhs, 1.0, a, lda, b, ldb);
    orc_dtrsm('L', 'U', 'N', 'N', n, nrhs, 1.0, a, lda, b, ldb);
  } else {
    orc_dtrsm('L', 'U', 'T', 'N', n, nrhs, 1.0, a, lda, b, ldb);
    orc_dtrsm('L', 'L', 'T', 'U', n, nrhs, 1.0, a, lda, b, ldb);
    Laswp(nrhs, b, ldb, 1, n, ipiv, -1);
  }
}

// Componentwise backward error of X (src/lapack/gesvx.cc:395-425, trans =
// 'N'): berr_j = max_i |B - A X|_ij / (|A| |X| + |B|)_ij, with the safe1 /
// safe2 guard. The residual is accumulated in long double.
void orc_dgerfs_berr(int n, int nrhs, const double* a, int lda, const double* b, int ldb,
                     const double* x, int ldx, double* berr) {
  const double eps = std::numeric_limits<double>::epsilon() * 0.5;
  const double safmin = std::numeric_limits<double>::min();
  const double safe1 = double(n + 1) * safmin, safe2 = safe1 / eps;
  for (int j = 0; j < nrhs; ++j) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) {
      long double r = b[i + int64_t(j) * ldb];
      double w = std::fabs(b[i + int64_t(j) * ldb]);
      for (int k = 0; k < n; ++k) {
        const double aik = a[i + int64_t(k) * lda], xk = x[k + int64_t(j) * ldx];
        r -= static_cast<long double>(aik) * xk;
        w += std::fabs(aik) * std::fabs(xk);
      }
      const double ra = std::fabs(double(r));
      s = std::max(s, w > safe2 ? ra / w : (ra + safe1) / (w + safe1));
    }
    berr[j] = s;
  }
}

// xLACN2 for real data (src/lapack/trsen.cc:2235-2366): Hager's 1-norm
// estimator with Higham's refinements, by reverse communication. On return
// with kase = 1 the caller replaces x by M x, with kase = 2 by M^T x, and
// calls again; kase = 0 ends it with est (and v with M v = w,
// est = |v|_1 / |w|_1). isave[0] names the product the next call receives
// (1: M x0 with x0 = 1/n; 2: M^T sign; 3: M e_j; 4: M^T sign inside the
// main loop; 5: M b with the alternating vector b), isave[1] is j
// (1-based), isave[2] counts the main-loop iterations (at most 5). The sums
// |y|_1 run in index order (hwyBLAS's Asum is vectorized, so its last bit
// can differ). Planted mistake 9: the alternating-vector stage is skipped.
void orc_dlacn2(int n, double* v, double* x, int* isgn, double* est, int* kase, int* isave) {
  constexpr int kItmax = 5;
  auto sign = [](double t) { return t >= 0.0 ? 1.0 : -1.0; };
  auto asum = [n](const double* y) {
    double s = 0.0;
    for (int i = 0; i < n; ++i) s += std::fabs(y[i]);
    return s;
  };
  auto ask_unit = [&](int j) {  // x = e_j, then M x
    std::fill(x, x + n, 0.0);
    x[j - 1] = 1.0;
    *kase = 1;
    isave[0] = 3;
  };
  auto ask_signs = [&](int step) {  // x = sign(x), then M^T x
    for (int i = 0; i < n; ++i) {
      x[i] = sign(x[i]);
      isgn[i] = int(x[i]);
    }
    *kase = 2;
    isave[0] = step;
  };
  auto ask_alternating = [&]() {  // b_i = (-1)^i (1 + i / (n - 1)), then M b
    if (g_mistake == 9) {
      *kase = 0;
      return;
    }
    double s = 1.0;
    for (int i = 0; i < n; ++i) {
      x[i] = s * (1.0 + double(i) / double(n - 1));
      s = -s;
    }
    *kase = 1;
    isave[0] = 5;
  };
  if (*kase == 0) {
    std::fill(x, x + n, 1.0 / double(n));
    *kase = 1;
    isave[0] = 1;
    return;
  }
  switch (isave[0]) {
    case 2:  // x = M^T sign(M x0)
      isave[1] = IamaxFirst(n, x) + 1;
      isave[2] = 2;
      ask_unit(isave[1]);
      return;
    case 3: {  // x = M e_j
      std::copy(x, x + n, v);
      const double estold = *est;
      *est = asum(v);
      bool repeated = true;
      for (int i = 0; i < n && repeated; ++i) repeated = int(sign(x[i])) == isgn[i];
      if (repeated || *est <= estold) {
        ask_alternating();
        return;
      }
      ask_signs(4);
      return;
    }
    case 4: {  // x = M^T sign(M e_j)
      const int jlast = isave[1];
      isave[1] = IamaxFirst(n, x) + 1;
      if (x[jlast - 1] != std::fabs(x[isave[1] - 1]) && isave[2] < kItmax) {
        ++isave[2];
        ask_unit(isave[1]);
        return;
      }
      ask_alternating();
      return;
    }
    case 5: {  // x = M b
      const double temp = 2.0 * (asum(x) / double(3 * n));
      if (temp > *est) {
        std::copy(x, x + n, v);
        *est = temp;
      }
      *kase = 0;
      return;
    }
    default:  // 1 (and, as there, any other value): x = M x0
      if (n == 1) {
        v[0] = x[0];
        *est = std::fabs(v[0]);
        *kase = 0;
        return;
      }
      *est = asum(x);
      ask_signs(2);
      return;
  }
}

// xGERFS for real data (src/lapack/gesvx.cc:333-521). Per column j: the
// FP64 residual R = B - op(A) X (:412-413; each row a dot product in index
// order), berr with the safe1 / safe2 guard (:415-450), a correction
// X += getrs(op, R) while berr > eps, berr at least halved and at most 5
// corrections made (:339, :457-463); then FERR = ||inv(op(A)) diag(W)||inf /
// max|X(:, j)| with W = |R| + nz eps (|op(A)| |X| + |B|), nz = n + 1, plus
// safe1 where |op(A)| |X| + |B| <= safe2 (:465-487), the norm estimated by
// orc_dlacn2 on M = diag(W) inv(op(A))^T (kase 1: getrs with the other
// transposition, then times W; kase 2: times W, then getrs with op) (:488-
// 518). Returns -1 for an unknown trans. Planted mistakes: 10, kase 1 solves
// with op itself; 11, nz = n.
int orc_dgerfs(char trans, int n, int nrhs, const double* a, int lda, const double* af,
               int ldaf, const int* ipiv, const double* b, int ldb, double* x, int ldx,
               double* ferr, double* berr) {
  const bool notran = trans == 'N' || trans == 'n';
  if (!notran && trans != 'T' && trans != 't' && trans != 'C' && trans != 'c') return -1;
  if (n == 0 || nrhs == 0) {
    for (int j = 0; j < nrhs; ++j) ferr[j] = berr[j] = 0.0;
    return 0;
  }
  const char transn = notran ? 'N' : 'T', transt = notran ? 'T' : 'N';
  const char kase1 = g_mistake == 10 ? transn : transt;
  const double nz = g_mistake == 11 ? double(n) : double(n + 1);
  const double eps = std::numeric_limits<double>::epsilon() * 0.5;  // dlamch('E')
  const double safmin = std::numeric_limits<double>::min();
  const double safe1 = nz * safmin, safe2 = safe1 / eps;
  auto opa = [&](int i, int k) {
    return notran ? a[i + int64_t(k) * lda] : a[k + int64_t(i) * lda];
  };
  std::vector<double> r(n), w(n), v(n);
  std::vector<int> isgn(n);
  for (int j = 0; j < nrhs; ++j) {
    double* xj = x + int64_t(j) * ldx;
    const double* bj = b + int64_t(j) * ldb;
    int count = 1;
    double lstres = 3.0;
    for (;;) {
      for (int i = 0; i < n; ++i) {
        double t = 0.0, d = std::fabs(bj[i]);
        for (int k = 0; k < n; ++k) {
          t += opa(i, k) * xj[k];
          d += std::fabs(opa(i, k)) * std::fabs(xj[k]);
        }
        r[i] = bj[i] - t;
        w[i] = d;
      }
      double s = 0.0;
      for (int i = 0; i < n; ++i)
        s = std::max(s, w[i] > safe2 ? std::fabs(r[i]) / w[i]
                                     : (std::fabs(r[i]) + safe1) / (w[i] + safe1));
      berr[j] = s;
      if (s > eps && 2.0 * s <= lstres && count <= 5) {
        orc_dgetrs(trans, n, 1, af, ldaf, ipiv, r.data(), n);
        for (int i = 0; i < n; ++i) xj[i] += r[i];
        lstres = s;
        ++count;
        continue;
      }
      break;
    }
    for (int i = 0; i < n; ++i)
      w[i] = std::fabs(r[i]) + nz * eps * w[i] + (w[i] > safe2 ? 0.0 : safe1);
    int kase = 0, isave[3] = {0, 0, 0};
    double est = 0.0;
    for (;;) {
      orc_dlacn2(n, v.data(), r.data(), isgn.data(), &est, &kase, isave);
      if (kase == 0) break;
      if (kase == 1) {
        orc_dgetrs(kase1, n, 1, af, ldaf, ipiv, r.data(), n);
        for (int i = 0; i < n; ++i) r[i] *= w[i];
      } else {
        for (int i = 0; i < n; ++i) r[i] *= w[i];
        orc_dgetrs(transn, n, 1, af, ldaf, ipiv, r.data(), n);
      }
    }
    double xn = 0.0;
    for (int i = 0; i < n; ++i) xn = std::max(xn, std::fabs(xj[i]));
    ferr[j] = xn != 0.0 ? est / xn : est;
  }
  return 0;
}

// xPOTRF2 (src/lapack/potrf.cc:49-122) as LAPACK states it: the recursion
// n1 = n / 2 down to n = 1 (without hwyBLAS's leaf kernel), A21 := A21 L11^-T
// (or A12 := U11^-T A12) by the oracle's TRSM, then the triangle of A22
// minus A21 A21^T (A12^T A12), each entry's dot product in index order, then
// A22. Returns 0, or the order of the first leading minor that is not
// positive (or NaN). Planted mistake 12: the update uses row i twice
// (A21(i, :) A21(i, :)^T).
int orc_dpotrf2(char uplo, int n, double* a, int lda) {
  if (n == 0) return 0;
  const bool upper = uplo == 'U' || uplo == 'u';
  if (n == 1) {
    if (!(a[0] > 0.0)) return 1;
    a[0] = std::sqrt(a[0]);
    return 0;
  }
  const int n1 = n / 2, n2 = n - n1;
  int info = orc_dpotrf2(uplo, n1, a, lda);
  if (info != 0) return info;
  double* a22 = a + n1 + int64_t(n1) * lda;
  if (upper) {
    double* a12 = a + int64_t(n1) * lda;
    orc_dtrsm('L', 'U', 'T', 'N', n1, n2, 1.0, a, lda, a12, lda);
    for (int j = 0; j < n2; ++j)
      for (int i = 0; i <= j; ++i) {
        const int ii = g_mistake == 12 ? j : i;
        double acc = 0.0;
        for (int k = 0; k < n1; ++k)
          acc = std::fma(a12[k + int64_t(ii) * lda], a12[k + int64_t(j) * lda], acc);
        a22[i + int64_t(j) * lda] = -acc + a22[i + int64_t(j) * lda];
      }
  } else {
    double* a21 = a + n1;
    orc_dtrsm('R', 'L', 'T', 'N', n2, n1, 1.0, a, lda, a21, lda);
    for (int j = 0; j < n2; ++j)
      for (int i = j; i < n2; ++i) {
        const int jj = g_mistake == 12 ? i : j;
        double acc = 0.0;
        for (int k = 0; k < n1; ++k)
          acc = std::fma(a21[i + int64_t(k) * lda], a21[jj + int64_t(k) * lda], acc);
        a22[i + int64_t(j) * lda] = -acc + a22[i + int64_t(j) * lda];
      }
  }
  info = orc_dpotrf2(uplo, n2, a22, lda);
  return info != 0 ? info + n1 : 0;
}

}  // extern "C"
