// The choice of the moduli (INT8) or slice (BF16) count of the emulated
// DGEMMs, restated on the CPU: exact statistics of every pair (i, j), the
// error bound of each count in units of |A||B|(i, j), and the rule that
// picks the count. Test infrastructure only.
//
// Statistics. Row i of op(A) has max |a_ip| = f 2^E_i (f in [0.5, 1); E_i = 0
// for a zero row); s_i = scale_bits - E_i with scale_bits 0 (INT8) or -1
// (BF16), bias = 7 - scale_bits, so that
//   q_ip  = floor(|a_ip| 2^(bias + s_i)) = floor(|a_ip| 2^(7 - E_i)) in [0, 127],
//   rs_i  = sum_p max(ceil(|a_ip| 2^(40 - E_i)), [a_ip != 0])  (an integer),
// likewise r_pj, cs_j, F_j for the columns of op(B), and G_ij = sum_p q_ip r_pj.
// Then 2^(E_i + F_j - 14) G_ij <= |A||B|(i, j) and rs_i 2^(E_i - 40) >=
// sum_p |a_ip|. Over the pairs whose row and column are not zero:
//   X = max (cs_j + rs_i) 2^-27 / G_ij (Inf when G_ij = 0), gmin = min G_ij.
// Error bounds per count c, with Z = 4096 k / gmin (0 without pairs):
//   INT8, bits = max(TotalBits(c, k), 2), bA = bits / 2:
//     t = 2^-bA X + 2^-bits Z;  typical t, worst t (1 + 2^-20);
//   BF16, n_c = ceil(k / 1024) c (c + 1) / 2 chunk sums:
//     trunc = 2^(1 - 8c) X, drop = (1.004 (c - 1) 2^-8c + 2^-16c) 4 Z,
//     typical trunc + drop + sqrt(n_c) u,
//     worst (trunc + drop)(1 + 2^-20) + (1.02 n_c + 3) u.
// Rule: NaN/Inf in either operand -> 0; forced > 0 -> forced clamped to
// [2, 24] (INT8) or [1, 16] (BF16); else the smallest c in that range with
// typical <= 2 sqrt(k) u and worst <= (4 (k + 2) - 24) u, else the smallest
// with worst <= (4 (k + 2) - 24) u, else 0.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <vector>

#include "oracle.h"

extern int g_mistake;

namespace {

struct Stats {
  std::vector<int> e_row, e_col;            // E_i, F_j
  std::vector<uint64_t> rs, cs;             // exact upper sums
  std::vector<uint8_t> q, r;                // q[i * k + p], r[j * k + p]
  bool finite = true;
};

double At(const double* x, int ld, bool t, int i, int p) {  // op(X)(i, p)
  return t ? x[p + size_t(i) * ld] : x[i + size_t(p) * ld];
}

bool Trans(char c) { return c == 'T' || c == 't' || c == 'C' || c == 'c'; }

// One side: R rows of length k, element (r, p) given by get(r, p).
template <typename Get>
void SideStats(int R, int k, Get get, std::vector<int>* e, std::vector<uint64_t>* sums,
               std::vector<uint8_t>* coarse, bool* finite) {
  e->assign(R, 0);
  sums->assign(R, 0);
  coarse->assign(size_t(R) * k, 0);
  for (int r = 0; r < R; ++r) {
    double mx = 0;
    for (int p = 0; p < k; ++p) {
      const double v = std::fabs(get(r, p));
      if (!(v <= std::numeric_limits<double>::max())) *finite = false;  // NaN or Inf
      if (v > mx) mx = v;
    }
    int E = 0;
    if (mx > 0 && mx <= std::numeric_limits<double>::max()) std::frexp(mx, &E);
    (*e)[r] = E;
    uint64_t s = 0;
    for (int p = 0; p < k; ++p) {
      const double ax = std::fabs(get(r, p));
      // Planted mistake 14: coarse magnitudes rounded to nearest, not down.
      const double v = std::ldexp(ax, 7 - E);
      const double vq = g_mistake == 14 ? std::min(std::nearbyint(v), 127.0) : v;
      (*coarse)[size_t(r) * k + p] = uint8_t((vq >= 0.0 && vq < 128.0) ? int(vq) : 0);
      const double w = std::ldexp(ax, 40 - E);
      uint64_t t = (w >= 0.0 && w <= 0x1p41) ? uint64_t(std::ceil(w)) : 0;
      if (ax > 0.0 && t == 0) t = 1;
      s += t;
    }
    (*sums)[r] = s;
  }
}

Stats Gather(char ta, char tb, int m, int n, int k, const double* a, int lda, const double* b,
             int ldb) {
  Stats st;
  const bool tA = Trans(ta), tB = Trans(tb);
  SideStats(m, k, [&](int i, int p) { return At(a, lda, tA, i, p); }, &st.e_row, &st.rs, &st.q,
            &st.finite);
  // op(B)(p, j): column j of op(B) as a row of op(B)^T.
  SideStats(n, k, [&](int j, int p) { return At(b, ldb, !tB, j, p); }, &st.e_col, &st.cs, &st.r,
            &st.finite);
  return st;
}

void Reduce(const Stats& st, int m, int n, int k, double* xmax, int* gmin) {
  uint64_t xb = 0;
  int g = std::numeric_limits<int>::max();
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < n; ++j) {
      if (st.rs[i] == 0 || st.cs[j] == 0) continue;
      int64_t G = 0;
      for (int p = 0; p < k; ++p) G += int64_t(st.q[size_t(i) * k + p]) * st.r[size_t(j) * k + p];
      // Planted mistake 13: the rows' sums (op(B)'s rounding term) dropped.
      const double rs = g_mistake == 13 ? 0.0 : double(st.rs[i]);
      const double x = G == 0 ? std::numeric_limits<double>::infinity()
                              : (double(st.cs[j]) + rs) * 0x1p-27 / double(G);
      uint64_t bits;
      std::memcpy(&bits, &x, 8);
      if (bits > xb) xb = bits;
      if (G < g) g = int(G);
    }
  std::memcpy(xmax, &xb, 8);
  *gmin = g;
}

void Bound(int kind, int count, int64_t k, double x, int gmin, double* typical, double* worst) {
  const double u = 0x1p-53;
  const double z = gmin == std::numeric_limits<int>::max() ? 0.0 : 4096.0 * double(k) / gmin;
  if (kind == 0) {
    const int bits = std::max(orc_ozaki_total_bits(count, k), 2), ba = bits / 2;
    const double t = std::ldexp(x, -ba) + std::ldexp(z, -bits);
    *typical = t;
    *worst = t * (1.0 + 0x1p-20);
  } else {
    const double s = count;
    const double trunc = std::ldexp(x, 1 - 8 * count);
    const double drop =
        (1.004 * (s - 1.0) * std::ldexp(1.0, -8 * count) + std::ldexp(1.0, -16 * count)) * 4.0 *
        z;
    const double chunks = double((k + 1023) / 1024) * (s * (s + 1.0) * 0.5);
    *typical = trunc + drop + std::sqrt(chunks) * u;
    *worst = (trunc + drop) * (1.0 + 0x1p-20) + (1.02 * chunks + 3.0) * u;
  }
}

}  // namespace

extern "C" {

int orc_emulated_choose(int kind, int forced, char ta, char tb, int m, int n, int k,
                        const double* a, int lda, const double* b, int ldb, double* xmax,
                        int* gmin) {
  const Stats st = Gather(ta, tb, m, n, k, a, lda, b, ldb);
  *xmax = 0;
  *gmin = std::numeric_limits<int>::max();
  if (!st.finite) return 0;
  const int lo = kind == 0 ? 2 : 1, hi = kind == 0 ? 24 : 16;
  if (forced > 0) return std::min(std::max(forced, lo), hi);
  Reduce(st, m, n, k, xmax, gmin);
  const double u = 0x1p-53, kk = double(k);
  const double t_typ = 2.0 * std::sqrt(kk) * u, t_bound = (4.0 * (kk + 2.0) - 24.0) * u;
  for (int c = lo; c <= hi; ++c) {
    double typ, worst;
    Bound(kind, c, k, *xmax, *gmin, &typ, &worst);
    if (typ <= t_typ && worst <= t_bound) return c;
  }
  for (int c = lo; c <= hi; ++c) {
    double typ, worst;
    Bound(kind, c, k, *xmax, *gmin, &typ, &worst);
    if (worst <= t_bound) return c;
  }
  return 0;
}

void orc_emulated_bound(int kind, int count, int64_t k, double xmax, int gmin, double* typical,
                        double* worst) {
  Bound(kind, count, k, xmax, gmin, typical, worst);
}

void orc_emulated_coarse(char ta, char tb, int m, int n, int k, const double* a, int lda,
                         const double* b, int ldb, double* lower, double* rowsum_up,
                         double* colsum_up) {
  const Stats st = Gather(ta, tb, m, n, k, a, lda, b, ldb);
  for (int i = 0; i < m; ++i) rowsum_up[i] = std::ldexp(double(st.rs[i]), st.e_row[i] - 40);
  for (int j = 0; j < n; ++j) colsum_up[j] = std::ldexp(double(st.cs[j]), st.e_col[j] - 40);
  for (int j = 0; j < n; ++j)
    for (int i = 0; i < m; ++i) {
      int64_t G = 0;
      for (int p = 0; p < k; ++p) G += int64_t(st.q[size_t(i) * k + p]) * st.r[size_t(j) * k + p];
      lower[i + size_t(j) * m] = std::ldexp(double(G), st.e_row[i] + st.e_col[j] - 14);
    }
}

}  // extern "C"
