// CBLAS wrappers over the Fortran-convention routines (hwyBLAS
// src/cblas.cc:517-537 for GEMM, :661-691 for TRSM).
//
// Column-major calls are forwarded unchanged. A row-major matrix is the
// column-major storage of its transpose, so row-major GEMM computes
// C^T = op(B)^T op(A)^T (A and B swapped, m and n swapped), and row-major
// TRSM solves with the other side and the other triangle (op unchanged).
// Invalid arguments are reported with CBLAS positions (the layout is
// argument 1): each wrapper tells the error path which CBLAS argument every
// Fortran argument came from.
#include <utility>

#include "common.h"

namespace hwy_amd {
namespace {

// Maps Fortran argument positions to CBLAS positions for one call.
class CblasScope {
 public:
  CblasScope(const char* name, std::initializer_list<int> positions) {
    int i = 0;
    for (int p : positions) positions_[i++] = p;
    ctx_.name = name;
    ctx_.positions = positions_;
    SetCblasContext(&ctx_);
  }
  ~CblasScope() { SetCblasContext(nullptr); }
  CblasScope(const CblasScope&) = delete;
  CblasScope& operator=(const CblasScope&) = delete;

 private:
  CblasContext ctx_;
  int positions_[16];
};

bool ValidLayout(const char* name, CBLAS_LAYOUT layout) {
  if (layout == CblasRowMajor || layout == CblasColMajor) return true;
  if (!CallErrorHandler(name, 1))
    cblas_xerbla(1, name, "Illegal layout setting, %d\n", static_cast<int>(layout));
  return false;
}

// Invalid enum values become '?', which the Fortran routine rejects.
char TransChar(CBLAS_TRANSPOSE t) {
  return t == CblasNoTrans ? 'N' : t == CblasTrans ? 'T' : t == CblasConjTrans ? 'C' : '?';
}
char UploChar(CBLAS_UPLO u) { return u == CblasUpper ? 'U' : u == CblasLower ? 'L' : '?'; }
char DiagChar(CBLAS_DIAG d) { return d == CblasNonUnit ? 'N' : d == CblasUnit ? 'U' : '?'; }
char SideChar(CBLAS_SIDE s) { return s == CblasLeft ? 'L' : s == CblasRight ? 'R' : '?'; }
char FlipUplo(char c) { return c == 'U' ? 'L' : c == 'L' ? 'U' : c; }
char FlipSide(char c) { return c == 'L' ? 'R' : c == 'R' ? 'L' : c; }

// Row-major SYRK/SYR2K take the other transposition (real: C is T).
char FlipRankTrans(char c) { return c == 'N' ? 'T' : (c == 'T' || c == 'C') ? 'N' : c; }

template <typename T>
struct Fortran;
template <>
struct Fortran<double> {
  static constexpr auto gemv = dgemv_;
  static constexpr auto gemm = dgemm_;
  static constexpr auto trsm = dtrsm_;
  static constexpr auto gemmtr = dgemmtr_;
  static constexpr auto gemmt = dgemmt_;
  static constexpr auto symm = dsymm_;
  static constexpr auto syrk = dsyrk_;
  static constexpr auto syr2k = dsyr2k_;
  static constexpr auto trmm = dtrmm_;
};
template <>
struct Fortran<float> {
  static constexpr auto gemv = sgemv_;
  static constexpr auto gemm = sgemm_;
  static constexpr auto trsm = strsm_;
  static constexpr auto gemmtr = sgemmtr_;
  static constexpr auto gemmt = sgemmt_;
  static constexpr auto symm = ssymm_;
  static constexpr auto syrk = ssyrk_;
  static constexpr auto syr2k = ssyr2k_;
  static constexpr auto trmm = strmm_;
};

template <typename T>
void Gemm(const char* name, CBLAS_LAYOUT layout, CBLAS_TRANSPOSE transa,
          CBLAS_TRANSPOSE transb, blasint m, blasint n, blasint k, T alpha,
          const T* a, blasint lda, const T* b, blasint ldb, T beta, T* c,
          blasint ldc) {
  if (!ValidLayout(name, layout)) return;
  const char ta = TransChar(transa), tb = TransChar(transb);
  if (layout == CblasColMajor) {
    CblasScope scope(name, {2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14});
    Fortran<T>::gemm(&ta, &tb, &m, &n, &k, &alpha, a, &lda, b, &ldb, &beta, c, &ldc);
  } else {
    CblasScope scope(name, {3, 2, 5, 4, 6, 7, 10, 11, 8, 9, 12, 13, 14});
    Fortran<T>::gemm(&tb, &ta, &n, &m, &k, &alpha, b, &ldb, a, &lda, &beta, c, &ldc);
  }
}

// Complex GEMM: the same mapping and positions as Gemm, the scalars and
// matrices as pointers to interleaved (re, im) reals.
template <typename R>
void ComplexGemm(const char* name, CBLAS_LAYOUT layout, CBLAS_TRANSPOSE transa,
                 CBLAS_TRANSPOSE transb, blasint m, blasint n, blasint k, const void* alpha,
                 const void* a, blasint lda, const void* b, blasint ldb, const void* beta,
                 void* c, blasint ldc) {
  using F = void (*)(const char*, const char*, const blasint*, const blasint*, const blasint*,
                     const R*, const R*, const blasint*, const R*, const blasint*, const R*,
                     R*, const blasint*);
  F gemm;
  if constexpr (sizeof(R) == 8) gemm = zgemm_;
  else gemm = cgemm_;
  if (!ValidLayout(name, layout)) return;
  const char ta = TransChar(transa), tb = TransChar(transb);
  const R* al = static_cast<const R*>(alpha);
  const R* be = static_cast<const R*>(beta);
  const R* ar = static_cast<const R*>(a);
  const R* br = static_cast<const R*>(b);
  R* cr = static_cast<R*>(c);
  if (layout == CblasColMajor) {
    CblasScope scope(name, {2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14});
    gemm(&ta, &tb, &m, &n, &k, al, ar, &lda, br, &ldb, be, cr, &ldc);
  } else {
    CblasScope scope(name, {3, 2, 5, 4, 6, 7, 10, 11, 8, 9, 12, 13, 14});
    gemm(&tb, &ta, &n, &m, &k, al, br, &ldb, ar, &lda, be, cr, &ldc);
  }
}

template <typename T>
void Trsm(const char* name, CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo,
          CBLAS_TRANSPOSE transa, CBLAS_DIAG diag, blasint m, blasint n, T alpha,
          const T* a, blasint lda, T* b, blasint ldb) {
  if (!ValidLayout(name, layout)) return;
  char sc = SideChar(side), uc = UploChar(uplo);
  const char tc = TransChar(transa), dc = DiagChar(diag);
  const bool row = layout == CblasRowMajor;
  if (row) {
    sc = FlipSide(sc);
    uc = FlipUplo(uc);
    std::swap(m, n);
  }
  CblasScope scope(name, row ? std::initializer_list<int>{2, 3, 4, 5, 7, 6, 8, 9, 10, 11, 12}
                             : std::initializer_list<int>{2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12});
  Fortran<T>::trsm(&sc, &uc, &tc, &dc, &m, &n, &alpha, a, &lda, b, &ldb);
}

// Complex TRSM: the same mapping and positions as Trsm.
template <typename R>
void ComplexTrsm(const char* name, CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo,
                 CBLAS_TRANSPOSE transa, CBLAS_DIAG diag, blasint m, blasint n,
                 const void* alpha, const void* a, blasint lda, void* b, blasint ldb) {
  using F = void (*)(const char*, const char*, const char*, const char*, const blasint*,
                     const blasint*, const R*, const R*, const blasint*, R*, const blasint*);
  F trsm;
  if constexpr (sizeof(R) == 8) trsm = ztrsm_;
  else trsm = ctrsm_;
  if (!ValidLayout(name, layout)) return;
  char sc = SideChar(side), uc = UploChar(uplo);
  const char tc = TransChar(transa), dc = DiagChar(diag);
  const bool row = layout == CblasRowMajor;
  if (row) {
    sc = FlipSide(sc);
    uc = FlipUplo(uc);
    std::swap(m, n);
  }
  CblasScope scope(name, row ? std::initializer_list<int>{2, 3, 4, 5, 7, 6, 8, 9, 10, 11, 12}
                             : std::initializer_list<int>{2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12});
  trsm(&sc, &uc, &tc, &dc, &m, &n, static_cast<const R*>(alpha), static_cast<const R*>(a), &lda,
       static_cast<R*>(b), &ldb);
}

// Complex GEMV. Row-major: N and T swap, and ConjTrans becomes 'R'
// (conjugated, not transposed) on the column-major storage A^T; m and n swap.
// Positions as Gemv.
template <typename R>
void ComplexGemv(const char* name, const char* fname, CBLAS_LAYOUT layout, CBLAS_TRANSPOSE trans,
                 blasint m, blasint n, const void* alpha, const void* a, blasint lda,
                 const void* x, blasint incx, const void* beta, void* y, blasint incy) {
  if (!ValidLayout(name, layout)) return;
  char tc = TransChar(trans);
  const bool row = layout == CblasRowMajor;
  if (row) {
    tc = tc == 'N' ? 'T' : tc == 'T' ? 'N' : tc == 'C' ? 'R' : tc;
    std::swap(m, n);
  }
  CblasScope scope(name, row ? std::initializer_list<int>{2, 4, 3, 5, 6, 7, 8, 9, 10, 11, 12}
                             : std::initializer_list<int>{2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12});
  ZgemvEntry<R>(fname, tc, m, n, static_cast<const R*>(alpha), static_cast<const R*>(a), lda,
                static_cast<const R*>(x), incx, static_cast<const R*>(beta), static_cast<R*>(y),
                incy);
}

// GEMV. Row-major: op(A) of the row-major A is the other op of the
// column-major A^T (n x m). Positions: layout 1, trans 2, m 3, n 4, alpha 5,
// a 6, lda 7, x 8, incx 9, beta 10, y 11, incy 12.
template <typename T>
void Gemv(const char* name, CBLAS_LAYOUT layout, CBLAS_TRANSPOSE trans, blasint m, blasint n,
          T alpha, const T* a, blasint lda, const T* x, blasint incx, T beta, T* y,
          blasint incy) {
  if (!ValidLayout(name, layout)) return;
  char tc = TransChar(trans);
  const bool row = layout == CblasRowMajor;
  if (row) {
    tc = tc == 'N' ? 'T' : (tc == 'T' || tc == 'C') ? 'N' : tc;
    std::swap(m, n);
  }
  CblasScope scope(name, row ? std::initializer_list<int>{2, 4, 3, 5, 6, 7, 8, 9, 10, 11, 12}
                             : std::initializer_list<int>{2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12});
  Fortran<T>::gemv(&tc, &m, &n, &alpha, a, &lda, x, &incx, &beta, y, &incy);
}

// GEMMTR (and its GEMMT name, `alias`). Row-major: the triangle of C^T =
// op(B)^T op(A)^T, the other uplo, A and B swapped with their
// transpositions. Positions: layout 1, uplo 2, transa 3, transb 4, n 5,
// k 6, alpha 7, a 8, lda 9, b 10, ldb 11, beta 12, c 13, ldc 14.
template <typename T>
void Gemmtr(const char* name, bool alias, CBLAS_LAYOUT layout, CBLAS_UPLO uplo,
            CBLAS_TRANSPOSE transa, CBLAS_TRANSPOSE transb, blasint n, blasint k, T alpha,
            const T* a, blasint lda, const T* b, blasint ldb, T beta, T* c, blasint ldc) {
  if (!ValidLayout(name, layout)) return;
  const char ta = TransChar(transa), tb = TransChar(transb), uc = UploChar(uplo);
  auto f = alias ? Fortran<T>::gemmt : Fortran<T>::gemmtr;
  if (layout == CblasColMajor) {
    CblasScope scope(name, {2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14});
    f(&uc, &ta, &tb, &n, &k, &alpha, a, &lda, b, &ldb, &beta, c, &ldc);
  } else {
    const char ur = FlipUplo(uc);
    CblasScope scope(name, {2, 4, 3, 5, 6, 7, 10, 11, 8, 9, 12, 13, 14});
    f(&ur, &tb, &ta, &n, &k, &alpha, b, &ldb, a, &lda, &beta, c, &ldc);
  }
}

// SYMM. Row-major: C^T = B^T A (side flipped), A's storage is the other
// triangle of the same symmetric matrix, m and n swapped. Positions: layout
// 1, side 2, uplo 3, m 4, n 5, alpha 6, a 7, lda 8, b 9, ldb 10, beta 11,
// c 12, ldc 13.
template <typename T>
void Symm(const char* name, CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo, blasint m,
          blasint n, T alpha, const T* a, blasint lda, const T* b, blasint ldb, T beta, T* c,
          blasint ldc) {
  if (!ValidLayout(name, layout)) return;
  char sc = SideChar(side), uc = UploChar(uplo);
  const bool row = layout == CblasRowMajor;
  if (row) {
    sc = FlipSide(sc);
    uc = FlipUplo(uc);
    std::swap(m, n);
  }
  CblasScope scope(name, row ? std::initializer_list<int>{2, 3, 5, 4, 6, 7, 8, 9, 10, 11, 12, 13}
                             : std::initializer_list<int>{2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13});
  Fortran<T>::symm(&sc, &uc, &m, &n, &alpha, a, &lda, b, &ldb, &beta, c, &ldc);
}

// SYRK. Row-major: C^T is the same product of the column-major A^T: the
// other triangle and transposition. Positions: layout 1, uplo 2, trans 3,
// n 4, k 5, alpha 6, a 7, lda 8, beta 9, c 10, ldc 11.
template <typename T>
void Syrk(const char* name, CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans,
          blasint n, blasint k, T alpha, const T* a, blasint lda, T beta, T* c, blasint ldc) {
  if (!ValidLayout(name, layout)) return;
  char uc = UploChar(uplo), tc = TransChar(trans);
  if (layout == CblasRowMajor) {
    uc = FlipUplo(uc);
    tc = FlipRankTrans(tc);
  }
  CblasScope scope(name, {2, 3, 4, 5, 6, 7, 8, 9, 10, 11});
  Fortran<T>::syrk(&uc, &tc, &n, &k, &alpha, a, &lda, &beta, c, &ldc);
}

// SYR2K: as SYRK. Positions: layout 1, uplo 2, trans 3, n 4, k 5, alpha 6,
// a 7, lda 8, b 9, ldb 10, beta 11, c 12, ldc 13.
template <typename T>
void Syr2k(const char* name, CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans,
           blasint n, blasint k, T alpha, const T* a, blasint lda, const T* b, blasint ldb,
           T beta, T* c, blasint ldc) {
  if (!ValidLayout(name, layout)) return;
  char uc = UploChar(uplo), tc = TransChar(trans);
  if (layout == CblasRowMajor) {
    uc = FlipUplo(uc);
    tc = FlipRankTrans(tc);
  }
  CblasScope scope(name, {2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13});
  Fortran<T>::syr2k(&uc, &tc, &n, &k, &alpha, a, &lda, b, &ldb, &beta, c, &ldc);
}

// TRMM: as TRSM (the other side and triangle, op unchanged).
template <typename T>
void Trmm(const char* name, CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo,
          CBLAS_TRANSPOSE transa, CBLAS_DIAG diag, blasint m, blasint n, T alpha, const T* a,
          blasint lda, T* b, blasint ldb) {
  if (!ValidLayout(name, layout)) return;
  char sc = SideChar(side), uc = UploChar(uplo);
  const char tc = TransChar(transa), dc = DiagChar(diag);
  const bool row = layout == CblasRowMajor;
  if (row) {
    sc = FlipSide(sc);
    uc = FlipUplo(uc);
    std::swap(m, n);
  }
  CblasScope scope(name, row ? std::initializer_list<int>{2, 3, 4, 5, 7, 6, 8, 9, 10, 11, 12}
                             : std::initializer_list<int>{2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12});
  Fortran<T>::trmm(&sc, &uc, &tc, &dc, &m, &n, &alpha, a, &lda, b, &ldb);
}

}  // namespace
}  // namespace hwy_amd

extern "C" {

#define HWY_CBLAS_GEMV(NAME, T)                                                               \
  HWYBLAS_AMD_API void NAME(const CBLAS_LAYOUT layout, const CBLAS_TRANSPOSE trans,           \
                            const hwyblas_int m, const hwyblas_int n, const T alpha,          \
                            const T* a, const hwyblas_int lda, const T* x,                    \
                            const hwyblas_int incx, const T beta, T* y,                       \
                            const hwyblas_int incy) {                                         \
    hwy_amd::Gemv<T>(#NAME, layout, trans, m, n, alpha, a, lda, x, incx, beta, y, incy);      \
  }
HWY_CBLAS_GEMV(cblas_sgemv, float)
HWY_CBLAS_GEMV(cblas_dgemv, double)
#undef HWY_CBLAS_GEMV

#define HWY_CBLAS_GEMMTR(NAME, T, ALIAS)                                                      \
  HWYBLAS_AMD_API void NAME(const CBLAS_LAYOUT layout, const CBLAS_UPLO uplo,                 \
                            const CBLAS_TRANSPOSE transa, const CBLAS_TRANSPOSE transb,       \
                            const hwyblas_int n, const hwyblas_int k, const T alpha,          \
                            const T* a, const hwyblas_int lda, const T* b,                    \
                            const hwyblas_int ldb, const T beta, T* c, const hwyblas_int ldc) { \
    hwy_amd::Gemmtr<T>(#NAME, ALIAS, layout, uplo, transa, transb, n, k, alpha, a, lda, b,    \
                       ldb, beta, c, ldc);                                                    \
  }
HWY_CBLAS_GEMMTR(cblas_sgemmtr, float, false)
HWY_CBLAS_GEMMTR(cblas_dgemmtr, double, false)
HWY_CBLAS_GEMMTR(cblas_sgemmt, float, true)
HWY_CBLAS_GEMMTR(cblas_dgemmt, double, true)
#undef HWY_CBLAS_GEMMTR

#define HWY_CBLAS_SYMM(NAME, T)                                                               \
  HWYBLAS_AMD_API void NAME(const CBLAS_LAYOUT layout, const CBLAS_SIDE side,                 \
                            const CBLAS_UPLO uplo, const hwyblas_int m, const hwyblas_int n,  \
                            const T alpha, const T* a, const hwyblas_int lda, const T* b,     \
                            const hwyblas_int ldb, const T beta, T* c, const hwyblas_int ldc) { \
    hwy_amd::Symm<T>(#NAME, layout, side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc);   \
  }
HWY_CBLAS_SYMM(cblas_ssymm, float)
HWY_CBLAS_SYMM(cblas_dsymm, double)
#undef HWY_CBLAS_SYMM

#define HWY_CBLAS_SYRK(NAME, T)                                                               \
  HWYBLAS_AMD_API void NAME(const CBLAS_LAYOUT layout, const CBLAS_UPLO uplo,                 \
                            const CBLAS_TRANSPOSE trans, const hwyblas_int n,                 \
                            const hwyblas_int k, const T alpha, const T* a,                   \
                            const hwyblas_int lda, const T beta, T* c, const hwyblas_int ldc) { \
    hwy_amd::Syrk<T>(#NAME, layout, uplo, trans, n, k, alpha, a, lda, beta, c, ldc);          \
  }
HWY_CBLAS_SYRK(cblas_ssyrk, float)
HWY_CBLAS_SYRK(cblas_dsyrk, double)
#undef HWY_CBLAS_SYRK

#define HWY_CBLAS_SYR2K(NAME, T)                                                              \
  HWYBLAS_AMD_API void NAME(const CBLAS_LAYOUT layout, const CBLAS_UPLO uplo,                 \
                            const CBLAS_TRANSPOSE trans, const hwyblas_int n,                 \
                            const hwyblas_int k, const T alpha, const T* a,                   \
                            const hwyblas_int lda, const T* b, const hwyblas_int ldb,         \
                            const T beta, T* c, const hwyblas_int ldc) {                      \
    hwy_amd::Syr2k<T>(#NAME, layout, uplo, trans, n, k, alpha, a, lda, b, ldb, beta, c, ldc); \
  }
HWY_CBLAS_SYR2K(cblas_ssyr2k, float)
HWY_CBLAS_SYR2K(cblas_dsyr2k, double)
#undef HWY_CBLAS_SYR2K

#define HWY_CBLAS_TRMM(NAME, T)                                                               \
  HWYBLAS_AMD_API void NAME(const CBLAS_LAYOUT layout, const CBLAS_SIDE side,                 \
                            const CBLAS_UPLO uplo, const CBLAS_TRANSPOSE transa,              \
                            const CBLAS_DIAG diag, const hwyblas_int m, const hwyblas_int n,  \
                            const T alpha, const T* a, const hwyblas_int lda, T* b,           \
                            const hwyblas_int ldb) {                                          \
    hwy_amd::Trmm<T>(#NAME, layout, side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb);   \
  }
HWY_CBLAS_TRMM(cblas_strmm, float)
HWY_CBLAS_TRMM(cblas_dtrmm, double)
#undef HWY_CBLAS_TRMM


HWYBLAS_AMD_API void cblas_sgemm(const CBLAS_LAYOUT layout, const CBLAS_TRANSPOSE transa,
                                 const CBLAS_TRANSPOSE transb, const hwyblas_int m,
                                 const hwyblas_int n, const hwyblas_int k, const float alpha,
                                 const float* a, const hwyblas_int lda, const float* b,
                                 const hwyblas_int ldb, const float beta, float* c,
                                 const hwyblas_int ldc) {
  hwy_amd::Gemm<float>("cblas_sgemm", layout, transa, transb, m, n, k, alpha, a, lda, b,
                       ldb, beta, c, ldc);
}

HWYBLAS_AMD_API void cblas_dgemm(const CBLAS_LAYOUT layout, const CBLAS_TRANSPOSE transa,
                                 const CBLAS_TRANSPOSE transb, const hwyblas_int m,
                                 const hwyblas_int n, const hwyblas_int k, const double alpha,
                                 const double* a, const hwyblas_int lda, const double* b,
                                 const hwyblas_int ldb, const double beta, double* c,
                                 const hwyblas_int ldc) {
  hwy_amd::Gemm<double>("cblas_dgemm", layout, transa, transb, m, n, k, alpha, a, lda, b,
                        ldb, beta, c, ldc);
}

HWYBLAS_AMD_API void cblas_cgemm(const CBLAS_LAYOUT layout, const CBLAS_TRANSPOSE transa,
                                 const CBLAS_TRANSPOSE transb, const hwyblas_int m,
                                 const hwyblas_int n, const hwyblas_int k, const void* alpha,
                                 const void* a, const hwyblas_int lda, const void* b,
                                 const hwyblas_int ldb, const void* beta, void* c,
                                 const hwyblas_int ldc) {
  hwy_amd::ComplexGemm<float>("cblas_cgemm", layout, transa, transb, m, n, k, alpha, a, lda,
                              b, ldb, beta, c, ldc);
}

HWYBLAS_AMD_API void cblas_zgemm(const CBLAS_LAYOUT layout, const CBLAS_TRANSPOSE transa,
                                 const CBLAS_TRANSPOSE transb, const hwyblas_int m,
                                 const hwyblas_int n, const hwyblas_int k, const void* alpha,
                                 const void* a, const hwyblas_int lda, const void* b,
                                 const hwyblas_int ldb, const void* beta, void* c,
                                 const hwyblas_int ldc) {
  hwy_amd::ComplexGemm<double>("cblas_zgemm", layout, transa, transb, m, n, k, alpha, a, lda,
                               b, ldb, beta, c, ldc);
}

HWYBLAS_AMD_API void cblas_ctrsm(const CBLAS_LAYOUT layout, const CBLAS_SIDE side,
                                 const CBLAS_UPLO uplo, const CBLAS_TRANSPOSE transa,
                                 const CBLAS_DIAG diag, const hwyblas_int m,
                                 const hwyblas_int n, const void* alpha, const void* a,
                                 const hwyblas_int lda, void* b, const hwyblas_int ldb) {
  hwy_amd::ComplexTrsm<float>("cblas_ctrsm", layout, side, uplo, transa, diag, m, n, alpha, a,
                              lda, b, ldb);
}

HWYBLAS_AMD_API void cblas_ztrsm(const CBLAS_LAYOUT layout, const CBLAS_SIDE side,
                                 const CBLAS_UPLO uplo, const CBLAS_TRANSPOSE transa,
                                 const CBLAS_DIAG diag, const hwyblas_int m,
                                 const hwyblas_int n, const void* alpha, const void* a,
                                 const hwyblas_int lda, void* b, const hwyblas_int ldb) {
  hwy_amd::ComplexTrsm<double>("cblas_ztrsm", layout, side, uplo, transa, diag, m, n, alpha, a,
                               lda, b, ldb);
}

HWYBLAS_AMD_API void cblas_cgemv(const CBLAS_LAYOUT layout, const CBLAS_TRANSPOSE trans,
                                 const hwyblas_int m, const hwyblas_int n, const void* alpha,
                                 const void* a, const hwyblas_int lda, const void* x,
                                 const hwyblas_int incx, const void* beta, void* y,
                                 const hwyblas_int incy) {
  hwy_amd::ComplexGemv<float>("cblas_cgemv", "CGEMV", layout, trans, m, n, alpha, a, lda, x,
                              incx, beta, y, incy);
}

HWYBLAS_AMD_API void cblas_zgemv(const CBLAS_LAYOUT layout, const CBLAS_TRANSPOSE trans,
                                 const hwyblas_int m, const hwyblas_int n, const void* alpha,
                                 const void* a, const hwyblas_int lda, const void* x,
                                 const hwyblas_int incx, const void* beta, void* y,
                                 const hwyblas_int incy) {
  hwy_amd::ComplexGemv<double>("cblas_zgemv", "ZGEMV", layout, trans, m, n, alpha, a, lda, x,
                               incx, beta, y, incy);
}

HWYBLAS_AMD_API void cblas_strsm(const CBLAS_LAYOUT layout, const CBLAS_SIDE side,
                                 const CBLAS_UPLO uplo, const CBLAS_TRANSPOSE transa,
                                 const CBLAS_DIAG diag, const hwyblas_int m,
                                 const hwyblas_int n, const float alpha, const float* a,
                                 const hwyblas_int lda, float* b, const hwyblas_int ldb) {
  hwy_amd::Trsm<float>("cblas_strsm", layout, side, uplo, transa, diag, m, n, alpha, a,
                       lda, b, ldb);
}

HWYBLAS_AMD_API void cblas_dtrsm(const CBLAS_LAYOUT layout, const CBLAS_SIDE side,
                                 const CBLAS_UPLO uplo, const CBLAS_TRANSPOSE transa,
                                 const CBLAS_DIAG diag, const hwyblas_int m,
                                 const hwyblas_int n, const double alpha, const double* a,
                                 const hwyblas_int lda, double* b, const hwyblas_int ldb) {
  hwy_amd::Trsm<double>("cblas_dtrsm", layout, side, uplo, transa, diag, m, n, alpha, a,
                        lda, b, ldb);
}

}  // extern "C"
