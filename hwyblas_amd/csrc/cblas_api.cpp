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

template <typename T>
struct Fortran;
template <>
struct Fortran<double> {
  static constexpr auto gemm = dgemm_;
  static constexpr auto trsm = dtrsm_;
};
template <>
struct Fortran<float> {
  static constexpr auto gemm = sgemm_;
  static constexpr auto trsm = strsm_;
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

}  // namespace
}  // namespace hwy_amd

extern "C" {

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
