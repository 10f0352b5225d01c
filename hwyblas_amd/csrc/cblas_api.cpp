// CBLAS forms of the routines (hwyBLAS src/cblas.cc: GEMV :279-330, GEMM
// :517-537, GEMMTR :543-560, SYMM :569-600, SYRK/SYR2K :606-655, TRMM/TRSM
// :661-691).
//
// A CBLAS call is forwarded to the Fortran-convention core (entries.h) as the
// column-major problem it denotes. Column-major calls keep their arguments.
// A row-major matrix is the column-major storage of its transpose, so a
// row-major call becomes the transposed equation:
//   GEMM, GEMMTR   C^T = op(B)^T op(A)^T: A and B exchange places with their
//                  transpositions and leading dimensions, m and n exchange,
//                  and GEMMTR's triangle is the other one;
//   SYMM, TRMM,    the other side and the other triangle, m and n exchanged
//   TRSM           (op is unchanged);
//   SYRK, SYR2K    the other triangle and the other transposition;
//   GEMV           the other transposition, m and n exchanged; a complex
//                  ConjTrans becomes 'R', conjugation without transposition.
//
// Invalid arguments are reported at the CBLAS position the program passed
// them in, the layout being argument 1. A Forward keeps, next to the Fortran
// argument list it builds, the CBLAS argument each entry holds; the rewrite
// moves values with Exchange, which moves those records along, so the
// position map follows from the rewrite itself.
#include <string.h>

#include <utility>

#include "entries.h"

namespace hwy_amd {
namespace {

class Forward {
 public:
  // `nargs`: the Fortran routine's argument count; CBLAS adds the layout in
  // front, so Fortran argument i starts as CBLAS argument i + 1.
  Forward(const char* name, CBLAS_LAYOUT layout, int nargs) : name_(name) {
    for (int i = 0; i < nargs; ++i) from_[i] = static_cast<unsigned char>(i + 2);
    valid_ = layout == CblasColMajor || layout == CblasRowMajor;
    row_major_ = layout == CblasRowMajor;
    if (!valid_) RejectLayout(name, static_cast<int>(layout));
  }
  Forward(const Forward&) = delete;
  Forward& operator=(const Forward&) = delete;

  bool valid() const { return valid_; }
  bool row_major() const { return row_major_; }

  // Fortran arguments i and j (1-based), holding x and y, trade places.
  template <typename V>
  void Exchange(int i, V& x, int j, V& y) {
    std::swap(x, y);
    std::swap(from_[i - 1], from_[j - 1]);
  }

  // The identity the core reports under; valid while this Forward lives.
  Caller caller() const { return Caller(name_, from_); }

 private:
  const char* name_;
  unsigned char from_[16] = {};
  bool valid_ = false;
  bool row_major_ = false;
};

// CBLAS enumerators to Fortran option letters. An enumeration's values are
// consecutive from `first`; any other value becomes '?', which the core then
// reports as invalid at that argument's position.
char Letter(int value, int first, const char* letters) {
  const int i = value - first;
  return (i >= 0 && i < static_cast<int>(strlen(letters))) ? letters[i] : '?';
}
char TransLetter(CBLAS_TRANSPOSE v) { return Letter(v, CblasNoTrans, "NTC"); }
char UploLetter(CBLAS_UPLO v) { return Letter(v, CblasUpper, "UL"); }
char DiagLetter(CBLAS_DIAG v) { return Letter(v, CblasNonUnit, "NU"); }
char SideLetter(CBLAS_SIDE v) { return Letter(v, CblasLeft, "LR"); }

// The other letter of the pair {x, y}; a letter outside it (invalid) stays.
char Other(char c, char x, char y) { return c == x ? y : c == y ? x : c; }

// The transposition of the transposed equation for a real operand: N <-> T
// (ConjTrans is Trans).
char OtherRealTrans(char c) { return c == 'N' ? 'T' : (c == 'T' || c == 'C') ? 'N' : c; }

// T: the real type of the routine; kComplex selects the complex core, whose
// scalars and matrices are (re, im) pairs of T.
template <typename T, bool kComplex>
void GemmCall(const char* name, CBLAS_LAYOUT layout, CBLAS_TRANSPOSE transa,
              CBLAS_TRANSPOSE transb, blasint m, blasint n, blasint k, const T* alpha,
              const T* a, blasint lda, const T* b, blasint ldb, const T* beta, T* c,
              blasint ldc) {
  Forward f(name, layout, 13);
  if (!f.valid()) return;
  char ta = TransLetter(transa), tb = TransLetter(transb);
  if (f.row_major()) {
    f.Exchange(1, ta, 2, tb);
    f.Exchange(3, m, 4, n);
    f.Exchange(7, a, 9, b);
    f.Exchange(8, lda, 10, ldb);
  }
  if constexpr (kComplex)
    ComplexGemm(f.caller(), &ta, &tb, &m, &n, &k, alpha, a, &lda, b, &ldb, beta, c, &ldc);
  else
    Gemm(f.caller(), &ta, &tb, &m, &n, &k, alpha, a, &lda, b, &ldb, beta, c, &ldc);
}

// TRSM (kind 0, real or complex) and TRMM (kind 1, real): Fortran arguments
// side 1, uplo 2, transa 3, diag 4, m 5, n 6, alpha 7, a 8, lda 9, b 10, ldb 11.
enum class TriKind { kSolve, kComplexSolve, kMultiply };

template <typename T, TriKind kKind>
void TriangularCall(const char* name, CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo,
                    CBLAS_TRANSPOSE transa, CBLAS_DIAG diag, blasint m, blasint n,
                    const T* alpha, const T* a, blasint lda, T* b, blasint ldb) {
  Forward f(name, layout, 11);
  if (!f.valid()) return;
  char sc = SideLetter(side), uc = UploLetter(uplo);
  const char tc = TransLetter(transa), dc = DiagLetter(diag);
  if (f.row_major()) {
    sc = Other(sc, 'L', 'R');
    uc = Other(uc, 'U', 'L');
    f.Exchange(5, m, 6, n);
  }
  const Caller who = f.caller();
  if constexpr (kKind == TriKind::kSolve)
    Trsm(who, &sc, &uc, &tc, &dc, &m, &n, alpha, a, &lda, b, &ldb);
  else if constexpr (kKind == TriKind::kComplexSolve)
    ComplexTrsm(who, &sc, &uc, &tc, &dc, &m, &n, alpha, a, &lda, b, &ldb);
  else
    Trmm(who, &sc, &uc, &tc, &dc, &m, &n, alpha, a, &lda, b, &ldb);
}

// Fortran arguments: trans 1, m 2, n 3, alpha 4, a 5, lda 6, x 7, incx 8,
// beta 9, y 10, incy 11.
template <typename T, bool kComplex>
void GemvCall(const char* name, CBLAS_LAYOUT layout, CBLAS_TRANSPOSE trans, blasint m,
              blasint n, const T* alpha, const T* a, blasint lda, const T* x, blasint incx,
              const T* beta, T* y, blasint incy) {
  Forward f(name, layout, 11);
  if (!f.valid()) return;
  char tc = TransLetter(trans);
  if (f.row_major()) {
    tc = kComplex ? (tc == 'C' ? 'R' : Other(tc, 'N', 'T')) : OtherRealTrans(tc);
    f.Exchange(2, m, 3, n);
  }
  if constexpr (kComplex)
    ComplexGemv(f.caller(), &tc, &m, &n, alpha, a, &lda, x, &incx, beta, y, &incy);
  else
    Gemv(f.caller(), &tc, &m, &n, alpha, a, &lda, x, &incx, beta, y, &incy);
}

// Fortran arguments: uplo 1, transa 2, transb 3, n 4, k 5, alpha 6, a 7,
// lda 8, b 9, ldb 10, beta 11, c 12, ldc 13.
template <typename T>
void GemmtrCall(const char* name, CBLAS_LAYOUT layout, CBLAS_UPLO uplo,
                CBLAS_TRANSPOSE transa, CBLAS_TRANSPOSE transb, blasint n, blasint k,
                const T* alpha, const T* a, blasint lda, const T* b, blasint ldb,
                const T* beta, T* c, blasint ldc) {
  Forward f(name, layout, 13);
  if (!f.valid()) return;
  char uc = UploLetter(uplo), ta = TransLetter(transa), tb = TransLetter(transb);
  if (f.row_major()) {
    uc = Other(uc, 'U', 'L');
    f.Exchange(2, ta, 3, tb);
    f.Exchange(7, a, 9, b);
    f.Exchange(8, lda, 10, ldb);
  }
  Gemmtr(f.caller(), &uc, &ta, &tb, &n, &k, alpha, a, &lda, b, &ldb, beta, c, &ldc);
}

// Fortran arguments: side 1, uplo 2, m 3, n 4, alpha 5, a 6, lda 7, b 8,
// ldb 9, beta 10, c 11, ldc 12. Row-major A is the other triangle's storage
// of the same symmetric matrix.
template <typename T>
void SymmCall(const char* name, CBLAS_LAYOUT layout, CBLAS_SIDE side, CBLAS_UPLO uplo,
              blasint m, blasint n, const T* alpha, const T* a, blasint lda, const T* b,
              blasint ldb, const T* beta, T* c, blasint ldc) {
  Forward f(name, layout, 12);
  if (!f.valid()) return;
  char sc = SideLetter(side), uc = UploLetter(uplo);
  if (f.row_major()) {
    sc = Other(sc, 'L', 'R');
    uc = Other(uc, 'U', 'L');
    f.Exchange(3, m, 4, n);
  }
  Symm(f.caller(), &sc, &uc, &m, &n, alpha, a, &lda, b, &ldb, beta, c, &ldc);
}

// SYRK (b == nullptr; Fortran arguments uplo 1, trans 2, n 3, k 4, alpha 5,
// a 6, lda 7, beta 8, c 9, ldc 10) and SYR2K (b 8, ldb 9 inserted after
// lda). Row-major: C^T is the same product of the column-major A^T.
template <typename T, bool kTwo>
void RankCall(const char* name, CBLAS_LAYOUT layout, CBLAS_UPLO uplo, CBLAS_TRANSPOSE trans,
              blasint n, blasint k, const T* alpha, const T* a, blasint lda, const T* b,
              blasint ldb, const T* beta, T* c, blasint ldc) {
  Forward f(name, layout, kTwo ? 12 : 10);
  if (!f.valid()) return;
  char uc = UploLetter(uplo), tc = TransLetter(trans);
  if (f.row_major()) {
    uc = Other(uc, 'U', 'L');
    tc = OtherRealTrans(tc);
  }
  if constexpr (kTwo)
    Syr2k(f.caller(), &uc, &tc, &n, &k, alpha, a, &lda, b, &ldb, beta, c, &ldc);
  else
    Syrk(f.caller(), &uc, &tc, &n, &k, alpha, a, &lda, beta, c, &ldc);
}

}  // namespace
}  // namespace hwy_amd

using hwy_amd::GemmCall;
using hwy_amd::TriKind;
using hwy_amd::TriangularCall;

extern "C" {

// ---- real routines: scalars by value ----

#define HWY_CBLAS_REAL(P, T)                                                                  \
  HWYBLAS_AMD_API void cblas_##P##gemm(                                                       \
      const CBLAS_LAYOUT layout, const CBLAS_TRANSPOSE transa, const CBLAS_TRANSPOSE transb,  \
      const hwyblas_int m, const hwyblas_int n, const hwyblas_int k, const T alpha,           \
      const T* a, const hwyblas_int lda, const T* b, const hwyblas_int ldb, const T beta,     \
      T* c, const hwyblas_int ldc) {                                                          \
    GemmCall<T, false>("cblas_" #P "gemm", layout, transa, transb, m, n, k, &alpha, a, lda,   \
                       b, ldb, &beta, c, ldc);                                                \
  }                                                                                           \
  HWYBLAS_AMD_API void cblas_##P##trsm(                                                       \
      const CBLAS_LAYOUT layout, const CBLAS_SIDE side, const CBLAS_UPLO uplo,                \
      const CBLAS_TRANSPOSE transa, const CBLAS_DIAG diag, const hwyblas_int m,               \
      const hwyblas_int n, const T alpha, const T* a, const hwyblas_int lda, T* b,            \
      const hwyblas_int ldb) {                                                                \
    TriangularCall<T, TriKind::kSolve>("cblas_" #P "trsm", layout, side, uplo, transa, diag,  \
                                       m, n, &alpha, a, lda, b, ldb);                         \
  }                                                                                           \
  HWYBLAS_AMD_API void cblas_##P##trmm(                                                       \
      const CBLAS_LAYOUT layout, const CBLAS_SIDE side, const CBLAS_UPLO uplo,                \
      const CBLAS_TRANSPOSE transa, const CBLAS_DIAG diag, const hwyblas_int m,               \
      const hwyblas_int n, const T alpha, const T* a, const hwyblas_int lda, T* b,            \
      const hwyblas_int ldb) {                                                                \
    TriangularCall<T, TriKind::kMultiply>("cblas_" #P "trmm", layout, side, uplo, transa,     \
                                          diag, m, n, &alpha, a, lda, b, ldb);                \
  }                                                                                           \
  HWYBLAS_AMD_API void cblas_##P##gemv(                                                       \
      const CBLAS_LAYOUT layout, const CBLAS_TRANSPOSE trans, const hwyblas_int m,            \
      const hwyblas_int n, const T alpha, const T* a, const hwyblas_int lda, const T* x,      \
      const hwyblas_int incx, const T beta, T* y, const hwyblas_int incy) {                   \
    hwy_amd::GemvCall<T, false>("cblas_" #P "gemv", layout, trans, m, n, &alpha, a, lda, x,   \
                                incx, &beta, y, incy);                                        \
  }                                                                                           \
  HWYBLAS_AMD_API void cblas_##P##gemmtr(                                                     \
      const CBLAS_LAYOUT layout, const CBLAS_UPLO uplo, const CBLAS_TRANSPOSE transa,         \
      const CBLAS_TRANSPOSE transb, const hwyblas_int n, const hwyblas_int k, const T alpha,  \
      const T* a, const hwyblas_int lda, const T* b, const hwyblas_int ldb, const T beta,     \
      T* c, const hwyblas_int ldc) {                                                          \
    hwy_amd::GemmtrCall<T>("cblas_" #P "gemmtr", layout, uplo, transa, transb, n, k, &alpha,  \
                           a, lda, b, ldb, &beta, c, ldc);                                    \
  }                                                                                           \
  HWYBLAS_AMD_API void cblas_##P##gemmt(                                                      \
      const CBLAS_LAYOUT layout, const CBLAS_UPLO uplo, const CBLAS_TRANSPOSE transa,         \
      const CBLAS_TRANSPOSE transb, const hwyblas_int n, const hwyblas_int k, const T alpha,  \
      const T* a, const hwyblas_int lda, const T* b, const hwyblas_int ldb, const T beta,     \
      T* c, const hwyblas_int ldc) {                                                          \
    hwy_amd::GemmtrCall<T>("cblas_" #P "gemmt", layout, uplo, transa, transb, n, k, &alpha,   \
                           a, lda, b, ldb, &beta, c, ldc);                                    \
  }                                                                                           \
  HWYBLAS_AMD_API void cblas_##P##symm(                                                       \
      const CBLAS_LAYOUT layout, const CBLAS_SIDE side, const CBLAS_UPLO uplo,                \
      const hwyblas_int m, const hwyblas_int n, const T alpha, const T* a,                    \
      const hwyblas_int lda, const T* b, const hwyblas_int ldb, const T beta, T* c,           \
      const hwyblas_int ldc) {                                                                \
    hwy_amd::SymmCall<T>("cblas_" #P "symm", layout, side, uplo, m, n, &alpha, a, lda, b,     \
                         ldb, &beta, c, ldc);                                                 \
  }                                                                                           \
  HWYBLAS_AMD_API void cblas_##P##syrk(                                                       \
      const CBLAS_LAYOUT layout, const CBLAS_UPLO uplo, const CBLAS_TRANSPOSE trans,          \
      const hwyblas_int n, const hwyblas_int k, const T alpha, const T* a,                    \
      const hwyblas_int lda, const T beta, T* c, const hwyblas_int ldc) {                     \
    hwy_amd::RankCall<T, false>("cblas_" #P "syrk", layout, uplo, trans, n, k, &alpha, a,     \
                                lda, nullptr, 0, &beta, c, ldc);                              \
  }                                                                                           \
  HWYBLAS_AMD_API void cblas_##P##syr2k(                                                      \
      const CBLAS_LAYOUT layout, const CBLAS_UPLO uplo, const CBLAS_TRANSPOSE trans,          \
      const hwyblas_int n, const hwyblas_int k, const T alpha, const T* a,                    \
      const hwyblas_int lda, const T* b, const hwyblas_int ldb, const T beta, T* c,           \
      const hwyblas_int ldc) {                                                                \
    hwy_amd::RankCall<T, true>("cblas_" #P "syr2k", layout, uplo, trans, n, k, &alpha, a,     \
                               lda, b, ldb, &beta, c, ldc);                                   \
  }

HWY_CBLAS_REAL(s, float)
HWY_CBLAS_REAL(d, double)
#undef HWY_CBLAS_REAL

// ---- complex routines: scalars and matrices through void pointers ----

#define HWY_CBLAS_COMPLEX(P, R)                                                               \
  HWYBLAS_AMD_API void cblas_##P##gemm(                                                       \
      const CBLAS_LAYOUT layout, const CBLAS_TRANSPOSE transa, const CBLAS_TRANSPOSE transb,  \
      const hwyblas_int m, const hwyblas_int n, const hwyblas_int k, const void* alpha,       \
      const void* a, const hwyblas_int lda, const void* b, const hwyblas_int ldb,             \
      const void* beta, void* c, const hwyblas_int ldc) {                                     \
    GemmCall<R, true>("cblas_" #P "gemm", layout, transa, transb, m, n, k,                    \
                      static_cast<const R*>(alpha), static_cast<const R*>(a), lda,            \
                      static_cast<const R*>(b), ldb, static_cast<const R*>(beta),             \
                      static_cast<R*>(c), ldc);                                               \
  }                                                                                           \
  HWYBLAS_AMD_API void cblas_##P##trsm(                                                       \
      const CBLAS_LAYOUT layout, const CBLAS_SIDE side, const CBLAS_UPLO uplo,                \
      const CBLAS_TRANSPOSE transa, const CBLAS_DIAG diag, const hwyblas_int m,               \
      const hwyblas_int n, const void* alpha, const void* a, const hwyblas_int lda, void* b,  \
      const hwyblas_int ldb) {                                                                \
    TriangularCall<R, TriKind::kComplexSolve>(                                                \
        "cblas_" #P "trsm", layout, side, uplo, transa, diag, m, n,                           \
        static_cast<const R*>(alpha), static_cast<const R*>(a), lda, static_cast<R*>(b),      \
        ldb);                                                                                 \
  }                                                                                           \
  HWYBLAS_AMD_API void cblas_##P##gemv(                                                       \
      const CBLAS_LAYOUT layout, const CBLAS_TRANSPOSE trans, const hwyblas_int m,            \
      const hwyblas_int n, const void* alpha, const void* a, const hwyblas_int lda,           \
      const void* x, const hwyblas_int incx, const void* beta, void* y,                       \
      const hwyblas_int incy) {                                                               \
    hwy_amd::GemvCall<R, true>("cblas_" #P "gemv", layout, trans, m, n,                       \
                               static_cast<const R*>(alpha), static_cast<const R*>(a), lda,   \
                               static_cast<const R*>(x), incx, static_cast<const R*>(beta),   \
                               static_cast<R*>(y), incy);                                     \
  }

HWY_CBLAS_COMPLEX(c, float)
HWY_CBLAS_COMPLEX(z, double)
#undef HWY_CBLAS_COMPLEX

}  // extern "C"
