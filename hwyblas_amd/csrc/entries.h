// The Fortran-convention cores of the BLAS routines, one overload per type.
// The extern "C" Fortran entries call them with their own routine name as
// the Caller; the CBLAS layer (cblas_api.cpp) calls them with the CBLAS name
// and the argument map of the call it forwards. Arguments are passed as the
// Fortran ABI passes them: by pointer, characters as their first byte.
#ifndef HWYBLAS_AMD_CSRC_ENTRIES_H_
#define HWYBLAS_AMD_CSRC_ENTRIES_H_

#include "common.h"

namespace hwy_amd {

#define HWY_REAL_LEVEL3_CORES(T)                                                              \
  void Gemm(const Caller& who, const char* transa, const char* transb, const blasint* m,      \
            const blasint* n, const blasint* k, const T* alpha, const T* a,                   \
            const blasint* lda, const T* b, const blasint* ldb, const T* beta, T* c,          \
            const blasint* ldc);                                                              \
  void Trsm(const Caller& who, const char* side, const char* uplo, const char* transa,        \
            const char* diag, const blasint* m, const blasint* n, const T* alpha, const T* a, \
            const blasint* lda, T* b, const blasint* ldb);                                    \
  void Trmm(const Caller& who, const char* side, const char* uplo, const char* transa,        \
            const char* diag, const blasint* m, const blasint* n, const T* alpha, const T* a, \
            const blasint* lda, T* b, const blasint* ldb);                                    \
  void Symm(const Caller& who, const char* side, const char* uplo, const blasint* m,          \
            const blasint* n, const T* alpha, const T* a, const blasint* lda, const T* b,     \
            const blasint* ldb, const T* beta, T* c, const blasint* ldc);                     \
  void Syrk(const Caller& who, const char* uplo, const char* trans, const blasint* n,         \
            const blasint* k, const T* alpha, const T* a, const blasint* lda, const T* beta,  \
            T* c, const blasint* ldc);                                                        \
  void Syr2k(const Caller& who, const char* uplo, const char* trans, const blasint* n,        \
             const blasint* k, const T* alpha, const T* a, const blasint* lda, const T* b,    \
             const blasint* ldb, const T* beta, T* c, const blasint* ldc);                    \
  void Gemmtr(const Caller& who, const char* uplo, const char* transa, const char* transb,    \
              const blasint* n, const blasint* k, const T* alpha, const T* a,                 \
              const blasint* lda, const T* b, const blasint* ldb, const T* beta, T* c,        \
              const blasint* ldc);                                                            \
  void Gemv(const Caller& who, const char* trans, const blasint* m, const blasint* n,         \
            const T* alpha, const T* a, const blasint* lda, const T* x, const blasint* incx,  \
            const T* beta, T* y, const blasint* incy);

HWY_REAL_LEVEL3_CORES(float)
HWY_REAL_LEVEL3_CORES(double)
#undef HWY_REAL_LEVEL3_CORES

// Complex routines: R is the real type; scalars and matrices are interleaved
// (re, im) pairs of R.
#define HWY_COMPLEX_CORES(R)                                                                  \
  void ComplexGemm(const Caller& who, const char* transa, const char* transb,                 \
                   const blasint* m, const blasint* n, const blasint* k, const R* alpha,      \
                   const R* a, const blasint* lda, const R* b, const blasint* ldb,            \
                   const R* beta, R* c, const blasint* ldc);                                  \
  void ComplexTrsm(const Caller& who, const char* side, const char* uplo,                     \
                   const char* transa, const char* diag, const blasint* m, const blasint* n,  \
                   const R* alpha, const R* a, const blasint* lda, R* b, const blasint* ldb); \
  /* op is 'N', 'T', 'C', or 'R' (conjugated, not transposed: row-major ConjTrans). */        \
  void ComplexGemv(const Caller& who, const char* op, const blasint* m, const blasint* n,     \
                   const R* alpha, const R* a, const blasint* lda, const R* x,                \
                   const blasint* incx, const R* beta, R* y, const blasint* incy);

HWY_COMPLEX_CORES(float)
HWY_COMPLEX_CORES(double)
#undef HWY_COMPLEX_CORES

}  // namespace hwy_amd

#endif  // HWYBLAS_AMD_CSRC_ENTRIES_H_
