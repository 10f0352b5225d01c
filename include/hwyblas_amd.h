/*
 * libhwyblas_amd — the Level-3 hot path of hwyBLAS (?gemm_, ?trsm_) on
 * AMD Instinct MI355X (gfx950), behind the same Fortran and CBLAS ABI.
 *
 * Every declaration cites the hwyBLAS interface it replaces (paths relative
 * to the hwyBLAS source tree). Conventions are those of hwyBLAS:
 *  - LP64: dimensions, leading dimensions and positions are `int`
 *    (include/hwyblas/hwyblas.h:42);
 *  - Fortran routines take every argument by pointer and read only the first
 *    character of CHARACTER arguments; hidden length arguments are neither
 *    declared nor needed (include/hwyblas/hwyblas.h:14-16);
 *  - invalid arguments are reported through the installed handler, else
 *    cblas_xerbla / xerbla_, which print and return: the routine then
 *    returns with its outputs untouched (src/common.cc:33-44);
 *  - A and B are read only; only rows [0, m) of each column of C (of B for
 *    TRSM) are written; C is not read when beta == 0, A and B are not read
 *    when alpha == 0 (README.md:508-525).
 *
 * The Fortran and CBLAS routines accept host pointers (operands are staged
 * through device memory on a per-thread stream; the call returns when the
 * result is back) or device pointers (hipMalloc or managed memory: the work
 * runs in place). They are synchronous, reentrant and callable from many
 * threads. Large sgemm_/dgemm_ calls on host operands pipeline their copies
 * with two helper threads that end before the call returns
 * (HWYBLAS_AMD_HOST_PIPELINE=0 turns that off). The *_dev entry points take
 * device pointers and a stream and are asynchronous. A HIP failure (out of
 * device memory, failed launch) prints a message and aborts, as hwyBLAS does
 * when scratch allocation fails (src/common.cc:57-60).
 */
#ifndef HWYBLAS_AMD_H_
#define HWYBLAS_AMD_H_

#include <stddef.h>
#include <stdint.h>

#define HWYBLAS_AMD_API __attribute__((visibility("default")))

#ifdef __cplusplus
extern "C" {
#endif

/* include/hwyblas/hwyblas.h:42 */
typedef int hwyblas_int;

/* ------------------------------------------------------------------------ */
/* Fortran ABI                                                              */

/* C := alpha op(A) op(B) + beta C. Replaces sgemm_/dgemm_
 * (include/hwyblas/hwyblas.h:564-575, src/gemm.cc:142-158); argument checks
 * in the order of src/gemm.cc:89-106 (info 1, 2, 3, 4, 5, 8, 10, 13). */
HWYBLAS_AMD_API void sgemm_(const char* transa, const char* transb,
                            const hwyblas_int* m, const hwyblas_int* n,
                            const hwyblas_int* k, const float* alpha,
                            const float* a, const hwyblas_int* lda,
                            const float* b, const hwyblas_int* ldb,
                            const float* beta, float* c,
                            const hwyblas_int* ldc);
HWYBLAS_AMD_API void dgemm_(const char* transa, const char* transb,
                            const hwyblas_int* m, const hwyblas_int* n,
                            const hwyblas_int* k, const double* alpha,
                            const double* a, const hwyblas_int* lda,
                            const double* b, const hwyblas_int* ldb,
                            const double* beta, double* c,
                            const hwyblas_int* ldc);

/* Solves op(A) X = alpha B (side L) or X op(A) = alpha B (side R); X
 * overwrites B. Replaces strsm_/dtrsm_ (include/hwyblas/hwyblas.h:787-796,
 * src/trsm.cc:188-201); checks in the order of src/trsm.cc:151-168 (info 1,
 * 2, 3, 4, 5, 6, 9, 11). */
HWYBLAS_AMD_API void strsm_(const char* side, const char* uplo,
                            const char* transa, const char* diag,
                            const hwyblas_int* m, const hwyblas_int* n,
                            const float* alpha, const float* a,
                            const hwyblas_int* lda, float* b,
                            const hwyblas_int* ldb);
HWYBLAS_AMD_API void dtrsm_(const char* side, const char* uplo,
                            const char* transa, const char* diag,
                            const hwyblas_int* m, const hwyblas_int* n,
                            const double* alpha, const double* a,
                            const hwyblas_int* lda, double* b,
                            const hwyblas_int* ldb);

/* y := alpha op(A) x + beta y on the GPU (hwyBLAS src/gemv.cc:39-84: info 1
 * trans, 2 m, 3 n, 6 lda, 8 incx, 11 incy; quick return for m = 0, n = 0 or
 * alpha = 0 and beta = 1; beta = 0 never reads y; a negative increment walks
 * the vector from its end). Host or device operands as for dgemm_;
 * synchronous. Replaces sgemv_/dgemv_ (include/hwyblas/hwyblas.h:219-228). */
HWYBLAS_AMD_API void sgemv_(const char* trans, const hwyblas_int* m, const hwyblas_int* n,
                            const float* alpha, const float* a, const hwyblas_int* lda,
                            const float* x, const hwyblas_int* incx, const float* beta,
                            float* y, const hwyblas_int* incy);
HWYBLAS_AMD_API void dgemv_(const char* trans, const hwyblas_int* m, const hwyblas_int* n,
                            const double* alpha, const double* a, const hwyblas_int* lda,
                            const double* x, const hwyblas_int* incx, const double* beta,
                            double* y, const hwyblas_int* incy);

/* The other real Level-3 routines, on the same GEMM kernel, with hwyBLAS's
 * argument checks, info positions and quick returns (src/level3.cc:25-345).
 * Host or device operands as for dgemm_; synchronous.
 *  - ?symm_: C := alpha A B + beta C (side L) or alpha B A + beta C (R), A
 *    symmetric, read from its uplo triangle. Replaces ssymm_/dsymm_
 *    (include/hwyblas/hwyblas.h:591-602).
 *  - ?syrk_: the uplo triangle of C := alpha A A^T + beta C (trans N) or
 *    alpha A^T A + beta C (T, C); the other triangle is not written.
 *    Replaces ssyrk_/dsyrk_ (hwyblas.h:632-641).
 *  - ?syr2k_: the uplo triangle of C := alpha (op(A) op(B)^T + op(B) op(A)^T)
 *    + beta C. Replaces ssyr2k_/dsyr2k_ (hwyblas.h:668-679).
 *  - ?gemmtr_ and ?gemmt_ (its other name, reported as ?GEMMT): the uplo
 *    triangle of C := alpha op(A) op(B) + beta C, op(A) n x k. Replace
 *    sgemmtr_/dgemmtr_/sgemmt_/dgemmt_ (hwyblas.h:711-748).
 *  - ?trmm_: B := alpha op(A) B (side L) or alpha B op(A), A triangular
 *    (diag U: unit, not read). Replaces strmm_/dtrmm_ (hwyblas.h:764-774). */
HWYBLAS_AMD_API void ssymm_(const char* side, const char* uplo, const hwyblas_int* m,
                            const hwyblas_int* n, const float* alpha, const float* a,
                            const hwyblas_int* lda, const float* b, const hwyblas_int* ldb,
                            const float* beta, float* c, const hwyblas_int* ldc);
HWYBLAS_AMD_API void dsymm_(const char* side, const char* uplo, const hwyblas_int* m,
                            const hwyblas_int* n, const double* alpha, const double* a,
                            const hwyblas_int* lda, const double* b, const hwyblas_int* ldb,
                            const double* beta, double* c, const hwyblas_int* ldc);
HWYBLAS_AMD_API void ssyrk_(const char* uplo, const char* trans, const hwyblas_int* n,
                            const hwyblas_int* k, const float* alpha, const float* a,
                            const hwyblas_int* lda, const float* beta, float* c,
                            const hwyblas_int* ldc);
HWYBLAS_AMD_API void dsyrk_(const char* uplo, const char* trans, const hwyblas_int* n,
                            const hwyblas_int* k, const double* alpha, const double* a,
                            const hwyblas_int* lda, const double* beta, double* c,
                            const hwyblas_int* ldc);
HWYBLAS_AMD_API void ssyr2k_(const char* uplo, const char* trans, const hwyblas_int* n,
                             const hwyblas_int* k, const float* alpha, const float* a,
                             const hwyblas_int* lda, const float* b, const hwyblas_int* ldb,
                             const float* beta, float* c, const hwyblas_int* ldc);
HWYBLAS_AMD_API void dsyr2k_(const char* uplo, const char* trans, const hwyblas_int* n,
                             const hwyblas_int* k, const double* alpha, const double* a,
                             const hwyblas_int* lda, const double* b, const hwyblas_int* ldb,
                             const double* beta, double* c, const hwyblas_int* ldc);
HWYBLAS_AMD_API void sgemmtr_(const char* uplo, const char* transa, const char* transb,
                              const hwyblas_int* n, const hwyblas_int* k, const float* alpha,
                              const float* a, const hwyblas_int* lda, const float* b,
                              const hwyblas_int* ldb, const float* beta, float* c,
                              const hwyblas_int* ldc);
HWYBLAS_AMD_API void dgemmtr_(const char* uplo, const char* transa, const char* transb,
                              const hwyblas_int* n, const hwyblas_int* k, const double* alpha,
                              const double* a, const hwyblas_int* lda, const double* b,
                              const hwyblas_int* ldb, const double* beta, double* c,
                              const hwyblas_int* ldc);
HWYBLAS_AMD_API void sgemmt_(const char* uplo, const char* transa, const char* transb,
                             const hwyblas_int* n, const hwyblas_int* k, const float* alpha,
                             const float* a, const hwyblas_int* lda, const float* b,
                             const hwyblas_int* ldb, const float* beta, float* c,
                             const hwyblas_int* ldc);
HWYBLAS_AMD_API void dgemmt_(const char* uplo, const char* transa, const char* transb,
                             const hwyblas_int* n, const hwyblas_int* k, const double* alpha,
                             const double* a, const hwyblas_int* lda, const double* b,
                             const hwyblas_int* ldb, const double* beta, double* c,
                             const hwyblas_int* ldc);
HWYBLAS_AMD_API void strmm_(const char* side, const char* uplo, const char* transa,
                            const char* diag, const hwyblas_int* m, const hwyblas_int* n,
                            const float* alpha, const float* a, const hwyblas_int* lda,
                            float* b, const hwyblas_int* ldb);
HWYBLAS_AMD_API void dtrmm_(const char* side, const char* uplo, const char* transa,
                            const char* diag, const hwyblas_int* m, const hwyblas_int* n,
                            const double* alpha, const double* a, const hwyblas_int* lda,
                            double* b, const hwyblas_int* ldb);

/* ------------------------------------------------------------------------ */
/* Error handling (include/hwyblas/hwyblas.h:888-902, cblas.h:45)           */

/* Prints " ** On entry to NAME parameter number I had an illegal value" and
 * returns (src/common.cc:71-79). */
HWYBLAS_AMD_API void xerbla_(const char* srname, const hwyblas_int* info,
                             size_t srname_len);
/* Prints "Parameter P to routine ROUT was incorrect" and then `form`
 * (src/cblas.cc:714-726). */
HWYBLAS_AMD_API void cblas_xerbla(hwyblas_int p, const char* rout,
                                  const char* form, ...);
typedef void (*hwyblas_error_handler)(const char* name, hwyblas_int position);
/* Installs a handler called instead of xerbla_/cblas_xerbla; returns the
 * previous one; NULL restores the default (src/common.cc:81-84). */
HWYBLAS_AMD_API hwyblas_error_handler
hwyblas_set_error_handler(hwyblas_error_handler handler);

/* The DGEMM mode of the drop-in entries (no hwyBLAS counterpart): 0, the
 * default, runs dgemm_, cblas_dgemm and hwyblas_dgemm_dev on the native FP64
 * kernel; 1 runs products with m, n >= 2048 and k >= 1024 on the emulated
 * DGEMM (INT8 matrix cores, count chosen on the device, hwyBLAS's
 * componentwise bound 4 (k + 2) u (|alpha||A||B| + |beta||C|) guaranteed, the
 * native kernel where it cannot be certified; see hwyblas_dgemm_emulated_dev),
 * and dgetrf_/dgesv_'s trailing updates on it. The environment variable
 * HWYBLAS_AMD_DGEMM=emulated selects 1 at load. The mode is process-wide.
 * hwyblas_set_dgemm_mode returns the previous mode, or -1 (and changes
 * nothing) for a mode other than 0 or 1. */
HWYBLAS_AMD_API int hwyblas_set_dgemm_mode(int mode);
HWYBLAS_AMD_API int hwyblas_get_dgemm_mode(void);

/* ------------------------------------------------------------------------ */
/* CBLAS (include/hwyblas/cblas.h:26-38 enums, :607-622, :858-871)          */

typedef enum CBLAS_LAYOUT { CblasRowMajor = 101, CblasColMajor = 102 } CBLAS_LAYOUT;
typedef enum CBLAS_TRANSPOSE {
  CblasNoTrans = 111,
  CblasTrans = 112,
  CblasConjTrans = 113,
  CblasConjNoTrans = 114
} CBLAS_TRANSPOSE;
typedef enum CBLAS_UPLO { CblasUpper = 121, CblasLower = 122 } CBLAS_UPLO;
typedef enum CBLAS_DIAG { CblasNonUnit = 131, CblasUnit = 132 } CBLAS_DIAG;
typedef enum CBLAS_SIDE { CblasLeft = 141, CblasRight = 142 } CBLAS_SIDE;

/* Row-major maps to the column-major call with A and B swapped
 * (src/cblas.cc:517-537); error positions count the layout as argument 1. */
HWYBLAS_AMD_API void cblas_sgemm(const CBLAS_LAYOUT layout,
                                 const CBLAS_TRANSPOSE transa,
                                 const CBLAS_TRANSPOSE transb,
                                 const hwyblas_int m, const hwyblas_int n,
                                 const hwyblas_int k, const float alpha,
                                 const float* a, const hwyblas_int lda,
                                 const float* b, const hwyblas_int ldb,
                                 const float beta, float* c,
                                 const hwyblas_int ldc);
HWYBLAS_AMD_API void cblas_dgemm(const CBLAS_LAYOUT layout,
                                 const CBLAS_TRANSPOSE transa,
                                 const CBLAS_TRANSPOSE transb,
                                 const hwyblas_int m, const hwyblas_int n,
                                 const hwyblas_int k, const double alpha,
                                 const double* a, const hwyblas_int lda,
                                 const double* b, const hwyblas_int ldb,
                                 const double beta, double* c,
                                 const hwyblas_int ldc);
/* Row-major flips side and uplo and swaps m and n (src/cblas.cc:661-691). */
HWYBLAS_AMD_API void cblas_strsm(const CBLAS_LAYOUT layout,
                                 const CBLAS_SIDE side, const CBLAS_UPLO uplo,
                                 const CBLAS_TRANSPOSE transa,
                                 const CBLAS_DIAG diag, const hwyblas_int m,
                                 const hwyblas_int n, const float alpha,
                                 const float* a, const hwyblas_int lda,
                                 float* b, const hwyblas_int ldb);
HWYBLAS_AMD_API void cblas_dtrsm(const CBLAS_LAYOUT layout,
                                 const CBLAS_SIDE side, const CBLAS_UPLO uplo,
                                 const CBLAS_TRANSPOSE transa,
                                 const CBLAS_DIAG diag, const hwyblas_int m,
                                 const hwyblas_int n, const double alpha,
                                 const double* a, const hwyblas_int lda,
                                 double* b, const hwyblas_int ldb);
/* GEMV (src/cblas.cc:279-330; include/hwyblas/cblas.h:227-240): row-major
 * takes the other transposition and swaps m and n. */
HWYBLAS_AMD_API void cblas_sgemv(const CBLAS_LAYOUT layout, const CBLAS_TRANSPOSE trans,
                                 const hwyblas_int m, const hwyblas_int n, const float alpha,
                                 const float* a, const hwyblas_int lda, const float* x,
                                 const hwyblas_int incx, const float beta, float* y,
                                 const hwyblas_int incy);
HWYBLAS_AMD_API void cblas_dgemv(const CBLAS_LAYOUT layout, const CBLAS_TRANSPOSE trans,
                                 const hwyblas_int m, const hwyblas_int n, const double alpha,
                                 const double* a, const hwyblas_int lda, const double* x,
                                 const hwyblas_int incx, const double beta, double* y,
                                 const hwyblas_int incy);
/* CBLAS forms of the other real Level-3 routines (src/cblas.cc:543-701;
 * declarations include/hwyblas/cblas.h:641-840). Row-major: GEMMTR flips
 * uplo and swaps A and B with their transpositions; SYMM flips side and
 * uplo and swaps m and n; SYRK and SYR2K flip uplo and the transposition
 * (N <-> T); TRMM flips side and uplo and swaps m and n. Errors are reported
 * at CBLAS positions (layout = 1). */
HWYBLAS_AMD_API void cblas_sgemmtr(const CBLAS_LAYOUT layout, const CBLAS_UPLO uplo,
                                   const CBLAS_TRANSPOSE transa, const CBLAS_TRANSPOSE transb,
                                   const hwyblas_int n, const hwyblas_int k, const float alpha,
                                   const float* a, const hwyblas_int lda, const float* b,
                                   const hwyblas_int ldb, const float beta, float* c,
                                   const hwyblas_int ldc);
HWYBLAS_AMD_API void cblas_dgemmtr(const CBLAS_LAYOUT layout, const CBLAS_UPLO uplo,
                                   const CBLAS_TRANSPOSE transa, const CBLAS_TRANSPOSE transb,
                                   const hwyblas_int n, const hwyblas_int k, const double alpha,
                                   const double* a, const hwyblas_int lda, const double* b,
                                   const hwyblas_int ldb, const double beta, double* c,
                                   const hwyblas_int ldc);
HWYBLAS_AMD_API void cblas_sgemmt(const CBLAS_LAYOUT layout, const CBLAS_UPLO uplo,
                                  const CBLAS_TRANSPOSE transa, const CBLAS_TRANSPOSE transb,
                                  const hwyblas_int n, const hwyblas_int k, const float alpha,
                                  const float* a, const hwyblas_int lda, const float* b,
                                  const hwyblas_int ldb, const float beta, float* c,
                                  const hwyblas_int ldc);
HWYBLAS_AMD_API void cblas_dgemmt(const CBLAS_LAYOUT layout, const CBLAS_UPLO uplo,
                                  const CBLAS_TRANSPOSE transa, const CBLAS_TRANSPOSE transb,
                                  const hwyblas_int n, const hwyblas_int k, const double alpha,
                                  const double* a, const hwyblas_int lda, const double* b,
                                  const hwyblas_int ldb, const double beta, double* c,
                                  const hwyblas_int ldc);
HWYBLAS_AMD_API void cblas_ssymm(const CBLAS_LAYOUT layout, const CBLAS_SIDE side,
                                 const CBLAS_UPLO uplo, const hwyblas_int m, const hwyblas_int n,
                                 const float alpha, const float* a, const hwyblas_int lda,
                                 const float* b, const hwyblas_int ldb, const float beta,
                                 float* c, const hwyblas_int ldc);
HWYBLAS_AMD_API void cblas_dsymm(const CBLAS_LAYOUT layout, const CBLAS_SIDE side,
                                 const CBLAS_UPLO uplo, const hwyblas_int m, const hwyblas_int n,
                                 const double alpha, const double* a, const hwyblas_int lda,
                                 const double* b, const hwyblas_int ldb, const double beta,
                                 double* c, const hwyblas_int ldc);
HWYBLAS_AMD_API void cblas_ssyrk(const CBLAS_LAYOUT layout, const CBLAS_UPLO uplo,
                                 const CBLAS_TRANSPOSE trans, const hwyblas_int n,
                                 const hwyblas_int k, const float alpha, const float* a,
                                 const hwyblas_int lda, const float beta, float* c,
                                 const hwyblas_int ldc);
HWYBLAS_AMD_API void cblas_dsyrk(const CBLAS_LAYOUT layout, const CBLAS_UPLO uplo,
                                 const CBLAS_TRANSPOSE trans, const hwyblas_int n,
                                 const hwyblas_int k, const double alpha, const double* a,
                                 const hwyblas_int lda, const double beta, double* c,
                                 const hwyblas_int ldc);
HWYBLAS_AMD_API void cblas_ssyr2k(const CBLAS_LAYOUT layout, const CBLAS_UPLO uplo,
                                  const CBLAS_TRANSPOSE trans, const hwyblas_int n,
                                  const hwyblas_int k, const float alpha, const float* a,
                                  const hwyblas_int lda, const float* b, const hwyblas_int ldb,
                                  const float beta, float* c, const hwyblas_int ldc);
HWYBLAS_AMD_API void cblas_dsyr2k(const CBLAS_LAYOUT layout, const CBLAS_UPLO uplo,
                                  const CBLAS_TRANSPOSE trans, const hwyblas_int n,
                                  const hwyblas_int k, const double alpha, const double* a,
                                  const hwyblas_int lda, const double* b, const hwyblas_int ldb,
                                  const double beta, double* c, const hwyblas_int ldc);
HWYBLAS_AMD_API void cblas_strmm(const CBLAS_LAYOUT layout, const CBLAS_SIDE side,
                                 const CBLAS_UPLO uplo, const CBLAS_TRANSPOSE transa,
                                 const CBLAS_DIAG diag, const hwyblas_int m, const hwyblas_int n,
                                 const float alpha, const float* a, const hwyblas_int lda,
                                 float* b, const hwyblas_int ldb);
HWYBLAS_AMD_API void cblas_dtrmm(const CBLAS_LAYOUT layout, const CBLAS_SIDE side,
                                 const CBLAS_UPLO uplo, const CBLAS_TRANSPOSE transa,
                                 const CBLAS_DIAG diag, const hwyblas_int m, const hwyblas_int n,
                                 const double alpha, const double* a, const hwyblas_int lda,
                                 double* b, const hwyblas_int ldb);

/* ------------------------------------------------------------------------ */
/* Device-pointer entry points (no hwyBLAS counterpart: the reference has no
 * device). Same arguments, meaning, checks and quick returns as the Fortran
 * routines, with values instead of pointers for scalars, device pointers for
 * the matrices and a HIP stream (hipStream_t; NULL = the null stream). The
 * work is enqueued and the call returns without waiting. Returns 0, or the
 * Fortran position of the first invalid argument (also reported as above,
 * under the Fortran routine name). */
HWYBLAS_AMD_API int hwyblas_dgemm_dev(char transa, char transb, int m, int n,
                                      int k, double alpha, const double* a,
                                      int lda, const double* b, int ldb,
                                      double beta, double* c, int ldc,
                                      void* stream);
HWYBLAS_AMD_API int hwyblas_sgemm_dev(char transa, char transb, int m, int n,
                                      int k, float alpha, const float* a,
                                      int lda, const float* b, int ldb,
                                      float beta, float* c, int ldc,
                                      void* stream);
HWYBLAS_AMD_API int hwyblas_dtrsm_dev(char side, char uplo, char transa,
                                      char diag, int m, int n, double alpha,
                                      const double* a, int lda, double* b,
                                      int ldb, void* stream);
HWYBLAS_AMD_API int hwyblas_strsm_dev(char side, char uplo, char transa,
                                      char diag, int m, int n, float alpha,
                                      const float* a, int lda, float* b,
                                      int ldb, void* stream);

/* ------------------------------------------------------------------------ */
/* Emulated DGEMM on the INT8 matrix cores (no hwyBLAS counterpart). Same
 * arguments, checks (reported as "DGEMM") and quick returns as
 * hwyblas_dgemm_dev, FP64 in and out. The product is computed by the Ozaki
 * scheme II: power-of-two row/column scaling to integers, symmetric residues
 * modulo `nmoduli` pairwise coprime moduli <= 256, exact INT8 products and a
 * Chinese-remainder reconstruction.
 * nmoduli 2..24 (larger values use 24) fixes the count, with no accuracy
 * guarantee. nmoduli 0 lets the device choose it from exact statistics of
 * every pair (i, j) (row and column sums, and an INT8 product of coarse
 * magnitudes that bounds |A||B| from below): the smallest count whose error
 * bound is within 2 sqrt(k) u of |A||B| (u = 2^-53), else the smallest that
 * keeps hwyBLAS's componentwise bound
 * |C - C_exact| <= 4 (k + 2) u (|alpha||A||B| + |beta||C|), else the native
 * FP64 kernel. Operands with NaN or Inf, and k >= 131072, also use the
 * native kernel. The decision stays on the device: with used_moduli NULL the
 * call is asynchronous (it can be captured in a HIP graph once its
 * workspace has grown to the problem's size); a non-NULL used_moduli makes it
 * synchronise `stream` at the end and receive the count, or 0 when the
 * native kernel ran. Returns 0 or the Fortran info. */
HWYBLAS_AMD_API int hwyblas_dgemm_emulated_dev(char transa, char transb, int m, int n,
                                               int k, double alpha, const double* a,
                                               int lda, const double* b, int ldb,
                                               double beta, double* c, int ldc,
                                               int nmoduli, void* stream,
                                               int* used_moduli);

/* The INT8 emulated DGEMM that allocates nothing (no hwyBLAS counterpart;
 * for callers that cannot allocate device memory: code under HIP graph
 * capture, serving loops, harnesses that forbid hipMalloc). Same arguments,
 * checks, quick returns, count rule and results (bit for bit) as
 * hwyblas_dgemm_emulated_dev, but its workspace is a fixed 64 MiB array
 * that every GPU the library runs on holds from load until the process ends,
 * instead of a grow-only per-thread buffer; the first call on a device looks
 * up that array's address (no allocation). It needs
 * S + 24 (MP KP + NP KP + MP NP) bytes, MP and NP = m and n rounded up to
 * 256, KP = k rounded up to 128, S a few (m + n) words of statistics: for
 * example m = n = 768 with k up to 1280 (61 MB). A problem that needs
 * more returns -1 with a message on stderr and enqueues nothing (k >= 131072
 * runs the native kernel and needs none); -2 if the array's address cannot be
 * obtained. Every call on one GPU shares that GPU's array: calls that overlap
 * on one GPU (several streams or threads at once) race and must be ordered by
 * the caller. With used_moduli NULL the call neither allocates nor
 * synchronizes. Returns 0, the Fortran info, -1 or -2. */
HWYBLAS_AMD_API int hwyblas_dgemm_emulated_noalloc_dev(char transa, char transb, int m, int n,
                                                       int k, double alpha, const double* a,
                                                       int lda, const double* b, int ldb,
                                                       double beta, double* c, int ldc,
                                                       int nmoduli, void* stream,
                                                       int* used_moduli);

/* Emulated DGEMM on the BF16 matrix cores (no hwyBLAS counterpart): the
 * Ozaki scheme I with BF16 slices. Same arguments, checks and quick returns
 * as hwyblas_dgemm_emulated_dev. Rows of op(A) and columns of op(B) are
 * scaled by powers of two below 1 in magnitude and cut into `slices`
 * (1..16; larger values use 16) 8-bit integer digits, exact in BF16; the
 * s (s + 1) / 2 digit products with t + u <= s + 1 run on the BF16 MFMA with
 * exact FP32 accumulation over k chunks of 1024, summed in FP64 in a fixed
 * order (the result is reproducible bit for bit). slices 0 chooses the count
 * on the device by the same rule and statistics as the INT8 method, or runs
 * the native FP64 kernel; NaN or Inf operands and k >= 131072 run it too.
 * Asynchronous with used_slices NULL; otherwise synchronises at the end and
 * *used_slices receives the slice count, or 0 when the native kernel ran.
 * Returns 0 or the Fortran info. */
HWYBLAS_AMD_API int hwyblas_dgemm_emulated_bf16_dev(char transa, char transb, int m, int n,
                                                    int k, double alpha, const double* a,
                                                    int lda, const double* b, int ldb,
                                                    double beta, double* c, int ldc,
                                                    int slices, void* stream,
                                                    int* used_slices);

/* ------------------------------------------------------------------------ */
/* Batched (grouped) GEMM and TRSM over many independent, uneven problems (no
 * hwyBLAS counterpart: hwyBLAS solves one problem per call). A plan checks
 * every problem as the Fortran routine would (the first invalid one is
 * reported under "DGEMM"/"SGEMM"/"DTRSM"/"STRSM" with its position; *info
 * and *bad_index receive them and NULL is returned), uploads descriptors and
 * a work list once, and then runs all problems in one persistent launch,
 * directly or through a captured HIP graph. Matrices are device pointers;
 * the problem array and `configs` are host memory. dtype_bytes is 8 (double)
 * or 4 (float; alpha and beta are converted). configs[i] selects the tile
 * configuration of GEMM problem i (see hwyblas_grouped_config_info; -1 or a
 * NULL array: the built-in choice). Problems must not overlap in their
 * output. Launches are asynchronous on `stream`; a plan may be launched
 * again (the operands may change, not the problem list). */
typedef struct hwyblas_gemm_problem {
  char transa, transb;
  int m, n, k;
  double alpha, beta;
  const void* a;
  int lda;
  const void* b;
  int ldb;
  void* c;
  int ldc;
} hwyblas_gemm_problem;

typedef struct hwyblas_trsm_problem {
  char side, uplo, transa, diag;
  int m, n;
  double alpha;
  const void* a;
  int lda;
  void* b;
  int ldb;
} hwyblas_trsm_problem;

HWYBLAS_AMD_API void* hwyblas_gemm_plan_create(int dtype_bytes, int count,
                                               const hwyblas_gemm_problem* problems,
                                               const int* configs, int* info,
                                               int* bad_index);
HWYBLAS_AMD_API void* hwyblas_trsm_plan_create(int dtype_bytes, int count,
                                               const hwyblas_trsm_problem* problems,
                                               int* info, int* bad_index);
HWYBLAS_AMD_API void hwyblas_plan_launch(void* plan, void* stream);
/* As hwyblas_plan_launch through a HIP graph captured at the first call and
 * replayed afterwards. */
HWYBLAS_AMD_API void hwyblas_plan_launch_graph(void* plan, void* stream);
HWYBLAS_AMD_API int hwyblas_plan_work_items(void* plan);
/* The tile configuration a GEMM plan uses for problem i (-1: no work). */
HWYBLAS_AMD_API int hwyblas_plan_problem_config(void* plan, int problem);
HWYBLAS_AMD_API void hwyblas_plan_destroy(void* plan);
HWYBLAS_AMD_API int hwyblas_grouped_config_count(void);
/* Tile configuration cfg: a bm x bn tile; block = 0: one wave computes it
 * (unroll = k steps of 4 per load batch); block = 1: a workgroup of four
 * waves computes it through LDS (unroll = k per LDS stage). */
HWYBLAS_AMD_API int hwyblas_grouped_config_info(int cfg, int* bm, int* bn, int* unroll,
                                                int* block);
HWYBLAS_AMD_API int hwyblas_grouped_default_config(int dtype_bytes, int m, int n, int k);

/* ------------------------------------------------------------------------ */
/* Autotuner and cost model for the grouped tile configurations (no hwyBLAS
 * counterpart: hwyBLAS fixes its blocking per SIMD target at compile time,
 * src/gemm_kernel-inl.h:60-75). */

/* Times every grouped configuration on each shape (5 ints per shape:
 * transa != 0, transb != 0, m, n, k; m, n, k >= 1): a plan of
 * hwyblas_grouped_tune_copies() problems of that shape with distinct device
 * operands, launched `reps` times after a warm-up. Writes microseconds per
 * problem to us[shape * hwyblas_grouped_config_count() + cfg]. Returns 0, or
 * -1 on invalid arguments. */
HWYBLAS_AMD_API int hwyblas_grouped_tune(int dtype_bytes, int nshapes, const int* shapes,
                                         int reps, float* us, void* stream);
HWYBLAS_AMD_API int hwyblas_grouped_tune_copies(int dtype_bytes, int m, int n, int k);
/* The cost model's 6 input features of a shape. */
HWYBLAS_AMD_API void hwyblas_grouped_features(int dtype_bytes, int transa, int transb, int m,
                                              int n, int k, float* features);
/* Fills configs[i] for each problem (host array, as for plan creation):
 * policy 0 the built-in rule; 1 the installed cost model (falls back to 2
 * when none is installed); 2 exhaustive tuning of each distinct shape, kept
 * for the process. Returns 0, or -1 on invalid arguments. */
HWYBLAS_AMD_API int hwyblas_grouped_choose_configs(int dtype_bytes, int count,
                                                   const hwyblas_gemm_problem* problems,
                                                   int policy, int* configs, void* stream);

/* The cost model: a 4-layer MLP, width 64, FP32 (6 inputs -> 64 -> 64 -> 64
 * -> `outputs`, ReLU), predicting log2(microseconds per problem) of each
 * configuration. Parameters row-major W[out][in], in the order W1 b1 W2 b2
 * W3 b3 W4 b4. Training, gradients and inference run on the GPU. */
HWYBLAS_AMD_API void* hwyblas_mlp_create(int outputs, uint64_t seed);
HWYBLAS_AMD_API void hwyblas_mlp_destroy(void* model);
HWYBLAS_AMD_API int hwyblas_mlp_num_params(void* model);
HWYBLAS_AMD_API int hwyblas_mlp_outputs(void* model);
HWYBLAS_AMD_API void hwyblas_mlp_get_params(void* model, float* params);
/* Replaces the parameters and resets the optimizer state. */
HWYBLAS_AMD_API void hwyblas_mlp_set_params(void* model, const float* params);
/* out[i * outputs + c]: prediction for features[i * 6 ...]. */
HWYBLAS_AMD_API int hwyblas_mlp_predict(void* model, int n, const float* features, float* out,
                                        void* stream);
/* Loss 0.5 mean_i (out[i][cfg[i]] - target[i])^2 and its gradient (no update). */
HWYBLAS_AMD_API int hwyblas_mlp_gradients(void* model, int n, const float* features,
                                          const int* cfg, const float* target, float* grads,
                                          float* loss, void* stream);
/* Adam (beta 0.9, 0.999) over shuffled mini-batches of that loss. */
HWYBLAS_AMD_API int hwyblas_mlp_train(void* model, int n, const float* features, const int* cfg,
                                      const float* target, int epochs, int batch, float lr,
                                      uint64_t seed, float* epoch_loss, void* stream);
/* Installs the model policy 1 uses (NULL uninstalls); the caller keeps it alive. */
HWYBLAS_AMD_API void hwyblas_grouped_set_model(void* model);

/* ------------------------------------------------------------------------ */
/* Cholesky on the GPU (hwyBLAS src/lapack/potrf.cc, posv.cc; LAPACK
 * signatures): ?potrf_ factors A = L L^T (uplo L) or U^T U (U) in place, the
 * other triangle untouched; info 0, -i for argument i (1 uplo, 2 n, 4 lda;
 * reported as ?POTRF), or the order of the first leading minor that is not
 * positive (or NaN; entries past that column are then not specified).
 * ?potrs_ solves with the factor (info positions 1 uplo, 2 n, 3 nrhs, 5 lda,
 * 7 ldb); ?posv_ factors, then solves if the factorization succeeded. Host
 * or device operands as for dgemm_; synchronous. */
HWYBLAS_AMD_API void spotrf_(const char* uplo, const hwyblas_int* n, float* a,
                             const hwyblas_int* lda, hwyblas_int* info);
HWYBLAS_AMD_API void dpotrf_(const char* uplo, const hwyblas_int* n, double* a,
                             const hwyblas_int* lda, hwyblas_int* info);
HWYBLAS_AMD_API void spotrs_(const char* uplo, const hwyblas_int* n, const hwyblas_int* nrhs,
                             const float* a, const hwyblas_int* lda, float* b,
                             const hwyblas_int* ldb, hwyblas_int* info);
HWYBLAS_AMD_API void dpotrs_(const char* uplo, const hwyblas_int* n, const hwyblas_int* nrhs,
                             const double* a, const hwyblas_int* lda, double* b,
                             const hwyblas_int* ldb, hwyblas_int* info);
HWYBLAS_AMD_API void sposv_(const char* uplo, const hwyblas_int* n, const hwyblas_int* nrhs,
                            float* a, const hwyblas_int* lda, float* b, const hwyblas_int* ldb,
                            hwyblas_int* info);
HWYBLAS_AMD_API void dposv_(const char* uplo, const hwyblas_int* n, const hwyblas_int* nrhs,
                            double* a, const hwyblas_int* lda, double* b, const hwyblas_int* ldb,
                            hwyblas_int* info);

/* ------------------------------------------------------------------------ */
/* LU consumer on device memory: hwyBLAS's LAPACK port calls dgemm_/dtrsm_
 * from dgetrf/dgetrs/dgerfs (src/lapack/getrf.cc:216-339,
 * src/lapack/gesvx.cc:333-463); these run the same algorithms on the GPU. */

/* A = P L U with partial pivoting (the pivot of a column is its first entry
 * of largest magnitude, as xIAMAX), the recursion of dgetrf2
 * (src/lapack/getrf.cc:112-207) with 64-column panels. ipiv: device array of
 * min(m, n) 1-based row indices. update: precision of the trailing-matrix
 * GEMMs (0 FP64, 1 emulated FP64 on the INT8 matrix cores where k >= 1024,
 * 2 FP32); panels and TRSMs are FP64. A panel is factored by one
 * cooperative launch when all its workgroups fit on the device at once
 * (HWYBLAS_AMD_LU_PANEL=steps selects one launch per column instead).
 * Returns info: 0, -i for an invalid argument i (reported as DGETRF), the
 * first exactly-zero pivot (1-based), or -1000 if a wait inside the
 * cooperative panel kernel exceeded its bound (not expected; the factors
 * are then invalid). Synchronizes the stream. */
HWYBLAS_AMD_API int hwyblas_dgetrf_dev(int m, int n, double* a, int lda, int* ipiv,
                                       int update, void* stream);
/* op(A) X = B with the factorization (src/lapack/getrf.cc:287-339); X
 * overwrites B. Returns 0 or -i (reported as DGETRS). Synchronizes (the row
 * interchanges are composed into one permutation on the host). */
HWYBLAS_AMD_API int hwyblas_dgetrs_dev(char trans, int n, int nrhs, const double* a, int lda,
                                       const int* ipiv, double* b, int ldb, void* stream);
/* A X = B for nrhs <= 16: af := LU of A (dgetrf with `update`), X := dgetrs,
 * then FP64 iterative refinement against A as dgerfs does it
 * (src/lapack/gesvx.cc:395-463): a column is corrected while its
 * componentwise backward error berr exceeds eps, halves at least, and fewer
 * than itmax corrections were made. Then, when ferr is not NULL, dgerfs's
 * forward error bound (src/lapack/gesvx.cc:465-518): ferr_j estimates
 * || |inv(A)| (|R| + (n + 1) eps (|A| |X| + |B|)) ||inf / max|X(:, j)| by
 * the xLACN2 estimator (src/lapack/trsen.cc:2235-2366), its products solved
 * with af on the GPU. Host outputs per right-hand side: ferr, berr,
 * nbe = ||B - A X||inf / (||A||inf ||X||inf), iterations = corrections
 * (nbe and iterations may be NULL). Returns 0, getrf's info (> 0, or -1000
 * as there), or -i (reported as DGESVIR). Synchronizes. */
HWYBLAS_AMD_API int hwyblas_dgesv_ir_dev(int n, int nrhs, const double* a, int lda, double* af,
                                         int ldaf, int* ipiv, const double* b, int ldb,
                                         double* x, int ldx, int update, int itmax,
                                         double* ferr, double* berr, double* nbe,
                                         int* iterations, void* stream);
/* Householder QR, A = Q R (hwyBLAS src/lapack/geqrf.cc:66-150; signature as
 * include/hwyblas/lapack.h:1281-1283): R on and above the diagonal of A,
 * the reflectors H(i) = I - tau_i v v^T (v(i) = 1 implicit, v(i+1:m) below
 * the diagonal) and tau (min(m, n) values), with xLARFG's signs. Info
 * positions: 1 m, 2 n, 4 lda, 7 lwork (lwork >= max(1, n), or -1 for the
 * workspace query, which returns n * 32 in work[0]); reported as DGEQRF. A
 * and tau may be host or device memory; work is only written (work[0]).
 * Each call synchronizes. The factorization runs on the GPU in panels of 64
 * columns (one cooperative launch per panel when its workgroups all fit on
 * the device; HWYBLAS_AMD_QR_PANEL=steps selects one launch per column, with
 * the same bits), block reflectors of 128 columns, and the trailing updates
 * on the MFMA GEMM. info = -1000 is not an argument position: a wait inside
 * the cooperative panel kernel exceeded its bound (not expected), a message
 * names it on stderr, the installed error handler (if any) receives
 * ("DGEQRF", -1000), the affected tau are NaN and A holds no valid
 * factorization. */
HWYBLAS_AMD_API void dgeqrf_(const hwyblas_int* m, const hwyblas_int* n, double* a,
                             const hwyblas_int* lda, double* tau, double* work,
                             const hwyblas_int* lwork, hwyblas_int* info);
/* The same on device memory (A and tau), asynchronous on `stream` (no
 * hwyBLAS counterpart as such). Returns 0 or -i (reported as DGEQRF); a
 * timed-out panel barrier (not expected) shows as NaN in tau. */
HWYBLAS_AMD_API int hwyblas_dgeqrf_dev(int m, int n, double* a, int lda, double* tau,
                                       void* stream);

/* xGERFS on device operands (no hwyBLAS counterpart as such; dgerfs_ above is
 * the LAPACK form): refines X (n x nrhs) for op(A) X = B (trans 'N' or 'T')
 * with the factors af and the 1-based pivots ipiv of hwyblas_dgetrf_dev,
 * all device memory; ferr and berr receive nrhs values in host memory.
 * Returns 0 or -i (reported as DGERFS, xGERFS's positions). Synchronizes. */
HWYBLAS_AMD_API int hwyblas_dgerfs_dev(char trans, int n, int nrhs, const double* a, int lda,
                                       const double* af, int ldaf, const int* ipiv,
                                       const double* b, int ldb, double* x, int ldx,
                                       double* ferr, double* berr, void* stream);

/* ------------------------------------------------------------------------ */
/* Tile configurations of the plain GEMM kernel (no hwyBLAS counterpart: its
 * blocking is compile-time, src/gemm_kernel-inl.h:72-74). For the autotuner
 * and the cost model: the configurations are numbered 0..count-1; every one
 * computes the same result (the k order is the same), only the time
 * differs. The *_dev_cfg forms take m, n, k > 0 and valid leading
 * dimensions (else they return -1 without reporting) and enqueue like the
 * *_dev entry points. */
HWYBLAS_AMD_API int hwyblas_gemm_config_count(void);
/* Macro-tile bm x bn, k slice bk, waves per workgroup. 0, or -1 if cfg is
 * out of range. */
HWYBLAS_AMD_API int hwyblas_gemm_config_info(int cfg, int* bm, int* bn, int* bk,
                                             int* waves);
/* The depth of a configuration's LDS ring when its slices are streamed by
 * direct global-to-LDS loads (2 or more; such a configuration runs the
 * full tiles of calls with k a multiple of bk, 16-byte aligned operands and
 * leading dimensions of whole 16-byte chunks, and hands the edge strips and
 * other calls to the plain kernel), 0 for a register-staged configuration,
 * -1 if cfg is out of range. */
HWYBLAS_AMD_API int hwyblas_gemm_config_stages(int cfg);
/* The configuration the Fortran/CBLAS/dev entry points use for a shape. */
HWYBLAS_AMD_API int hwyblas_gemm_default_config(int dtype_bytes, char transa,
                                                char transb, int m, int n, int k);
HWYBLAS_AMD_API int hwyblas_dgemm_dev_cfg(int cfg, char transa, char transb, int m,
                                          int n, int k, double alpha,
                                          const double* a, int lda, const double* b,
                                          int ldb, double beta, double* c, int ldc,
                                          void* stream);
HWYBLAS_AMD_API int hwyblas_sgemm_dev_cfg(int cfg, char transa, char transb, int m,
                                          int n, int k, float alpha, const float* a,
                                          int lda, const float* b, int ldb, float beta,
                                          float* c, int ldc, void* stream);

/* ------------------------------------------------------------------------ */
/* Device utilities (no hwyBLAS counterpart): memory, streams, timing and the
 * seeded generator shared with the CPU oracle, for callers and benchmarks
 * that hold no other HIP binding. Errors abort like any HIP failure.        */

HWYBLAS_AMD_API int hwyblas_amd_device_count(void);
HWYBLAS_AMD_API void* hwyblas_amd_malloc(size_t bytes);
HWYBLAS_AMD_API void hwyblas_amd_free(void* p);
/* kind: 1 host->device, 2 device->host, 3 device->device. Returns when the
 * copy is complete. */
HWYBLAS_AMD_API void hwyblas_amd_memcpy(void* dst, const void* src,
                                        size_t bytes, int kind);
HWYBLAS_AMD_API void hwyblas_amd_memset(void* dst, int value, size_t bytes);
HWYBLAS_AMD_API void hwyblas_amd_synchronize(void);
/* The calling thread's library stream (used by the Fortran/CBLAS routines). */
HWYBLAS_AMD_API void* hwyblas_amd_thread_stream(void);
HWYBLAS_AMD_API void* hwyblas_amd_stream_create(void);
HWYBLAS_AMD_API void hwyblas_amd_stream_destroy(void* stream);
HWYBLAS_AMD_API void hwyblas_amd_stream_synchronize(void* stream);
HWYBLAS_AMD_API void* hwyblas_amd_event_create(void);
HWYBLAS_AMD_API void hwyblas_amd_event_destroy(void* event);
HWYBLAS_AMD_API void hwyblas_amd_event_record(void* event, void* stream);
/* Milliseconds between two recorded events (waits for the second). */
HWYBLAS_AMD_API float hwyblas_amd_event_elapsed_ms(void* start, void* stop);
/* x(i, j) = fma(u, hi - lo, lo) for i < rows, j < cols of a column-major
 * device matrix; u in [0,1) from the counter-based splitmix64 stream `seed`
 * at compact index i + j * rows (the same values as the CPU oracle's
 * generator). dtype_bytes is 8 (double) or 4 (float, rounded from double). */
HWYBLAS_AMD_API void hwyblas_amd_fill_uniform(void* x, int dtype_bytes,
                                              int64_t rows, int64_t cols,
                                              int64_t ld, uint64_t seed,
                                              double lo, double hi,
                                              void* stream);
/* Per-phase timing of the calling thread's last emulated DGEMM, for
 * profiling (off by default): 0 operand statistics and moduli choice,
 * 1 scaling and INT8 slicing, 2 the INT8 GEMMs (one launch for all moduli),
 * 3 CRT reconstruction and epilogue. HIP events are recorded on the call's
 * stream; hwyblas_amd_profile_ms waits for them and returns milliseconds, or
 * -1 if nothing was recorded. */
HWYBLAS_AMD_API void hwyblas_amd_profile_enable(int on);
HWYBLAS_AMD_API float hwyblas_amd_profile_ms(int phase);
/* Seeded test data: x(i, j) = +-u 2^e with u uniform in [1, 2) (52 random
 * bits), random sign and e a uniform integer in [emin, emax], a function of
 * the seed and the compact index i + j * rows only. */
HWYBLAS_AMD_API void hwyblas_amd_fill_pow2(double* x, int64_t rows, int64_t cols, int64_t ld,
                                           uint64_t seed, int emin, int emax, void* stream);
/* x(i, j) = (x(i, j) * row_scale[i]) * col_scale[j] (device vectors; NULL
 * skips one). */
HWYBLAS_AMD_API void hwyblas_amd_scale_rows_cols(double* x, int64_t rows, int64_t cols,
                                                 int64_t ld, const double* row_scale,
                                                 const double* col_scale, void* stream);
/* Stream capture into a hipGraph (for loops of calls on device pointers):
 * begin on a stream, enqueue calls on it, end to get an instantiated graph,
 * launch it any number of times, destroy it. */
HWYBLAS_AMD_API void hwyblas_amd_capture_begin(void* stream);
HWYBLAS_AMD_API void* hwyblas_amd_capture_end(void* stream);
HWYBLAS_AMD_API void hwyblas_amd_graph_launch(void* exec, void* stream);
HWYBLAS_AMD_API void hwyblas_amd_graph_destroy(void* exec);
/* ------------------------------------------------------------------------ */
/* Complex GEMM (hwyBLAS src/gemm.cc:73-132, :160-176; signatures as
 * include/hwyblas/hwyblas.h:576-587 and cblas.h:623-638). C := alpha op(A)
 * op(B) + beta C, op = N, T or C (conjugate transpose); matrices of
 * interleaved (re, im) reals with leading dimensions in complex elements,
 * alpha and beta as pointers to (re, im). Info positions as for ?gemm_
 * (1 transa, 2 transb, 3 m, 4 n, 5 k, 8 lda, 10 ldb, 13 ldc). Host or device
 * operands; the Fortran and CBLAS forms synchronize. Workspace from the
 * calling thread's device buffers: 2 (m + n) k + 4 m n reals. */
HWYBLAS_AMD_API void cgemm_(const char* transa, const char* transb,
                            const hwyblas_int* m, const hwyblas_int* n,
                            const hwyblas_int* k, const float* alpha,
                            const float* a, const hwyblas_int* lda, const float* b,
                            const hwyblas_int* ldb, const float* beta, float* c,
                            const hwyblas_int* ldc);
HWYBLAS_AMD_API void zgemm_(const char* transa, const char* transb,
                            const hwyblas_int* m, const hwyblas_int* n,
                            const hwyblas_int* k, const double* alpha,
                            const double* a, const hwyblas_int* lda,
                            const double* b, const hwyblas_int* ldb,
                            const double* beta, double* c, const hwyblas_int* ldc);
/* Row-major swaps A and B and m and n (src/cblas.cc:517-537, :1607-1630). */
HWYBLAS_AMD_API void cblas_cgemm(const CBLAS_LAYOUT layout,
                                 const CBLAS_TRANSPOSE transa,
                                 const CBLAS_TRANSPOSE transb, const hwyblas_int m,
                                 const hwyblas_int n, const hwyblas_int k,
                                 const void* alpha, const void* a,
                                 const hwyblas_int lda, const void* b,
                                 const hwyblas_int ldb, const void* beta, void* c,
                                 const hwyblas_int ldc);
HWYBLAS_AMD_API void cblas_zgemm(const CBLAS_LAYOUT layout,
                                 const CBLAS_TRANSPOSE transa,
                                 const CBLAS_TRANSPOSE transb, const hwyblas_int m,
                                 const hwyblas_int n, const hwyblas_int k,
                                 const void* alpha, const void* a,
                                 const hwyblas_int lda, const void* b,
                                 const hwyblas_int ldb, const void* beta, void* c,
                                 const hwyblas_int ldc);
/* Device-pointer forms (no hwyBLAS counterpart), as hwyblas_dgemm_dev with
 * alpha and beta as host pointers to (re, im). Asynchronous on `stream`;
 * calls that share the calling thread's workspace must be ordered (one
 * stream, or a synchronization between them). */
HWYBLAS_AMD_API int hwyblas_cgemm_dev(char transa, char transb, int m, int n, int k,
                                      const float* alpha, const float* a, int lda,
                                      const float* b, int ldb, const float* beta, float* c,
                                      int ldc, void* stream);
HWYBLAS_AMD_API int hwyblas_zgemm_dev(char transa, char transb, int m, int n, int k,
                                      const double* alpha, const double* a, int lda,
                                      const double* b, int ldb, const double* beta,
                                      double* c, int ldc, void* stream);

/* Complex TRSM (hwyBLAS src/trsm.cc:87-180; signatures as
 * include/hwyblas/hwyblas.h:797-806 and cblas.h:872-885): op(A) X = alpha B
 * (side L) or X op(A) = alpha B (side R), op = N, T or C, interleaved
 * (re, im) storage, alpha as a pointer to (re, im). Info positions as for
 * ?trsm_ (1 side, 2 uplo, 3 transa, 4 diag, 5 m, 6 n, 9 lda, 11 ldb);
 * alpha == 0 sets B to zero without reading A. Host or device operands; the
 * Fortran and CBLAS forms synchronize. Row-major CBLAS flips side and uplo
 * and swaps m and n. */
HWYBLAS_AMD_API void ctrsm_(const char* side, const char* uplo, const char* transa,
                            const char* diag, const hwyblas_int* m, const hwyblas_int* n,
                            const float* alpha, const float* a, const hwyblas_int* lda,
                            float* b, const hwyblas_int* ldb);
HWYBLAS_AMD_API void ztrsm_(const char* side, const char* uplo, const char* transa,
                            const char* diag, const hwyblas_int* m, const hwyblas_int* n,
                            const double* alpha, const double* a, const hwyblas_int* lda,
                            double* b, const hwyblas_int* ldb);
HWYBLAS_AMD_API void cblas_ctrsm(const CBLAS_LAYOUT layout, const CBLAS_SIDE side,
                                 const CBLAS_UPLO uplo, const CBLAS_TRANSPOSE transa,
                                 const CBLAS_DIAG diag, const hwyblas_int m,
                                 const hwyblas_int n, const void* alpha, const void* a,
                                 const hwyblas_int lda, void* b, const hwyblas_int ldb);
HWYBLAS_AMD_API void cblas_ztrsm(const CBLAS_LAYOUT layout, const CBLAS_SIDE side,
                                 const CBLAS_UPLO uplo, const CBLAS_TRANSPOSE transa,
                                 const CBLAS_DIAG diag, const hwyblas_int m,
                                 const hwyblas_int n, const void* alpha, const void* a,
                                 const hwyblas_int lda, void* b, const hwyblas_int ldb);
/* Device-pointer forms (no hwyBLAS counterpart), as hwyblas_dtrsm_dev with
 * alpha as a host pointer to (re, im); the same workspace rule as
 * hwyblas_zgemm_dev. */
HWYBLAS_AMD_API int hwyblas_ctrsm_dev(char side, char uplo, char transa, char diag, int m, int n,
                                      const float* alpha, const float* a, int lda, float* b,
                                      int ldb, void* stream);
HWYBLAS_AMD_API int hwyblas_ztrsm_dev(char side, char uplo, char transa, char diag, int m, int n,
                                      const double* alpha, const double* a, int lda, double* b,
                                      int ldb, void* stream);

/* ------------------------------------------------------------------------ */
/* LAPACKE for the Cholesky routines (hwyBLAS src/lapack/lapacke.cc:112-133,
 * 2065-2092, src/lapack/lapacke_symmetric.cc:22-70; signatures as
 * include/hwyblas/lapacke.h:100-106, 620-740). matrix_layout 101 (row-major)
 * or 102 (column-major); an invalid layout returns -1. Unless the NaN check is
 * off (LAPACKE_set_nancheck(0), or LAPACKE_NANCHECK=0 in the environment), a
 * NaN in A's referenced triangle or in B returns minus that argument's
 * position. Otherwise the return is the routine's info, one lower for an
 * invalid argument. Row-major operands are copied to column-major
 * temporaries: a short leading dimension returns minus its position in the
 * _work form. Host operands. */
HWYBLAS_AMD_API void LAPACKE_xerbla(const char* name, hwyblas_int info);
HWYBLAS_AMD_API void LAPACKE_set_nancheck(int flag);
HWYBLAS_AMD_API int LAPACKE_get_nancheck(void);
HWYBLAS_AMD_API hwyblas_int LAPACKE_spotrf(int matrix_layout, char uplo, hwyblas_int n, float* a,
                                           hwyblas_int lda);
HWYBLAS_AMD_API hwyblas_int LAPACKE_dpotrf(int matrix_layout, char uplo, hwyblas_int n, double* a,
                                           hwyblas_int lda);
HWYBLAS_AMD_API hwyblas_int LAPACKE_spotrf_work(int matrix_layout, char uplo, hwyblas_int n,
                                                float* a, hwyblas_int lda);
HWYBLAS_AMD_API hwyblas_int LAPACKE_dpotrf_work(int matrix_layout, char uplo, hwyblas_int n,
                                                double* a, hwyblas_int lda);
HWYBLAS_AMD_API hwyblas_int LAPACKE_spotrs(int matrix_layout, char uplo, hwyblas_int n,
                                           hwyblas_int nrhs, const float* a, hwyblas_int lda,
                                           float* b, hwyblas_int ldb);
HWYBLAS_AMD_API hwyblas_int LAPACKE_dpotrs(int matrix_layout, char uplo, hwyblas_int n,
                                           hwyblas_int nrhs, const double* a, hwyblas_int lda,
                                           double* b, hwyblas_int ldb);
HWYBLAS_AMD_API hwyblas_int LAPACKE_spotrs_work(int matrix_layout, char uplo, hwyblas_int n,
                                                hwyblas_int nrhs, const float* a,
                                                hwyblas_int lda, float* b, hwyblas_int ldb);
HWYBLAS_AMD_API hwyblas_int LAPACKE_dpotrs_work(int matrix_layout, char uplo, hwyblas_int n,
                                                hwyblas_int nrhs, const double* a,
                                                hwyblas_int lda, double* b, hwyblas_int ldb);
HWYBLAS_AMD_API hwyblas_int LAPACKE_sposv(int matrix_layout, char uplo, hwyblas_int n,
                                          hwyblas_int nrhs, float* a, hwyblas_int lda, float* b,
                                          hwyblas_int ldb);
HWYBLAS_AMD_API hwyblas_int LAPACKE_dposv(int matrix_layout, char uplo, hwyblas_int n,
                                          hwyblas_int nrhs, double* a, hwyblas_int lda, double* b,
                                          hwyblas_int ldb);
HWYBLAS_AMD_API hwyblas_int LAPACKE_sposv_work(int matrix_layout, char uplo, hwyblas_int n,
                                               hwyblas_int nrhs, float* a, hwyblas_int lda,
                                               float* b, hwyblas_int ldb);
HWYBLAS_AMD_API hwyblas_int LAPACKE_dposv_work(int matrix_layout, char uplo, hwyblas_int n,
                                               hwyblas_int nrhs, double* a, hwyblas_int lda,
                                               double* b, hwyblas_int ldb);

/* LU through LAPACK's Fortran and LAPACKE interfaces (hwyBLAS
 * src/lapack/getrf.cc:112-367, :500-540; src/lapack/lapacke_linear.cc:14-75;
 * signatures as include/hwyblas/lapack.h:93-95, 139-142, 158-161 and
 * lapacke.h). FP64 only: the factorization is hwyblas_dgetrf_dev with FP64
 * updates, the solve hwyblas_dgetrs_dev. Info positions as LAPACK's: dgetrf
 * 1 m, 2 n, 4 lda; dgetrs 1 trans, 2 n, 3 nrhs, 5 lda, 8 ldb; dgesv 1 n,
 * 2 nrhs, 4 lda, 7 ldb. Host or device operands (matrices and pivots);
 * each call synchronizes. info = -1000 (dgetrf_, dgesv_) is not an argument
 * position: a wait inside the cooperative panel kernel exceeded its bound, a
 * message names it on stderr, the installed error handler (if any) receives
 * ("DGETRF" or "DGESV", -1000), and A and ipiv hold no valid factorization.
 * The LAPACKE forms follow the rules stated for the Cholesky routines above
 * (host operands). */
HWYBLAS_AMD_API void dgetrf_(const hwyblas_int* m, const hwyblas_int* n, double* a,
                             const hwyblas_int* lda, hwyblas_int* ipiv, hwyblas_int* info);
HWYBLAS_AMD_API void dgetrs_(const char* trans, const hwyblas_int* n, const hwyblas_int* nrhs,
                             const double* a, const hwyblas_int* lda, const hwyblas_int* ipiv,
                             double* b, const hwyblas_int* ldb, hwyblas_int* info);
HWYBLAS_AMD_API void dgesv_(const hwyblas_int* n, const hwyblas_int* nrhs, double* a,
                            const hwyblas_int* lda, hwyblas_int* ipiv, double* b,
                            const hwyblas_int* ldb, hwyblas_int* info);
/* xGERFS: iterative refinement of X for op(A) X = B with the LU factors of
 * dgetrf_ (af, ipiv), with the componentwise backward error (berr) and the
 * forward error bound (ferr) of every column (hwyBLAS src/lapack/gesvx.cc:
 * 333-521; signature as include/hwyblas/lapack.h:465-471). Up to 5
 * corrections per column while berr > eps and halves; residuals in FP64 on
 * the GPU; ferr by xLACN2's estimator. Info positions as xGERFS's: 1 trans,
 * 2 n, 3 nrhs, 5 lda, 7 ldaf, 10 ldb, 12 ldx (reported as DGERFS). A, af, B,
 * X and ipiv may be host or device memory; ferr and berr are host arrays;
 * work and iwork are not used (the work arrays live on the device). Each
 * call synchronizes. */
HWYBLAS_AMD_API void dgerfs_(const char* trans, const hwyblas_int* n, const hwyblas_int* nrhs,
                             const double* a, const hwyblas_int* lda, const double* af,
                             const hwyblas_int* ldaf, const hwyblas_int* ipiv, const double* b,
                             const hwyblas_int* ldb, double* x, const hwyblas_int* ldx,
                             double* ferr, double* berr, double* work, hwyblas_int* iwork,
                             hwyblas_int* info);
HWYBLAS_AMD_API hwyblas_int LAPACKE_dgetrf(int matrix_layout, hwyblas_int m, hwyblas_int n,
                                           double* a, hwyblas_int lda, hwyblas_int* ipiv);
HWYBLAS_AMD_API hwyblas_int LAPACKE_dgetrf_work(int matrix_layout, hwyblas_int m, hwyblas_int n,
                                                double* a, hwyblas_int lda, hwyblas_int* ipiv);
HWYBLAS_AMD_API hwyblas_int LAPACKE_dgetrs(int matrix_layout, char trans, hwyblas_int n,
                                           hwyblas_int nrhs, const double* a, hwyblas_int lda,
                                           const hwyblas_int* ipiv, double* b, hwyblas_int ldb);
HWYBLAS_AMD_API hwyblas_int LAPACKE_dgetrs_work(int matrix_layout, char trans, hwyblas_int n,
                                                hwyblas_int nrhs, const double* a,
                                                hwyblas_int lda, const hwyblas_int* ipiv,
                                                double* b, hwyblas_int ldb);
HWYBLAS_AMD_API hwyblas_int LAPACKE_dgesv(int matrix_layout, hwyblas_int n, hwyblas_int nrhs,
                                          double* a, hwyblas_int lda, hwyblas_int* ipiv, double* b,
                                          hwyblas_int ldb);
HWYBLAS_AMD_API hwyblas_int LAPACKE_dgesv_work(int matrix_layout, hwyblas_int n, hwyblas_int nrhs,
                                               double* a, hwyblas_int lda, hwyblas_int* ipiv,
                                               double* b, hwyblas_int ldb);

/* Complex GEMV (hwyBLAS src/gemv.cc:39-94, src/cblas.cc:279-316, 1030-1053;
 * signatures as include/hwyblas/hwyblas.h:229-238 and cblas.h:241-252):
 * y := alpha op(A) x + beta y, op = N, T or C; interleaved (re, im) storage,
 * alpha and beta as
 * pointers to (re, im). Info positions as ?gemv_ (1 trans, 2 m, 3 n, 6 lda,
 * 8 incx, 11 incy). Row-major CBLAS ConjTrans runs as the conjugated,
 * untransposed product on the column-major storage (no conjugated copies of
 * x or y). Host or device operands; synchronous. */
HWYBLAS_AMD_API void cgemv_(const char* trans, const hwyblas_int* m, const hwyblas_int* n,
                            const float* alpha, const float* a, const hwyblas_int* lda,
                            const float* x, const hwyblas_int* incx, const float* beta, float* y,
                            const hwyblas_int* incy);
HWYBLAS_AMD_API void zgemv_(const char* trans, const hwyblas_int* m, const hwyblas_int* n,
                            const double* alpha, const double* a, const hwyblas_int* lda,
                            const double* x, const hwyblas_int* incx, const double* beta,
                            double* y, const hwyblas_int* incy);
HWYBLAS_AMD_API void cblas_cgemv(const CBLAS_LAYOUT layout, const CBLAS_TRANSPOSE trans,
                                 const hwyblas_int m, const hwyblas_int n, const void* alpha,
                                 const void* a, const hwyblas_int lda, const void* x,
                                 const hwyblas_int incx, const void* beta, void* y,
                                 const hwyblas_int incy);
HWYBLAS_AMD_API void cblas_zgemv(const CBLAS_LAYOUT layout, const CBLAS_TRANSPOSE trans,
                                 const hwyblas_int m, const hwyblas_int n, const void* alpha,
                                 const void* a, const hwyblas_int lda, const void* x,
                                 const hwyblas_int incx, const void* beta, void* y,
                                 const hwyblas_int incy);

/* Complex SYRK and HERK (hwyBLAS src/level3.cc:78-135; signatures as
 * include/hwyblas/hwyblas.h:642-663): C := alpha op(A) op(A)' + beta C on the
 * uplo triangle, ' = ^T (?syrk_: trans N or T, complex alpha and beta as
 * pointers to (re, im)) or ^H (?herk_: trans N or C, real alpha and beta,
 * the diagonal's imaginary parts set to zero). Info positions 1 uplo,
 * 2 trans, 3 n, 4 k, 7 lda, 10 ldc. The other triangle is not touched. The
 * product runs as the complex GEMM over all of C (twice the triangle's
 * flops). Host or device operands; synchronous. */
HWYBLAS_AMD_API void csyrk_(const char* uplo, const char* trans, const hwyblas_int* n,
                            const hwyblas_int* k, const float* alpha, const float* a,
                            const hwyblas_int* lda, const float* beta, float* c,
                            const hwyblas_int* ldc);
HWYBLAS_AMD_API void zsyrk_(const char* uplo, const char* trans, const hwyblas_int* n,
                            const hwyblas_int* k, const double* alpha, const double* a,
                            const hwyblas_int* lda, const double* beta, double* c,
                            const hwyblas_int* ldc);
HWYBLAS_AMD_API void cherk_(const char* uplo, const char* trans, const hwyblas_int* n,
                            const hwyblas_int* k, const float* alpha, const float* a,
                            const hwyblas_int* lda, const float* beta, float* c,
                            const hwyblas_int* ldc);
HWYBLAS_AMD_API void zherk_(const char* uplo, const char* trans, const hwyblas_int* n,
                            const hwyblas_int* k, const double* alpha, const double* a,
                            const hwyblas_int* lda, const double* beta, double* c,
                            const hwyblas_int* ldc);

/* Library version string. */
HWYBLAS_AMD_API const char* hwyblas_amd_version(void);

#ifdef __cplusplus
} /* extern "C" */
#endif

#endif /* HWYBLAS_AMD_H_ */
