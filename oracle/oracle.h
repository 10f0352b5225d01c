/*
 * CPU oracle for the Level-3 hot path (?gemm_, ?trsm_) and its getrf/getrs/
 * gerfs consumer. TEST INFRASTRUCTURE ONLY: tests, the smoke check and the
 * CPU-baseline leg of the benchmark may load this library, as the checker;
 * the product library never does.
 *
 * Three groups of functions:
 *  - orc_fill_*      the counter-based generator shared with the device fill
 *                    kernel (element (i, j) of a matrix depends only on the
 *                    seed and the compact index i + j * rows);
 *  - orc_?gemm/trsm  a single-threaded restatement of hwyBLAS's algorithms
 *                    (Goto/BLIS blocked GEMM with an MR x NR FMA register tile,
 *                    recursive TRSM halving at multiples of 32 with a
 *                    reciprocal-multiply substitution base);
 *  - orc_*_ref_*     double-double recomputations with per-element error
 *                    scales, used to check any result against the
 *                    componentwise bound |C - C_ref| <= c * k * u * scale.
 *
 * Each function cites the reference file:line it follows (paths relative to
 * the hwyBLAS source tree).
 */
#ifndef HWYBLAS_AMD_ORACLE_H_
#define HWYBLAS_AMD_ORACLE_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- generator (shared bit-for-bit with the device fill kernel) -------- */
uint64_t orc_splitmix64(uint64_t x);
/* u in [0,1) with 53 random bits for element `index` of stream `seed`. */
double orc_uniform01(uint64_t seed, uint64_t index);
/* x(i,j) = fma(u, hi - lo, lo) for i < rows, j < cols; u from index i+j*rows. */
void orc_fill_uniform_f64(double* x, int64_t rows, int64_t cols, int64_t ld,
                          uint64_t seed, double lo, double hi);
void orc_fill_uniform_f32(float* x, int64_t rows, int64_t cols, int64_t ld,
                          uint64_t seed, double lo, double hi);
/* x(i,j) = +-u 2^e, u in [1,2) with 52 random bits, e uniform in [emin, emax]
 * (two hashes of the compact index; the sign is bit 0 of the first). */
void orc_fill_pow2(double* x, int64_t rows, int64_t cols, int64_t ld, uint64_t seed,
                   int emin, int emax);

/* ---- hwyBLAS restatement (column-major, Fortran semantics; arguments are
 * assumed valid: validation is the ABI layer's job) ------------------------ */
void orc_dgemm(char transa, char transb, int m, int n, int k, double alpha,
               const double* a, int lda, const double* b, int ldb,
               double beta, double* c, int ldc);
void orc_sgemm(char transa, char transb, int m, int n, int k, float alpha,
               const float* a, int lda, const float* b, int ldb, float beta,
               float* c, int ldc);
void orc_dtrsm(char side, char uplo, char transa, char diag, int m, int n,
               double alpha, const double* a, int lda, double* b, int ldb);
void orc_strsm(char side, char uplo, char transa, char diag, int m, int n,
               float alpha, const float* a, int lda, float* b, int ldb);

/* ---- double-double references --------------------------------------- */
/* expected/bound are compact m x n (leading dimension m). bound(i,j) =
 * |alpha| sum_p |a_ip b_pj| + |beta| |c_ij| + tiny, the scale of the
 * componentwise bound; C is not read when beta == 0. */
void orc_dgemm_ref(char transa, char transb, int m, int n, int k, double alpha,
                   const double* a, int lda, const double* b, int ldb,
                   double beta, const double* c, int ldc, double* expected,
                   double* bound);
void orc_sgemm_ref(char transa, char transb, int m, int n, int k, float alpha,
                   const float* a, int lda, const float* b, int ldb,
                   float beta, const float* c, int ldc, double* expected,
                   double* bound);
/* As orc_dgemm_ref for `count` entries (rows[t], cols[t]) only. */
void orc_dgemm_ref_entries(char transa, char transb, int k, double alpha,
                           const double* a, int lda, const double* b, int ldb,
                           double beta, const double* c, int ldc,
                           const int* rows, const int* cols, int count,
                           double* expected, double* bound);
void orc_sgemm_ref_entries(char transa, char transb, int k, float alpha,
                           const float* a, int lda, const float* b, int ldb,
                           float beta, const float* c, int ldc,
                           const int* rows, const int* cols, int count,
                           double* expected, double* bound);

/* TRSM residual check: for the `count` right-hand sides listed in `which`
 * (columns of B for side L, rows for side R), returns in ratio[t] the largest
 * |op(A) X - alpha B0| / (|op(A)| |X| + |alpha B0| + tiny) over that right-hand
 * side, computed in double-double. X has leading dimension ldx. */
void orc_dtrsm_residual(char side, char uplo, char transa, char diag, int m,
                        int n, double alpha, const double* a, int lda,
                        const double* b0, int ldb, const double* x, int ldx,
                        const int* which, int count, double* ratio);
void orc_strsm_residual(char side, char uplo, char transa, char diag, int m,
                        int n, float alpha, const float* a, int lda,
                        const float* b0, int ldb, const float* x, int ldx,
                        const int* which, int count, double* ratio);

/* ---- emulated DGEMM (Ozaki scheme II), restated on the CPU ---------------
 * The same algorithm as the product's INT8 emulation: per-row/column power-
 * of-two scaling to integers with bA + bB = floor(log2 P - 1 - log2 k) bits,
 * symmetric residues modulo the first nmod of the pairwise coprime moduli
 * 256, 255, 253, 251, 247, 241, 239, 233, 229, 227, 223, 217, 211, 199,
 * 197, 193, 191, 181, 179, 173, 167, 163, 157, 151, exact integer products
 * reduced mod p, and the CRT reconstruction through 160-bit fixed-point
 * fractions y_l / p_l (four 40-bit limbs summed exactly in doubles).
 * Deterministic bit for bit. */
int orc_ozaki_modulus(int l);
int orc_ozaki_total_bits(int nmod, int64_t k);
/* Reconstructs the centred integer X (as a double) from its residues
 * res[l] in [0, p_l), l < nmod. */
double orc_ozaki_crt(int nmod, const int* res);
void orc_dgemm_ozaki(char transa, char transb, int m, int n, int k, double alpha,
                     const double* a, int lda, const double* b, int ldb, double beta,
                     double* c, int ldc, int nmod);

/* ---- emulated DGEMM with BF16 slices (Ozaki scheme I), restated on the CPU
 * (oracle/emulated_bf16.cc) ------------------------------------------------
 * The product's algorithm: rows of op(A) / columns of op(B) scaled by
 * 2^-(E+1) (max magnitude f 2^E, f in [0.5, 1)), `slices` balanced 8-bit
 * digits per entry (d_t = rint(256 r), ties to even; r = 256 r - d_t;
 * |d_t| <= 128), the pairs t + u <= s + 1 by level from s + 1 down to 2
 * (t ascending), exact chunk sums over k chunks of 1024 accumulated as
 * acc = fma(chunk, 2^-8l, acc), then C = alpha 2^(E_i + F_j + 2) acc + beta C.
 * Deterministic bit for bit. */
void orc_bf16_digits(double x, int sc, int slices, int* digits);
void orc_dgemm_ozaki_bf16(char transa, char transb, int m, int n, int k, double alpha,
                          const double* a, int lda, const double* b, int ldb, double beta,
                          double* c, int ldc, int slices);

/* ---- the count choice of the emulated DGEMMs (oracle/emulated_choice.cc) --
 * Exact statistics of every pair (i, j): the upper row/column sums, the
 * coarse INT8 product G that bounds |A||B| from below, X and gmin; the error
 * bound of a count in units of |A||B|(i, j); the rule that picks the count
 * (kind 0: INT8 moduli, 1: BF16 slices; 0 returned: native FP64). */
int orc_emulated_choose(int kind, int forced, char transa, char transb, int m, int n, int k,
                        const double* a, int lda, const double* b, int ldb, double* xmax,
                        int* gmin);
void orc_emulated_bound(int kind, int count, int64_t k, double xmax, int gmin, double* typical,
                        double* worst);
/* lower[i + j m] = 2^(E_i + F_j - 14) G_ij; rowsum_up[i] = rs_i 2^(E_i - 40),
 * colsum_up[j] = cs_j 2^(F_j - 40). */
void orc_emulated_coarse(char transa, char transb, int m, int n, int k, const double* a,
                         int lda, const double* b, int ldb, double* lower, double* rowsum_up,
                         double* colsum_up);

/* ---- Householder QR (oracle/qr.cc) ------------------------------------------
 * xGEQRF (NB 32, NX 128), xGEQR2 and xLARFT ('F', 'C') as hwyBLAS's LAPACK
 * port computes them: R on and above the diagonal of A, the reflectors'
 * vectors below it (unit first element implicit) and tau. */
int orc_dgeqrf(int m, int n, double* a, int lda, double* tau);
void orc_dgeqr2(int m, int n, double* a, int lda, double* tau);
void orc_dlarft(int n, int k, const double* v, int ldv, const double* tau, double* t, int ldt);

/* ---- LU with partial pivoting and its solve (oracle/lu.cc) ---------------
 * hwyBLAS's dgetrf (blocked NB = 64 over the recursive dgetrf2 with a leaf of
 * 16), dlaswp (1-based k1..k2, incx = +-1), dgetrs and gerfs's componentwise
 * backward error. ipiv is 1-based. */
int orc_dgetrf(int m, int n, double* a, int lda, int* ipiv);
void orc_dlaswp(int n, double* a, int lda, int k1, int k2, const int* ipiv, int incx);
void orc_dgetrs(char trans, int n, int nrhs, const double* a, int lda, const int* ipiv,
                double* b, int ldb);
void orc_dgerfs_berr(int n, int nrhs, const double* a, int lda, const double* b, int ldb,
                     const double* x, int ldx, double* berr);
/* xLACN2 (reverse-communication 1-norm estimate; kase 1: x := M x, kase 2:
 * x := M^T x; isave[3] keeps the state) and xGERFS (refines X in place with
 * the LU af/ipiv of A; per column the forward error bound ferr and the
 * componentwise backward error berr; trans N, T or C; returns -1 for any
 * other trans). */
void orc_dlacn2(int n, double* v, double* x, int* isgn, double* est, int* kase, int* isave);
int orc_dgerfs(char trans, int n, int nrhs, const double* a, int lda, const double* af,
               int ldaf, const int* ipiv, const double* b, int ldb, double* x, int ldx,
               double* ferr, double* berr);
/* xPOTRF2 (Cholesky by recursive halving down to n = 1) in place: the uplo
 * triangle of A becomes L (A = L L^T) or U (A = U^T U); returns 0 or the
 * order of the first non-positive (or NaN) leading minor. */
int orc_dpotrf2(char uplo, int n, double* a, int lda);

/* ---- planted mistakes (for tests that show the checks catch them) ------- */
/* 0: none; 1: drop the last k term of every GEMM tile; 2: use op(A) with
 * the wrong transposition; 3: ignore beta; 4: TRSM skips the reciprocal of
 * the diagonal; 5: the double-double reference drops its low word; 6: the
 * emulated DGEMM's CRT drops the last modulus; 7: the LU leaf does not swap
 * the columns left of the pivot column; 8: the BF16-slice emulation leaves
 * out its top level of slice pairs (t + u = s + 1); 9: the 1-norm estimator
 * skips its alternating-vector stage; 10: gerfs's forward error bound solves
 * with op(A) where op(A)^T belongs; 11: that bound takes nz = n; 12: the
 * Cholesky update multiplies a row of the off-diagonal block by itself. */
void orc_plant_mistake(int which);

#ifdef __cplusplus
}
#endif

#endif /* HWYBLAS_AMD_ORACLE_H_ */
