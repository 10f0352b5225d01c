// The Level-3 routines hwyBLAS builds on its GEMM driver (src/level3.cc),
// for real types: SYMM, SYRK, SYR2K, GEMMTR (also under its GEMMT name) and
// TRMM. Validation order, info positions and quick returns follow
// src/level3.cc. The arithmetic is the GEMM kernel's:
//  - SYMM and TRMM expand their structured operand into a dense device copy
//    (the referenced triangle mirrored, or the other triangle zero and a
//    unit diagonal written as ones without reading it), then multiply;
//  - SYRK, SYR2K and GEMMTR compute and write only the referenced triangle
//    of C, skipping the tiles beyond the diagonal, as hwyBLAS's
//    GemmProblem::part does.
// Host operands are staged into the calling thread's device buffers as in
// blas_api.cpp; device operands are used in place. The triangle routines
// always copy C in, so its other triangle goes back unchanged.
#include <algorithm>

#include "entries.h"

namespace hwy_amd {
namespace {

template <typename T>
void CopyIn(T* dst, const T* src, int64_t rows, int64_t cols, int64_t ld, hipStream_t s) {
  HWY_HIP_CHECK(hwy_amd::Memcpy2DAsync(dst, rows * sizeof(T), src, ld * sizeof(T), rows * sizeof(T),
                                       cols, hipMemcpyHostToDevice, s));
}

template <typename T>
void CopyOut(T* dst, int64_t ld, const T* src, int64_t rows, int64_t cols, hipStream_t s) {
  HWY_HIP_CHECK(hwy_amd::Memcpy2DAsync(dst, ld * sizeof(T), src, rows * sizeof(T), rows * sizeof(T),
                                       cols, hipMemcpyDeviceToHost, s));
}

// A host operand (rows x cols, ld) staged into `slot` at leading dimension
// rows; a device operand as it is.
template <typename T>
const T* Operand(const T* x, int64_t rows, int64_t cols, int64_t ld, int slot, int64_t* ld_out,
                 hipStream_t s) {
  if (IsDevicePointer(x)) {
    *ld_out = ld;
    return x;
  }
  T* buf = static_cast<T*>(Staging(slot, size_t(std::max<int64_t>(1, rows * cols)) * sizeof(T)));
  CopyIn(buf, x, rows, cols, ld, s);
  *ld_out = rows;
  return buf;
}

enum class Expand { kSymmetric, kTriangular, kUnitTriangular };

// out (n x n, leading dimension n) := the structured n x n operand stored in
// one triangle of a: mirrored (symmetric), or with zeros in the other
// triangle (triangular; a unit diagonal is written as ones, not read).
template <typename T>
__global__ void ExpandKernel(int64_t n, const T* __restrict__ a, int64_t lda, bool lower,
                             Expand mode, T* __restrict__ out) {
  const int64_t total = n * n;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
       idx += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = idx % n, j = idx / n;
    const bool stored = lower ? i >= j : i <= j;
    T v;
    if (mode == Expand::kSymmetric) v = stored ? a[i + j * lda] : a[j + i * lda];
    else if (mode == Expand::kUnitTriangular && i == j) v = T(1);
    else v = stored ? a[i + j * lda] : T(0);
    out[idx] = v;
  }
}

template <typename T>
T* ExpandDevice(int64_t n, const T* a, int64_t lda, bool lower, Expand mode, hipStream_t s) {
  T* out = static_cast<T*>(Staging(kSlotD, size_t(n) * n * sizeof(T)));
  const int64_t blocks = std::min<int64_t>((n * n + 255) / 256, 65536);
  hipLaunchKernelGGL(ExpandKernel<T>, dim3(unsigned(blocks)), dim3(256), 0, s, n, a, lda, lower,
                     mode, out);
  HWY_HIP_CHECK(hipGetLastError());
  return out;
}

// C := beta C over one triangle of the n x n C (part 1 lower, 2 upper);
// beta == 0 writes zeros without reading (src/gemm.cc:48-64).
template <typename T>
__global__ void ScalePartKernel(int64_t n, T beta, T* __restrict__ c, int64_t ldc, int part) {
  const int64_t total = n * n;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
       idx += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = idx % n, j = idx / n;
    if (part == 1 ? i < j : i > j) continue;
    T* x = c + i + j * ldc;
    *x = (beta == T(0)) ? T(0) : *x * beta;
  }
}

template <typename T>
void ScalePartDevice(int64_t n, T beta, T* c, int64_t ldc, int part, hipStream_t s) {
  const int64_t blocks = std::min<int64_t>((n * n + 255) / 256, 65536);
  hipLaunchKernelGGL(ScalePartKernel<T>, dim3(unsigned(blocks)), dim3(256), 0, s, n, beta, c,
                     ldc, part);
  HWY_HIP_CHECK(hipGetLastError());
}

// C (rows x cols, ldc) on the device: the caller's array, or a staging
// copy (copied in when `copy_in`) that Finish copies back.
template <typename T>
struct OutputC {
  T* host = nullptr;
  int64_t host_ld = 0, rows = 0, cols = 0;
  T* dev = nullptr;
  int64_t ld = 0;
  OutputC(T* c, int64_t r, int64_t cl, int64_t ldc, bool copy_in, int slot, hipStream_t s)
      : rows(r), cols(cl) {
    if (IsDevicePointer(c)) {
      dev = c;
      ld = ldc;
    } else {
      host = c;
      host_ld = ldc;
      dev = static_cast<T*>(Staging(slot, size_t(std::max<int64_t>(1, r * cl)) * sizeof(T)));
      ld = r;
      if (copy_in) CopyIn(dev, c, r, cl, ldc, s);
    }
  }
  void Finish(hipStream_t s) {
    if (host) CopyOut(host, host_ld, dev, rows, cols, s);
    HWY_HIP_CHECK(hipStreamSynchronize(s));
  }
};

// src/level3.cc:25-72.
template <typename T>
void SymmHost(const Caller& who, const char* sidep, const char* uplop, const blasint* mp,
              const blasint* np, const T* alphap, const T* a, const blasint* ldap, const T* b,
              const blasint* ldbp, const T* betap, T* c, const blasint* ldcp) {
  const blasint m = *mp, n = *np, lda = *ldap, ldb = *ldbp, ldc = *ldcp;
  bool left, lower;
  const bool vs = DecodeSide(*sidep, &left);
  const bool vu = DecodeUplo(*uplop, &lower);
  const blasint ra = left ? m : n;
  blasint info = 0;
  if (!vs) info = 1;
  else if (!vu) info = 2;
  else if (m < 0) info = 3;
  else if (n < 0) info = 4;
  else if (lda < std::max(1, ra)) info = 7;
  else if (ldb < std::max(1, m)) info = 9;
  else if (ldc < std::max(1, m)) info = 12;
  if (info != 0) return Reject(who, info);
  const T alpha = *alphap, beta = *betap;
  if (m == 0 || n == 0 || (alpha == T(0) && beta == T(1))) return;
  hipStream_t s = Context().stream;
  OutputC<T> oc(c, m, n, ldc, beta != T(0), kSlotC, s);
  if (alpha == T(0)) {
    ScaleDevice<T>(m, n, beta, oc.dev, oc.ld, s);
  } else {
    int64_t lda_d, ldb_d;
    const T* da = Operand(a, ra, ra, lda, kSlotA, &lda_d, s);
    const T* db = Operand(b, m, n, ldb, kSlotB, &ldb_d, s);
    const T* sym = ExpandDevice<T>(ra, da, lda_d, lower, Expand::kSymmetric, s);
    if (left)
      GemmDevice<T>(false, false, m, n, ra, alpha, sym, ra, db, ldb_d, beta, oc.dev, oc.ld, s);
    else
      GemmDevice<T>(false, false, m, n, ra, alpha, db, ldb_d, sym, ra, beta, oc.dev, oc.ld, s);
  }
  oc.Finish(s);
}

// src/level3.cc:81-139 (real SYRK: trans N, T or C).
template <typename T>
void SyrkHost(const Caller& who, const char* uplop, const char* transp, const blasint* np,
              const blasint* kp, const T* alphap, const T* a, const blasint* ldap,
              const T* betap, T* c, const blasint* ldcp) {
  const blasint n = *np, k = *kp, lda = *ldap, ldc = *ldcp;
  bool lower;
  Op op;
  const bool vu = DecodeUplo(*uplop, &lower);
  const bool vt = DecodeOp(*transp, &op);
  const bool no_trans = op == Op::kNoTrans;
  const blasint ra = no_trans ? n : k;
  blasint info = 0;
  if (!vu) info = 1;
  else if (!vt) info = 2;
  else if (n < 0) info = 3;
  else if (k < 0) info = 4;
  else if (lda < std::max(1, ra)) info = 7;
  else if (ldc < std::max(1, n)) info = 10;
  if (info != 0) return Reject(who, info);
  const T alpha = *alphap, beta = *betap;
  const int part = lower ? 1 : 2;
  if (n == 0 || ((alpha == T(0) || k == 0) && beta == T(1))) return;
  hipStream_t s = Context().stream;
  OutputC<T> oc(c, n, n, ldc, true, kSlotC, s);
  if (alpha == T(0) || k == 0) {
    ScalePartDevice<T>(n, beta, oc.dev, oc.ld, part, s);
  } else {
    int64_t lda_d;
    const T* da = Operand(a, ra, no_trans ? k : n, lda, kSlotA, &lda_d, s);
    // op(A) op(A)^T: A A^T (N) or A^T A (T).
    GemmDevicePart<T>(part, !no_trans, no_trans, n, n, k, alpha, da, lda_d, da, lda_d, beta,
                      oc.dev, oc.ld, s);
  }
  oc.Finish(s);
}

// src/level3.cc:152-214 (real SYR2K): two passes over the triangle, the
// second adding alpha op(B) op(A)^T to the first's result.
template <typename T>
void Syr2kHost(const Caller& who, const char* uplop, const char* transp, const blasint* np,
               const blasint* kp, const T* alphap, const T* a, const blasint* ldap, const T* b,
               const blasint* ldbp, const T* betap, T* c, const blasint* ldcp) {
  const blasint n = *np, k = *kp, lda = *ldap, ldb = *ldbp, ldc = *ldcp;
  bool lower;
  Op op;
  const bool vu = DecodeUplo(*uplop, &lower);
  const bool vt = DecodeOp(*transp, &op);
  const bool no_trans = op == Op::kNoTrans;
  const blasint rows = no_trans ? n : k;
  blasint info = 0;
  if (!vu) info = 1;
  else if (!vt) info = 2;
  else if (n < 0) info = 3;
  else if (k < 0) info = 4;
  else if (lda < std::max(1, rows)) info = 7;
  else if (ldb < std::max(1, rows)) info = 9;
  else if (ldc < std::max(1, n)) info = 12;
  if (info != 0) return Reject(who, info);
  const T alpha = *alphap, beta = *betap;
  const int part = lower ? 1 : 2;
  if (n == 0 || ((alpha == T(0) || k == 0) && beta == T(1))) return;
  hipStream_t s = Context().stream;
  OutputC<T> oc(c, n, n, ldc, true, kSlotC, s);
  if (alpha == T(0) || k == 0) {
    ScalePartDevice<T>(n, beta, oc.dev, oc.ld, part, s);
  } else {
    const int64_t cols = no_trans ? k : n;
    int64_t lda_d, ldb_d;
    const T* da = Operand(a, rows, cols, lda, kSlotA, &lda_d, s);
    const T* db = Operand(b, rows, cols, ldb, kSlotB, &ldb_d, s);
    GemmDevicePart<T>(part, !no_trans, no_trans, n, n, k, alpha, da, lda_d, db, ldb_d, beta,
                      oc.dev, oc.ld, s);
    GemmDevicePart<T>(part, !no_trans, no_trans, n, n, k, alpha, db, ldb_d, da, lda_d, T(1),
                      oc.dev, oc.ld, s);
  }
  oc.Finish(s);
}

// src/level3.cc:219-268.
template <typename T>
void GemmtrHost(const Caller& who, const char* uplop, const char* transap, const char* transbp,
                const blasint* np, const blasint* kp, const T* alphap, const T* a,
                const blasint* ldap, const T* b, const blasint* ldbp, const T* betap, T* c,
                const blasint* ldcp) {
  const blasint n = *np, k = *kp, lda = *ldap, ldb = *ldbp, ldc = *ldcp;
  bool lower;
  Op oa, ob;
  const bool vu = DecodeUplo(*uplop, &lower);
  const bool va = DecodeOp(*transap, &oa);
  const bool vb = DecodeOp(*transbp, &ob);
  const blasint ra = oa == Op::kNoTrans ? n : k, rb = ob == Op::kNoTrans ? k : n;
  blasint info = 0;
  if (!vu) info = 1;
  else if (!va) info = 2;
  else if (!vb) info = 3;
  else if (n < 0) info = 4;
  else if (k < 0) info = 5;
  else if (lda < std::max(1, ra)) info = 8;
  else if (ldb < std::max(1, rb)) info = 10;
  else if (ldc < std::max(1, n)) info = 13;
  if (info != 0) return Reject(who, info);
  const T alpha = *alphap, beta = *betap;
  const int part = lower ? 1 : 2;
  if (n == 0 || ((alpha == T(0) || k == 0) && beta == T(1))) return;
  hipStream_t s = Context().stream;
  OutputC<T> oc(c, n, n, ldc, true, kSlotC, s);
  if (alpha == T(0) || k == 0) {
    ScalePartDevice<T>(n, beta, oc.dev, oc.ld, part, s);
  } else {
    const bool ta = oa == Op::kTrans, tb = ob == Op::kTrans;
    int64_t lda_d, ldb_d;
    const T* da = Operand(a, ra, ta ? n : k, lda, kSlotA, &lda_d, s);
    const T* db = Operand(b, rb, tb ? k : n, ldb, kSlotB, &ldb_d, s);
    GemmDevicePart<T>(part, ta, tb, n, n, k, alpha, da, lda_d, db, ldb_d, beta, oc.dev, oc.ld,
                      s);
  }
  oc.Finish(s);
}

// src/level3.cc:273-345: B := alpha op(A) B (left) or alpha B op(A), from a
// copy of B's input; alpha == 0 writes zeros without reading A or B.
template <typename T>
void TrmmHost(const Caller& who, const char* sidep, const char* uplop, const char* transp,
              const char* diagp, const blasint* mp, const blasint* np, const T* alphap,
              const T* a, const blasint* ldap, T* b, const blasint* ldbp) {
  const blasint m = *mp, n = *np, lda = *ldap, ldb = *ldbp;
  bool left, lower, unit;
  Op op;
  const bool vs = DecodeSide(*sidep, &left);
  const bool vu = DecodeUplo(*uplop, &lower);
  const bool vt = DecodeOp(*transp, &op);
  const bool vd = DecodeDiag(*diagp, &unit);
  const blasint ra = left ? m : n;
  blasint info = 0;
  if (!vs) info = 1;
  else if (!vu) info = 2;
  else if (!vt) info = 3;
  else if (!vd) info = 4;
  else if (m < 0) info = 5;
  else if (n < 0) info = 6;
  else if (lda < std::max(1, ra)) info = 9;
  else if (ldb < std::max(1, m)) info = 11;
  if (info != 0) return Reject(who, info);
  const T alpha = *alphap;
  if (m == 0 || n == 0) return;
  hipStream_t s = Context().stream;
  OutputC<T> ob(b, m, n, ldb, alpha != T(0), kSlotB, s);
  if (alpha == T(0)) {
    ScaleDevice<T>(m, n, T(0), ob.dev, ob.ld, s);
  } else {
    int64_t lda_d;
    const T* da = Operand(a, ra, ra, lda, kSlotA, &lda_d, s);
    const T* tri = ExpandDevice<T>(ra, da, lda_d, lower,
                                   unit ? Expand::kUnitTriangular : Expand::kTriangular, s);
    T* x = static_cast<T*>(Staging(kSlotC, size_t(m) * n * sizeof(T)));
    HWY_HIP_CHECK(hwy_amd::Memcpy2DAsync(x, size_t(m) * sizeof(T), ob.dev,
                                         size_t(ob.ld) * sizeof(T), size_t(m) * sizeof(T), n,
                                         hipMemcpyDeviceToDevice, s));
    // op(A) is upper triangular when (uplo U, N) or (uplo L, T). Its zero
    // k-range is skipped per tile, as hwyBLAS's packing does
    // (src/gemm-inl.h:66-130): left, row i needs k >= i (upper) or k <= i;
    // right, column j needs k <= j (upper) or k >= j.
    const bool t = op == Op::kTrans;
    const bool upper = lower == t;
    if (left)
      GemmDeviceKRange<T>(upper ? 1 : 2, t, false, m, n, m, alpha, tri, ra, x, m, T(0), ob.dev,
                          ob.ld, s);
    else
      GemmDeviceKRange<T>(upper ? 3 : 4, false, t, m, n, n, alpha, x, m, tri, ra, T(0), ob.dev,
                          ob.ld, s);
  }
  ob.Finish(s);
}

}  // namespace

#define HWY_DEFINE_CORES(T)                                                                   \
  void Trmm(const Caller& who, const char* side, const char* uplo, const char* transa,        \
            const char* diag, const blasint* m, const blasint* n, const T* alpha, const T* a, \
            const blasint* lda, T* b, const blasint* ldb) {                                   \
    TrmmHost(who, side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb);                     \
  }                                                                                           \
  void Symm(const Caller& who, const char* side, const char* uplo, const blasint* m,          \
            const blasint* n, const T* alpha, const T* a, const blasint* lda, const T* b,     \
            const blasint* ldb, const T* beta, T* c, const blasint* ldc) {                    \
    SymmHost(who, side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc);                     \
  }                                                                                           \
  void Syrk(const Caller& who, const char* uplo, const char* trans, const blasint* n,         \
            const blasint* k, const T* alpha, const T* a, const blasint* lda, const T* beta,  \
            T* c, const blasint* ldc) {                                                       \
    SyrkHost(who, uplo, trans, n, k, alpha, a, lda, beta, c, ldc);                            \
  }                                                                                           \
  void Syr2k(const Caller& who, const char* uplo, const char* trans, const blasint* n,        \
             const blasint* k, const T* alpha, const T* a, const blasint* lda, const T* b,    \
             const blasint* ldb, const T* beta, T* c, const blasint* ldc) {                   \
    Syr2kHost(who, uplo, trans, n, k, alpha, a, lda, b, ldb, beta, c, ldc);                   \
  }                                                                                           \
  void Gemmtr(const Caller& who, const char* uplo, const char* transa, const char* transb,    \
              const blasint* n, const blasint* k, const T* alpha, const T* a,                 \
              const blasint* lda, const T* b, const blasint* ldb, const T* beta, T* c,        \
              const blasint* ldc) {                                                           \
    GemmtrHost(who, uplo, transa, transb, n, k, alpha, a, lda, b, ldb, beta, c, ldc);         \
  }
HWY_DEFINE_CORES(float)
HWY_DEFINE_CORES(double)
#undef HWY_DEFINE_CORES

}  // namespace hwy_amd

using hwy_amd::blasint;

extern "C" {

HWYBLAS_AMD_API void ssymm_(const char* side, const char* uplo, const blasint* m,
                            const blasint* n, const float* alpha, const float* a,
                            const blasint* lda, const float* b, const blasint* ldb,
                            const float* beta, float* c, const blasint* ldc) {
  hwy_amd::Symm("SSYMM", side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc);
}
HWYBLAS_AMD_API void dsymm_(const char* side, const char* uplo, const blasint* m,
                            const blasint* n, const double* alpha, const double* a,
                            const blasint* lda, const double* b, const blasint* ldb,
                            const double* beta, double* c, const blasint* ldc) {
  hwy_amd::Symm("DSYMM", side, uplo, m, n, alpha, a, lda, b, ldb, beta, c, ldc);
}
HWYBLAS_AMD_API void ssyrk_(const char* uplo, const char* trans, const blasint* n,
                            const blasint* k, const float* alpha, const float* a,
                            const blasint* lda, const float* beta, float* c,
                            const blasint* ldc) {
  hwy_amd::Syrk("SSYRK", uplo, trans, n, k, alpha, a, lda, beta, c, ldc);
}
HWYBLAS_AMD_API void dsyrk_(const char* uplo, const char* trans, const blasint* n,
                            const blasint* k, const double* alpha, const double* a,
                            const blasint* lda, const double* beta, double* c,
                            const blasint* ldc) {
  hwy_amd::Syrk("DSYRK", uplo, trans, n, k, alpha, a, lda, beta, c, ldc);
}
HWYBLAS_AMD_API void ssyr2k_(const char* uplo, const char* trans, const blasint* n,
                             const blasint* k, const float* alpha, const float* a,
                             const blasint* lda, const float* b, const blasint* ldb,
                             const float* beta, float* c, const blasint* ldc) {
  hwy_amd::Syr2k("SSYR2K", uplo, trans, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}
HWYBLAS_AMD_API void dsyr2k_(const char* uplo, const char* trans, const blasint* n,
                             const blasint* k, const double* alpha, const double* a,
                             const blasint* lda, const double* b, const blasint* ldb,
                             const double* beta, double* c, const blasint* ldc) {
  hwy_amd::Syr2k("DSYR2K", uplo, trans, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}
HWYBLAS_AMD_API void sgemmtr_(const char* uplo, const char* transa, const char* transb,
                              const blasint* n, const blasint* k, const float* alpha,
                              const float* a, const blasint* lda, const float* b,
                              const blasint* ldb, const float* beta, float* c,
                              const blasint* ldc) {
  hwy_amd::Gemmtr("SGEMMTR", uplo, transa, transb, n, k, alpha, a, lda, b, ldb, beta, c,
                      ldc);
}
HWYBLAS_AMD_API void dgemmtr_(const char* uplo, const char* transa, const char* transb,
                              const blasint* n, const blasint* k, const double* alpha,
                              const double* a, const blasint* lda, const double* b,
                              const blasint* ldb, const double* beta, double* c,
                              const blasint* ldc) {
  hwy_amd::Gemmtr("DGEMMTR", uplo, transa, transb, n, k, alpha, a, lda, b, ldb, beta, c,
                      ldc);
}
HWYBLAS_AMD_API void sgemmt_(const char* uplo, const char* transa, const char* transb,
                             const blasint* n, const blasint* k, const float* alpha,
                             const float* a, const blasint* lda, const float* b,
                             const blasint* ldb, const float* beta, float* c,
                             const blasint* ldc) {
  hwy_amd::Gemmtr("SGEMMT", uplo, transa, transb, n, k, alpha, a, lda, b, ldb, beta, c,
                      ldc);
}
HWYBLAS_AMD_API void dgemmt_(const char* uplo, const char* transa, const char* transb,
                             const blasint* n, const blasint* k, const double* alpha,
                             const double* a, const blasint* lda, const double* b,
                             const blasint* ldb, const double* beta, double* c,
                             const blasint* ldc) {
  hwy_amd::Gemmtr("DGEMMT", uplo, transa, transb, n, k, alpha, a, lda, b, ldb, beta, c,
                      ldc);
}
HWYBLAS_AMD_API void strmm_(const char* side, const char* uplo, const char* transa,
                            const char* diag, const blasint* m, const blasint* n,
                            const float* alpha, const float* a, const blasint* lda, float* b,
                            const blasint* ldb) {
  hwy_amd::Trmm("STRMM", side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb);
}
HWYBLAS_AMD_API void dtrmm_(const char* side, const char* uplo, const char* transa,
                            const char* diag, const blasint* m, const blasint* n,
                            const double* alpha, const double* a, const blasint* lda,
                            double* b, const blasint* ldb) {
  hwy_amd::Trmm("DTRMM", side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb);
}

}  // extern "C"
