// Fortran-convention entry points (sgemm_, dgemm_, strsm_, dtrsm_) and their
// device-pointer forms.
//
// Validation order, info positions and quick returns follow hwyBLAS
// (src/gemm.cc:73-132, src/trsm.cc:137-180). Host operands are staged into
// the calling thread's device buffers: only the referenced rows x cols of
// each operand travel (at compact leading dimension), C is copied in only
// when beta != 0, and only rows [0, m) of each column are copied back, so
// the padding of C between m and ldc is never written. Device operands
// (hipMalloc or managed memory) are used in place.
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <thread>
#include <type_traits>
#include <vector>

#include "entries.h"

namespace hwy_amd {

// How DGEMM-shaped work on the drop-in entries runs: 0 native FP64 (the
// default), 1 the emulated DGEMM on the INT8 matrix cores with the count
// chosen on the device and hwyBLAS's componentwise bound guaranteed (the
// native kernel where it cannot be certified), for products large enough
// to gain (m, n >= 2048 and k >= 1024). The environment variable
// HWYBLAS_AMD_DGEMM=emulated sets 1 at load; hwyblas_set_dgemm_mode changes
// it for every thread.
std::atomic<int>& DgemmModeFlag() {
  static std::atomic<int> mode{[] {
    const char* e = getenv("HWYBLAS_AMD_DGEMM");
    return (e && strcmp(e, "emulated") == 0) ? 1 : 0;
  }()};
  return mode;
}

int DgemmMode() { return DgemmModeFlag().load(std::memory_order_relaxed); }

namespace {

// C = alpha op(A) op(B) + beta C on device operands, by the current mode.
// `split`: whether a skinny product may split k (the host pipeline's column
// blocks may not: their bits must be the whole product's).
template <typename T>
void GemmByMode(bool ta, bool tb, int64_t m, int64_t n, int64_t k, T alpha, const T* a,
                int64_t lda, const T* b, int64_t ldb, T beta, T* c, int64_t ldc, hipStream_t s,
                bool split = true) {
  if constexpr (std::is_same_v<T, double>) {
    if (DgemmMode() == 1 && m >= 2048 && n >= 2048 && k >= 1024) {
      ozaki::EmulatedDgemm(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, 0, s, nullptr);
      return;
    }
  }
  // Skinny products with a long k (too few tiles of C to fill the device):
  // k split in ranges computed at once, the partials summed in order.
  if (split && GemmSplitPays<T>(m, n, k)) {
    GemmSplitKPublic<T>(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, s);
    return;
  }
  GemmDevice<T>(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, s);
}

template <typename T>
void CopyIn(T* dst, const T* src, int64_t rows, int64_t cols, int64_t ld,
            hipStream_t s) {
  HWY_HIP_CHECK(hwy_amd::Memcpy2DAsync(dst, rows * sizeof(T), src, ld * sizeof(T), rows * sizeof(T),
                                       cols, hipMemcpyHostToDevice, s));
}

template <typename T>
void CopyOut(T* dst, int64_t ld, const T* src, int64_t rows, int64_t cols,
             hipStream_t s) {
  HWY_HIP_CHECK(hwy_amd::Memcpy2DAsync(dst, ld * sizeof(T), src, rows * sizeof(T), rows * sizeof(T),
                                       cols, hipMemcpyDeviceToHost, s));
}

// Returns the Fortran info of the first invalid GEMM argument, or 0
// (src/gemm.cc:84-106).
blasint GemmInfo(char transa, char transb, blasint m, blasint n, blasint k,
                 blasint lda, blasint ldb, blasint ldc, Op* oa, Op* ob) {
  const bool va = DecodeOp(transa, oa);
  const bool vb = DecodeOp(transb, ob);
  const blasint rows_a = (*oa == Op::kNoTrans) ? m : k;
  const blasint rows_b = (*ob == Op::kNoTrans) ? k : n;
  if (!va) return 1;
  if (!vb) return 2;
  if (m < 0) return 3;
  if (n < 0) return 4;
  if (k < 0) return 5;
  if (lda < std::max(1, rows_a)) return 8;
  if (ldb < std::max(1, rows_b)) return 10;
  if (ldc < std::max(1, m)) return 13;
  return 0;
}

// Returns the Fortran info of the first invalid TRSM argument, or 0
// (src/trsm.cc:145-168).
blasint TrsmInfo(char side, char uplo, char transa, char diag, blasint m,
                 blasint n, blasint lda, blasint ldb, bool* left, bool* lower,
                 Op* op, bool* unit) {
  const bool vs = DecodeSide(side, left);
  const bool vu = DecodeUplo(uplo, lower);
  const bool vt = DecodeOp(transa, op);
  const bool vd = DecodeDiag(diag, unit);
  const blasint rows_a = *left ? m : n;
  if (!vs) return 1;
  if (!vu) return 2;
  if (!vt) return 3;
  if (!vd) return 4;
  if (m < 0) return 5;
  if (n < 0) return 6;
  if (lda < std::max(1, rows_a)) return 9;
  if (ldb < std::max(1, m)) return 11;
  return 0;
}

// Large products on host operands, pipelined over blocks of C's columns:
// a helper thread copies op(B)'s block (and C's, when beta != 0) in on one
// stream, the calling thread runs the block's product on its own stream,
// and a second helper copies the block of C out on a third, through two
// device buffers per operand, so that the copies of neighbouring blocks
// overlap the product (copies from pageable memory block the thread that
// issues them, hence one thread per direction). A goes in first. Events
// order the streams; host counters tell each thread which events have been
// recorded. Every block runs the same k order as the whole product, so the
// bits equal the unblocked call's. False (nothing done) when the product is
// small, an operand is device memory, or HWYBLAS_AMD_HOST_PIPELINE=0 was set
// at the first call (the helper threads are then never created).
template <typename T>
bool GemmHostPipelined(bool ta, bool tb, int64_t m, int64_t n, int64_t k, T alpha, const T* a,
                       int64_t lda, const T* b, int64_t ldb, T beta, T* c, int64_t ldc) {
  constexpr int64_t kMinBlock = 1024;
  if (m * n * k < (int64_t(1) << 36) || n < 4 * kMinBlock) return false;
  // A product the device call splits over k: its column blocks would not.
  if (GemmSplitPays<T>(m, n, k)) return false;
  static const bool enabled = [] {
    const char* e = getenv("HWYBLAS_AMD_HOST_PIPELINE");
    return !(e && strcmp(e, "0") == 0);
  }();
  if (!enabled) return false;
  if (IsDevicePointer(a) || IsDevicePointer(b) || IsDevicePointer(c)) return false;
  const int64_t nb = std::max<int64_t>(kMinBlock, ((n + 7) / 8 + 255) / 256 * 256);
  const int nblk = int((n + nb - 1) / nb);
  int dev = 0;
  HWY_HIP_CHECK(hipGetDevice(&dev));
  ThreadContext& ctx = Context();
  if (ctx.copy_in == nullptr) {
    HWY_HIP_CHECK(hipStreamCreateWithFlags(&ctx.copy_in, hipStreamNonBlocking));
    HWY_HIP_CHECK(hipStreamCreateWithFlags(&ctx.copy_out, hipStreamNonBlocking));
  }
  hipStream_t s = ctx.stream, sin = ctx.copy_in, sout = ctx.copy_out;
  const int64_t ra = ta ? k : m, ca = ta ? m : k;
  T* da = static_cast<T*>(Staging(kSlotA, size_t(ra) * ca * sizeof(T)));
  T* bring = static_cast<T*>(Staging(kSlotB, size_t(2) * k * nb * sizeof(T)));
  T* cring = static_cast<T*>(Staging(kSlotC, size_t(2) * m * nb * sizeof(T)));
  auto width = [&](int j) { return std::min(nb, n - j * nb); };
  auto bslot = [&](int j) { return bring + (j % 2) * k * nb; };
  auto cslot = [&](int j) { return cring + (j % 2) * m * nb; };

  std::vector<hipEvent_t> ev_in(nblk), ev_gemm(nblk), ev_out(nblk);
  for (int j = 0; j < nblk; ++j) {
    HWY_HIP_CHECK(hipEventCreateWithFlags(&ev_in[j], hipEventDisableTiming));
    HWY_HIP_CHECK(hipEventCreateWithFlags(&ev_gemm[j], hipEventDisableTiming));
    HWY_HIP_CHECK(hipEventCreateWithFlags(&ev_out[j], hipEventDisableTiming));
  }
  // Blocks whose copy-in / product / copy-out event has been recorded.
  std::atomic<int> in_rec{0}, gemm_rec{0}, out_rec{0};
  auto wait_for = [](const std::atomic<int>& count, int need) {
    while (count.load(std::memory_order_acquire) < need) std::this_thread::yield();
  };

  CopyIn(da, a, ra, ca, lda, s);
  std::thread copier_in([&] {
    HWY_HIP_CHECK(hipSetDevice(dev));
    for (int j = 0; j < nblk; ++j) {
      const int64_t j0 = j * nb, w = width(j);
      if (j >= 2) {  // the slots of block j - 2 are free once its product (and copy-out) ran
        wait_for(gemm_rec, j - 1);
        HWY_HIP_CHECK(hipStreamWaitEvent(sin, ev_gemm[j - 2], 0));
        if (beta != T(0)) {
          wait_for(out_rec, j - 1);
          HWY_HIP_CHECK(hipStreamWaitEvent(sin, ev_out[j - 2], 0));
        }
      }
      if (tb) CopyIn(bslot(j), b + j0, w, k, ldb, sin);  // rows j0.. of B: w x k
      else CopyIn(bslot(j), b + j0 * ldb, k, w, ldb, sin);
      if (beta != T(0)) CopyIn(cslot(j), c + j0 * ldc, m, w, ldc, sin);
      HWY_HIP_CHECK(hipEventRecord(ev_in[j], sin));
      in_rec.store(j + 1, std::memory_order_release);
    }
  });
  std::thread copier_out([&] {
    HWY_HIP_CHECK(hipSetDevice(dev));
    for (int j = 0; j < nblk; ++j) {
      wait_for(gemm_rec, j + 1);
      HWY_HIP_CHECK(hipStreamWaitEvent(sout, ev_gemm[j], 0));
      CopyOut(c + j * nb * ldc, ldc, cslot(j), m, width(j), sout);
      HWY_HIP_CHECK(hipEventRecord(ev_out[j], sout));
      out_rec.store(j + 1, std::memory_order_release);
    }
  });
  for (int j = 0; j < nblk; ++j) {
    const int64_t w = width(j);
    wait_for(in_rec, j + 1);
    HWY_HIP_CHECK(hipStreamWaitEvent(s, ev_in[j], 0));
    if (j >= 2 && beta == T(0)) {  // C's slot is free once block j - 2 went out
      wait_for(out_rec, j - 1);
      HWY_HIP_CHECK(hipStreamWaitEvent(s, ev_out[j - 2], 0));
    }
    GemmByMode<T>(ta, tb, m, w, k, alpha, da, ra, bslot(j), tb ? w : k, beta, cslot(j), m, s,
                  false);
    HWY_HIP_CHECK(hipEventRecord(ev_gemm[j], s));
    gemm_rec.store(j + 1, std::memory_order_release);
  }
  copier_in.join();
  copier_out.join();
  HWY_HIP_CHECK(hipStreamSynchronize(sout));
  HWY_HIP_CHECK(hipStreamSynchronize(sin));
  HWY_HIP_CHECK(hipStreamSynchronize(s));
  for (int j = 0; j < nblk; ++j) {
    HWY_HIP_CHECK(hipEventDestroy(ev_in[j]));
    HWY_HIP_CHECK(hipEventDestroy(ev_gemm[j]));
    HWY_HIP_CHECK(hipEventDestroy(ev_out[j]));
  }
  return true;
}

template <typename T>
void GemmHost(const Caller& who, const char* transa, const char* transb,
              const blasint* mp, const blasint* np, const blasint* kp,
              const T* alphap, const T* a, const blasint* ldap, const T* b,
              const blasint* ldbp, const T* betap, T* c, const blasint* ldcp) {
  const blasint m = *mp, n = *np, k = *kp, lda = *ldap, ldb = *ldbp, ldc = *ldcp;
  Op oa, ob;
  const blasint info = GemmInfo(*transa, *transb, m, n, k, lda, ldb, ldc, &oa, &ob);
  if (info != 0) {
    Reject(who, info);
    return;
  }
  const T alpha = *alphap, beta = *betap;
  if (m == 0 || n == 0) return;
  if ((k == 0 || alpha == T(0)) && beta == T(1)) return;

  if (k != 0 && alpha != T(0) &&
      GemmHostPipelined<T>(oa == Op::kTrans, ob == Op::kTrans, m, n, k, alpha, a, lda, b, ldb,
                           beta, c, ldc))
    return;
  ThreadContext& ctx = Context();
  hipStream_t s = ctx.stream;
  const bool c_dev = IsDevicePointer(c);
  T* dc = c;
  int64_t ldc_d = ldc;
  if (!c_dev) {
    dc = static_cast<T*>(Staging(kSlotC, size_t(m) * n * sizeof(T)));
    ldc_d = m;
    if (beta != T(0)) CopyIn(dc, c, m, n, ldc, s);
  }
  if (k == 0 || alpha == T(0)) {
    ScaleDevice<T>(m, n, beta, dc, ldc_d, s);
  } else {
    const bool ta = oa == Op::kTrans, tb = ob == Op::kTrans;
    const int64_t ra = ta ? k : m, ca = ta ? m : k;
    const int64_t rb = tb ? n : k, cb = tb ? k : n;
    const T* da = a;
    const T* db = b;
    int64_t lda_d = lda, ldb_d = ldb;
    if (!IsDevicePointer(a)) {
      T* buf = static_cast<T*>(Staging(kSlotA, size_t(ra) * ca * sizeof(T)));
      CopyIn(buf, a, ra, ca, lda, s);
      da = buf;
      lda_d = ra;
    }
    if (!IsDevicePointer(b)) {
      T* buf = static_cast<T*>(Staging(kSlotB, size_t(rb) * cb * sizeof(T)));
      CopyIn(buf, b, rb, cb, ldb, s);
      db = buf;
      ldb_d = rb;
    }
    GemmByMode<T>(ta, tb, m, n, k, alpha, da, lda_d, db, ldb_d, beta, dc, ldc_d, s);
  }
  if (!c_dev) CopyOut(c, ldc, dc, m, n, s);
  HWY_HIP_CHECK(hipStreamSynchronize(s));
}

template <typename T>
int GemmDev(const Caller& who, char transa, char transb, int m, int n, int k,
            T alpha, const T* a, int lda, const T* b, int ldb, T beta, T* c,
            int ldc, hipStream_t s) {
  Op oa, ob;
  const blasint info = GemmInfo(transa, transb, m, n, k, lda, ldb, ldc, &oa, &ob);
  if (info != 0) {
    Reject(who, info);
    return info;
  }
  if (m == 0 || n == 0) return 0;
  if ((k == 0 || alpha == T(0)) && beta == T(1)) return 0;
  if (k == 0 || alpha == T(0)) {
    ScaleDevice<T>(m, n, beta, c, ldc, s);
    return 0;
  }
  GemmByMode<T>(oa == Op::kTrans, ob == Op::kTrans, m, n, k, alpha, a, lda, b, ldb, beta, c,
                ldc, s);
  return 0;
}

template <typename T>
void TrsmHost(const Caller& who, const char* sidep, const char* uplop,
              const char* transp, const char* diagp, const blasint* mp,
              const blasint* np, const T* alphap, const T* a,
              const blasint* ldap, T* b, const blasint* ldbp) {
  const blasint m = *mp, n = *np, lda = *ldap, ldb = *ldbp;
  bool left, lower, unit;
  Op op;
  const blasint info = TrsmInfo(*sidep, *uplop, *transp, *diagp, m, n, lda, ldb,
                                &left, &lower, &op, &unit);
  if (info != 0) {
    Reject(who, info);
    return;
  }
  const T alpha = *alphap;
  if (m == 0 || n == 0) return;

  ThreadContext& ctx = Context();
  hipStream_t s = ctx.stream;
  const bool b_dev = IsDevicePointer(b);
  T* db = b;
  int64_t ldb_d = ldb;
  if (!b_dev) {
    db = static_cast<T*>(Staging(kSlotB, size_t(m) * n * sizeof(T)));
    ldb_d = m;
    if (alpha != T(0)) CopyIn(db, b, m, n, ldb, s);
  }
  // B := alpha B first; alpha == 0 leaves zeros and never reads A
  // (src/trsm.cc:172-175).
  if (alpha != T(1)) ScaleDevice<T>(m, n, alpha, db, ldb_d, s);
  if (alpha != T(0)) {
    const int64_t ka = left ? m : n;
    const T* da = a;
    int64_t lda_d = lda;
    if (!IsDevicePointer(a)) {
      T* buf = static_cast<T*>(Staging(kSlotA, size_t(ka) * ka * sizeof(T)));
      CopyIn(buf, a, ka, ka, lda, s);
      da = buf;
      lda_d = ka;
    }
    TrsmDevice<T>(left, lower, op == Op::kTrans, unit, m, n, da, lda_d, db, ldb_d, s);
  }
  if (!b_dev) CopyOut(b, ldb, db, m, n, s);
  HWY_HIP_CHECK(hipStreamSynchronize(s));
}

template <typename T>
int TrsmDev(const Caller& who, char side, char uplo, char transa, char diag,
            int m, int n, T alpha, const T* a, int lda, T* b, int ldb,
            hipStream_t s) {
  bool left, lower, unit;
  Op op;
  const blasint info =
      TrsmInfo(side, uplo, transa, diag, m, n, lda, ldb, &left, &lower, &op, &unit);
  if (info != 0) {
    Reject(who, info);
    return info;
  }
  if (m == 0 || n == 0) return 0;
  if (alpha != T(1)) ScaleDevice<T>(m, n, alpha, b, ldb, s);
  if (alpha == T(0)) return 0;
  TrsmDevice<T>(left, lower, op == Op::kTrans, unit, m, n, a, lda, b, ldb, s);
  return 0;
}

}  // namespace

#define HWY_DEFINE_CORES(T)                                                                   \
  void Gemm(const Caller& who, const char* transa, const char* transb, const blasint* m,      \
            const blasint* n, const blasint* k, const T* alpha, const T* a,                   \
            const blasint* lda, const T* b, const blasint* ldb, const T* beta, T* c,          \
            const blasint* ldc) {                                                             \
    GemmHost(who, transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);              \
  }                                                                                           \
  void Trsm(const Caller& who, const char* side, const char* uplo, const char* transa,        \
            const char* diag, const blasint* m, const blasint* n, const T* alpha, const T* a, \
            const blasint* lda, T* b, const blasint* ldb) {                                   \
    TrsmHost(who, side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb);                     \
  }
HWY_DEFINE_CORES(float)
HWY_DEFINE_CORES(double)
#undef HWY_DEFINE_CORES

}  // namespace hwy_amd

using hwy_amd::blasint;

extern "C" {

HWYBLAS_AMD_API int hwyblas_set_dgemm_mode(int mode) {
  if (mode != 0 && mode != 1) return -1;
  return hwy_amd::DgemmModeFlag().exchange(mode, std::memory_order_relaxed);
}

HWYBLAS_AMD_API int hwyblas_get_dgemm_mode(void) { return hwy_amd::DgemmMode(); }

HWYBLAS_AMD_API void sgemm_(const char* transa, const char* transb,
                            const blasint* m, const blasint* n, const blasint* k,
                            const float* alpha, const float* a, const blasint* lda,
                            const float* b, const blasint* ldb, const float* beta,
                            float* c, const blasint* ldc) {
  hwy_amd::Gemm("SGEMM", transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

HWYBLAS_AMD_API void dgemm_(const char* transa, const char* transb,
                            const blasint* m, const blasint* n, const blasint* k,
                            const double* alpha, const double* a, const blasint* lda,
                            const double* b, const blasint* ldb, const double* beta,
                            double* c, const blasint* ldc) {
  hwy_amd::Gemm("DGEMM", transa, transb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc);
}

HWYBLAS_AMD_API void strsm_(const char* side, const char* uplo, const char* transa,
                            const char* diag, const blasint* m, const blasint* n,
                            const float* alpha, const float* a, const blasint* lda,
                            float* b, const blasint* ldb) {
  hwy_amd::Trsm("STRSM", side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb);
}

HWYBLAS_AMD_API void dtrsm_(const char* side, const char* uplo, const char* transa,
                            const char* diag, const blasint* m, const blasint* n,
                            const double* alpha, const double* a, const blasint* lda,
                            double* b, const blasint* ldb) {
  hwy_amd::Trsm("DTRSM", side, uplo, transa, diag, m, n, alpha, a, lda, b, ldb);
}

HWYBLAS_AMD_API int hwyblas_dgemm_dev(char transa, char transb, int m, int n, int k,
                                      double alpha, const double* a, int lda,
                                      const double* b, int ldb, double beta,
                                      double* c, int ldc, void* stream) {
  return hwy_amd::GemmDev("DGEMM", transa, transb, m, n, k, alpha, a, lda, b, ldb,
                          beta, c, ldc, static_cast<hipStream_t>(stream));
}

HWYBLAS_AMD_API int hwyblas_sgemm_dev(char transa, char transb, int m, int n, int k,
                                      float alpha, const float* a, int lda,
                                      const float* b, int ldb, float beta,
                                      float* c, int ldc, void* stream) {
  return hwy_amd::GemmDev("SGEMM", transa, transb, m, n, k, alpha, a, lda, b, ldb,
                          beta, c, ldc, static_cast<hipStream_t>(stream));
}

HWYBLAS_AMD_API int hwyblas_dtrsm_dev(char side, char uplo, char transa, char diag,
                                      int m, int n, double alpha, const double* a,
                                      int lda, double* b, int ldb, void* stream) {
  return hwy_amd::TrsmDev("DTRSM", side, uplo, transa, diag, m, n, alpha, a, lda, b,
                          ldb, static_cast<hipStream_t>(stream));
}

HWYBLAS_AMD_API int hwyblas_strsm_dev(char side, char uplo, char transa, char diag,
                                      int m, int n, float alpha, const float* a,
                                      int lda, float* b, int ldb, void* stream) {
  return hwy_amd::TrsmDev("STRSM", side, uplo, transa, diag, m, n, alpha, a, lda, b,
                          ldb, static_cast<hipStream_t>(stream));
}

}  // extern "C"
