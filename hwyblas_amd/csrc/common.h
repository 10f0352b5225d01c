// Internal declarations shared by the host (C++) and device (HIP) parts of
// libhwyblas_amd: argument decoding, error reporting, the per-thread HIP
// context and the device-side launchers.
#ifndef HWYBLAS_AMD_CSRC_COMMON_H_
#define HWYBLAS_AMD_CSRC_COMMON_H_

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include "../../include/hwyblas_amd.h"

namespace hwy_amd {

using blasint = hwyblas_int;

// ---------------------------------------------------------------------------
// CHARACTER*1 decoding: only the first character is read, case-insensitively
// (hwyBLAS src/common.h:68-92, src/api.h:72-112). Each decoder returns false
// for an invalid value and still sets its output to the default.

inline char ToUpper(char c) { return (c >= 'a' && c <= 'z') ? char(c - 32) : c; }

enum class Op { kNoTrans, kTrans };  // ConjTrans is Trans for real types

inline bool DecodeOp(char c, Op* op) {
  switch (ToUpper(c)) {
    case 'N': *op = Op::kNoTrans; return true;
    case 'T':
    case 'C': *op = Op::kTrans; return true;
    default: *op = Op::kNoTrans; return false;
  }
}
inline bool DecodeUplo(char c, bool* lower) {
  switch (ToUpper(c)) {
    case 'L': *lower = true; return true;
    case 'U': *lower = false; return true;
    default: *lower = false; return false;
  }
}
inline bool DecodeSide(char c, bool* left) {
  switch (ToUpper(c)) {
    case 'L': *left = true; return true;
    case 'R': *left = false; return true;
    default: *left = false; return false;
  }
}
inline bool DecodeDiag(char c, bool* unit) {
  switch (ToUpper(c)) {
    case 'U': *unit = true; return true;
    case 'N': *unit = false; return true;
    default: *unit = false; return false;
  }
}

// ---------------------------------------------------------------------------
// Invalid-argument reports (errors.cpp)

// Whom a routine reports its first invalid argument to. A Fortran caller is
// identified by the routine name, and positions are Fortran argument
// numbers. A CBLAS caller is identified by the CBLAS name, and Fortran
// argument i was made from CBLAS argument cblas_arg[i - 1] (the layout is
// CBLAS argument 1), so the position reported is the one the program passed.
struct Caller {
  const char* name;
  const unsigned char* cblas_arg = nullptr;
  // Implicit: a Fortran routine name is a complete identity.
  Caller(const char* fortran_name) : name(fortran_name) {}  // NOLINT
  Caller(const char* cblas_name, const unsigned char* from_cblas)
      : name(cblas_name), cblas_arg(from_cblas) {}
  // The position reported for Fortran argument `info` (1-based).
  blasint Position(blasint info) const;
};

// Reports invalid argument `info` (1-based, in the Fortran numbering of the
// routine): to the installed handler if any, else printed by cblas_xerbla
// (CBLAS callers) or xerbla_ (Fortran callers). Never aborts.
void Reject(const Caller& who, blasint info);

// Reports a CBLAS layout that is neither row- nor column-major (argument 1).
void RejectLayout(const char* name, int layout);

// Reports a routine's internal failure `code` (negative, not an argument
// position) to the installed handler, if any, as (name, code); the routine
// has already printed what failed.
void ReportInternal(const char* name, int code);

// Prints the HIP error and aborts: device failures have no BLAS error code,
// and a routine that cannot run must not return as if it had.
[[noreturn]] void FatalHip(hipError_t e, const char* what, const char* file,
                           int line);

#define HWY_HIP_CHECK(expr)                                          \
  do {                                                               \
    hipError_t hwy_e_ = (expr);                                      \
    if (hwy_e_ != hipSuccess) ::hwy_amd::FatalHip(hwy_e_, #expr, __FILE__, __LINE__); \
  } while (0)

// ---------------------------------------------------------------------------
// Per-thread context (runtime.cpp): a stream of default priority and
// grow-only device staging buffers, reused across calls like hwyBLAS's
// thread-local scratch slots (src/common.cc:46-64).

// kSlotD: a dense copy of a structured operand (SYMM, TRMM). kSlotPackA,
// kSlotPackB, kSlotProd: the complex GEMM's real packs and real product.
// kSlotEmulated: the emulated DGEMM's statistics, slices and residues.
// kSlotTune: the autotuner's operand memory. kSlotSplit: the split-k
// partials of the GEMM entries (SplitPartials).
enum StagingSlot : int {
  kSlotA = 0,
  kSlotB = 1,
  kSlotC = 2,
  kSlotD = 3,
  kSlotPackA = 4,
  kSlotPackB = 5,
  kSlotProd = 6,
  kSlotEmulated = 7,
  kSlotTune = 8,
  kSlotSplit = 9,
  kNumSlots = 10
};

struct ThreadContext {
  hipStream_t stream = nullptr;
  // Streams of the pipelined host-operand GEMM's copies (created on first use).
  hipStream_t copy_in = nullptr, copy_out = nullptr;
  int device = -1;
  void* buf[kNumSlots] = {};
  size_t cap[kNumSlots] = {};
  // The last work that used kSlotSplit: its event and stream.
  hipEvent_t split_done = nullptr;
  hipStream_t split_stream = nullptr;
};

ThreadContext& Context();
// The calling thread's staging buffer `slot` on the current device, grown
// to at least `bytes`. Growing waits for the thread's stream and, if given,
// `user_stream` (work on either may still read the old buffer); `grown`
// (optional) tells whether the buffer was reallocated (its contents lost).
void* Staging(int slot, size_t bytes, hipStream_t user_stream = nullptr,
              bool* grown = nullptr);
// The calling thread's split-k partials (kSlotSplit) for work about to be
// queued on `stream`: `stream` first waits for the last work that used them
// on another stream, so calls of one thread on different streams never
// share them at once. SplitPartialsDone(stream) marks the end of that work.
void* SplitPartials(size_t bytes, hipStream_t stream);
void SplitPartialsDone(hipStream_t stream);

// Whether p can be dereferenced by kernels on the current device (device or
// managed memory); host memory, including pinned memory, is staged.
bool IsDevicePointer(const void* p);

// hipMemcpy2DAsync with the same arguments, except that a contiguous span
// (both pitches equal to the width, or a single row) goes as one 1-D copy.
hipError_t Memcpy2DAsync(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width,
                         size_t height, hipMemcpyKind kind, hipStream_t stream);

// ---------------------------------------------------------------------------
// Device operands and launchers (gemm.hip, trsm.hip). All take device
// pointers, enqueue work on `stream` and return without synchronising.
// Leading dimensions and offsets are 64-bit inside the kernels.

// C = alpha op(A) op(B) + beta C for m, n, k > 0 and alpha != 0; beta == 0
// never reads C. Op is encoded as trans = false (N) / true (T).
template <typename T>
void GemmDevice(bool ta, bool tb, int64_t m, int64_t n, int64_t k, T alpha,
                const T* a, int64_t lda, const T* b, int64_t ldb, T beta,
                T* c, int64_t ldc, hipStream_t stream);
template <>
void GemmDevice<double>(bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha,
                        const double* a, int64_t lda, const double* b, int64_t ldb,
                        double beta, double* c, int64_t ldc, hipStream_t stream);
template <>
void GemmDevice<float>(bool ta, bool tb, int64_t m, int64_t n, int64_t k, float alpha,
                       const float* a, int64_t lda, const float* b, int64_t ldb, float beta,
                       float* c, int64_t ldc, hipStream_t stream);

// As GemmDevice, computing and writing only part of C: 1 its lower
// triangle (row >= col), 2 its upper triangle (row <= col), 0 all of it.
template <typename T>
void GemmDevicePart(int part, bool ta, bool tb, int64_t m, int64_t n, int64_t k, T alpha,
                    const T* a, int64_t lda, const T* b, int64_t ldb, T beta, T* c,
                    int64_t ldc, hipStream_t stream);
template <>
void GemmDevicePart<double>(int part, bool ta, bool tb, int64_t m, int64_t n, int64_t k,
                            double alpha, const double* a, int64_t lda, const double* b,
                            int64_t ldb, double beta, double* c, int64_t ldc,
                            hipStream_t stream);
template <>
void GemmDevicePart<float>(int part, bool ta, bool tb, int64_t m, int64_t n, int64_t k,
                           float alpha, const float* a, int64_t lda, const float* b,
                           int64_t ldb, float beta, float* c, int64_t ldc,
                           hipStream_t stream);

// As GemmDevice, when one factor is triangular with explicit zeros: each
// tile multiplies only over the k slices that meet its nonzero k-range
// (krange as GemmParams::krange: 1 k >= the tile's first row, 2 k <= its
// last row, 3 k <= its last column, 4 k >= its first column).
template <typename T>
void GemmDeviceKRange(int krange, bool ta, bool tb, int64_t m, int64_t n, int64_t k, T alpha,
                      const T* a, int64_t lda, const T* b, int64_t ldb, T beta, T* c,
                      int64_t ldc, hipStream_t stream);
template <>
void GemmDeviceKRange<double>(int krange, bool ta, bool tb, int64_t m, int64_t n, int64_t k,
                              double alpha, const double* a, int64_t lda, const double* b,
                              int64_t ldb, double beta, double* c, int64_t ldc,
                              hipStream_t stream);
template <>
void GemmDeviceKRange<float>(int krange, bool ta, bool tb, int64_t m, int64_t n, int64_t k,
                             float alpha, const float* a, int64_t lda, const float* b,
                             int64_t ldb, float beta, float* c, int64_t ldc,
                             hipStream_t stream);

// Split k for skinny products (few 32 x 32 tiles of C, long k), FP64: the k
// range in up to SplitKCount(m, n, k) chunks of whole 32-deep slices (as
// many as `capacity` doubles of `partials` hold m x n each) computed at once
// (one launch, one grid row per chunk), then C = alpha (the chunks summed in
// ascending order) + beta C. The sum's rounding differs from a single
// chain's (it stays within the GEMM bound); deterministic. One chunk runs
// GemmDevice.
constexpr int kMaxSplitK = 64;
int SplitKCount(int64_t m, int64_t n, int64_t k);
// The GEMM entries' split k (FP64 and FP32): whether (m, n, k) is split, and
// the product on the thread's split-k partials.
template <typename T>
bool GemmSplitPays(int64_t m, int64_t n, int64_t k);
template <typename T>
void GemmSplitKPublic(bool ta, bool tb, int64_t m, int64_t n, int64_t k, T alpha, const T* a,
                      int64_t lda, const T* b, int64_t ldb, T beta, T* c, int64_t ldc,
                      hipStream_t stream);
void GemmSplitKDevice(bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha,
                      const double* a, int64_t lda, const double* b, int64_t ldb, double beta,
                      double* c, int64_t ldc, double* partials, int64_t capacity,
                      hipStream_t stream);

// As GemmDevice<double>, doing the work only if *gate == 0 when the kernel
// runs (a device-side decision made earlier on the stream; every workgroup
// exits at once otherwise).
template <typename T>
void GemmDeviceGated(const int* gate, bool ta, bool tb, int64_t m, int64_t n, int64_t k,
                     T alpha, const T* a, int64_t lda, const T* b, int64_t ldb, T beta, T* c,
                     int64_t ldc, hipStream_t stream);
template <>
void GemmDeviceGated<double>(const int* gate, bool ta, bool tb, int64_t m, int64_t n,
                             int64_t k, double alpha, const double* a, int64_t lda,
                             const double* b, int64_t ldb, double beta, double* c, int64_t ldc,
                             hipStream_t stream);

// The DGEMM mode of the drop-in entries (hwyblas_set_dgemm_mode): 0 native
// FP64, 1 the emulated DGEMM for large products.
int DgemmMode();

namespace ozaki {
// The emulated DGEMMs (ozaki.hip: INT8 moduli; ozaki_bf16.hip: BF16 slices)
// for m, n, k > 0 and alpha != 0, transpositions as GemmDevice. count > 0
// fixes the moduli (slices); 0 lets the device choose them, or the native
// kernel. Asynchronous unless `used` is non-null: it then receives the
// count that ran (0: the native kernel) after synchronizing the stream.
void EmulatedDgemm(bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha,
                   const double* a, int64_t lda, const double* b, int64_t ldb, double beta,
                   double* c, int64_t ldc, int count, hipStream_t stream, int* used);
void EmulatedDgemmBf16(bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha,
                       const double* a, int64_t lda, const double* b, int64_t ldb, double beta,
                       double* c, int64_t ldc, int count, hipStream_t stream, int* used);
}  // namespace ozaki

// Tile configuration GemmDevice uses for a shape (an index into the table of
// gemm_configs.h).
int ChooseGemmConfig(int dtype_bytes, bool ta, bool tb, int64_t m, int64_t n, int64_t k);

// C := beta C over m x n; beta == 0 writes zeros without reading C.
template <typename T>
void ScaleDevice(int64_t m, int64_t n, T beta, T* c, int64_t ldc,
                 hipStream_t stream);

// op(A) X = B (left) or X op(A) = B (right) in place, alpha already applied.
template <typename T>
void TrsmDevice(bool left, bool lower, bool trans, bool unit, int64_t m,
                int64_t n, const T* a, int64_t lda, T* b, int64_t ldb,
                hipStream_t stream);

}  // namespace hwy_amd

#endif  // HWYBLAS_AMD_CSRC_COMMON_H_
