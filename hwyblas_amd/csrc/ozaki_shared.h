// Pieces shared by the two emulated DGEMMs (INT8 Ozaki scheme II in
// ozaki.hip, BF16 slices in ozaki_bf16.hip): the operand statistics kernels,
// the XCD-aware tile order, the grow-only workspace and the per-phase timer
// (the direct-to-LDS loads and their waits are in dma.h).
#ifndef HWYBLAS_AMD_CSRC_OZAKI_SHARED_H_
#define HWYBLAS_AMD_CSRC_OZAKI_SHARED_H_

#include <algorithm>
#include <cmath>

#include "common.h"
#include "dma.h"

namespace hwy_amd {
namespace ozaki {

// ---------------------------------------------------------------------------
// Host side, defined once (ozaki.hip).

// The calling thread's emulated-DGEMM buffer on the current device (its
// kSlotEmulated staging buffer): grow-only; a larger request frees the old
// buffer (contents lost) after synchronizing `stream`, and sets *grown when
// given.
void* GetWorkspace(size_t bytes, hipStream_t stream, bool* grown = nullptr);

// Optional per-phase timing (hwyblas_amd_profile_*): HIP events recorded on
// the call's stream around the phases of the calling thread's last emulated
// call. Mark(kNumPhases) closes the record.
enum Phase { kPhaseStats = 0, kPhaseSlice = 1, kPhaseModGemm = 2, kPhaseReconstruct = 3,
             kNumPhases = 4 };
void Mark(int phase_boundary, hipStream_t stream);
void ResetPhases();

inline size_t AlignUp(size_t x, size_t a) { return (x + a - 1) / a * a; }

constexpr uint64_t kInfBits = 0x7FF0000000000000ull;

// Read by every kernel after the decision. count 0: native FP64 runs.
struct Decision {
  int count;   // moduli (INT8) or slices (BF16)
  int ba, bb;  // INT8: fixed-point bits of the rows of op(A), columns of op(B)
  int pad;
};

// The control block of one call.
struct EmuCtl {
  Decision dec;
  uint64_t amax_bits, bmax_bits;  // max |a|, max |b| as bits (NaN/Inf: >= kInfBits)
  uint64_t xmax_bits;             // X as the bits of a non-negative double
  int gmin;                       // INT_MAX when no pair counts
  int pad;
};

constexpr int kMaxCount = 24;  // table size of DecideArgs::bits

struct DecideArgs {
  EmuCtl* ctl;
  int kind;    // 0 INT8 moduli, 1 BF16 slices
  int forced;  // > 0: use this count (no guarantee)
  int64_t k;
  int bits[kMaxCount + 1];  // INT8: max(TotalBits(N, k), 2) for N = 0..24
};

// Operand views: op(A)(i, p) = a[i * a_rs + p * a_cs], op(B)(p, j) =
// b[p * b_rs + j * b_cs].
struct EmuOperands {
  const double* a;
  int64_t a_rs, a_cs;
  const double* b;
  int64_t b_rs, b_cs;
  int64_t m, n, k;
};

// Device buffers of the statistics: the control block; row maxima of
// op(A) then column maxima of op(B) (bits), followed directly by the exact
// sums rs then cs; the scale exponents of the rows and columns.
struct EmuStats {
  EmuCtl* ctl;
  uint64_t* rowmax;
  uint64_t* sums;
  int* sa;
  int* sb;
};
size_t StatsBytes(int64_t m, int64_t n);
EmuStats CarveStats(void* base, int64_t m, int64_t n);

// Statistics of op(A) and op(B), scale exponents bits - E (scale_bits: 0
// for INT8, -1 for BF16), and the decision of kind 0 (INT8) or 1 (BF16),
// forced when forced > 0. Otherwise the coarse planes (MPc x KPc and
// NPc x KPc bytes; MPc, NPc multiples of 256, KPc of 128) and their INT8
// product feed it. Everything stays on the device.
void Decide(int kind, int forced, const EmuOperands& op, const EmuStats& st, int scale_bits,
            uint8_t* coarse_a, uint8_t* coarse_b, int64_t MPc, int64_t NPc, int64_t KPc,
            hipStream_t stream);
// The decided count (0: native), after synchronizing the stream.
int ReportCount(const EmuStats& st, hipStream_t stream);

namespace {

typedef int i32x4 __attribute__((ext_vector_type(4)));

// ---------------------------------------------------------------------------
// Operand statistics

__global__ void FillU64(uint64_t* x, int64_t n, uint64_t v) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    x[i] = v;
}

// Max |x(r, p)| per row r of a strided operand (element at x[r*rs + p*cs]),
// as the bit pattern of a non-negative double (which orders like the
// value; NaN orders above Inf). Rows contiguous (rs == 1): one thread per
// row over a chunk of p. p contiguous (cs == 1): one wave per row.
template <bool ROWS_CONTIG>
__global__ void RowAbsMaxKernel(const double* x, int64_t ld, int64_t R, int64_t K,
                                int64_t pchunk, uint64_t* out) {
  if (ROWS_CONTIG) {
    const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= R) return;
    const int64_t p0 = int64_t(blockIdx.y) * pchunk;
    const int64_t p1 = min(K, p0 + pchunk);
    uint64_t mx = 0;
    for (int64_t p = p0; p < p1; ++p) {
      const uint64_t b = __double_as_longlong(fabs(x[r + p * ld]));
      mx = b > mx ? b : mx;
    }
    if (p1 > p0) atomicMax(reinterpret_cast<unsigned long long*>(out + r), mx);
  } else {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int64_t r = int64_t(blockIdx.x) * (blockDim.x >> 6) + wave;
    if (r >= R) return;
    const int64_t p0 = int64_t(blockIdx.y) * pchunk;
    const int64_t p1 = min(K, p0 + pchunk);
    uint64_t mx = 0;
    for (int64_t p = p0 + lane; p < p1; p += 64) {
      const uint64_t b = __double_as_longlong(fabs(x[r * ld + p]));
      mx = b > mx ? b : mx;
    }
    for (int off = 32; off > 0; off >>= 1) {
      const uint64_t o = __shfl_xor(mx, off, 64);
      mx = o > mx ? o : mx;
    }
    if (lane == 0 && p1 > p0) atomicMax(reinterpret_cast<unsigned long long*>(out + r), mx);
  }
}

// Per-row scale exponents s_r = bits - E_r, rowmax_r = f 2^E_r, f in
// [0.5, 1) (0 for zero rows), and the global max bit pattern in *gmax.
__global__ void ScaleExpKernel(const uint64_t* rowmax, int64_t R, int bits, int* s,
                               uint64_t* gmax) {
  const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  uint64_t b = 0;
  if (r < R) {
    b = rowmax[r];
    const double v = __longlong_as_double(b);
    int e = 0;
    if (v > 0 && b < 0x7FF0000000000000ull) frexp(v, &e);
    s[r] = (v > 0) ? bits - e : 0;
  }
  for (int off = 32; off > 0; off >>= 1) {
    const uint64_t o = __shfl_xor(b, off, 64);
    b = o > b ? o : b;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(reinterpret_cast<unsigned long long*>(gmax), b);
}

// ---------------------------------------------------------------------------
// The count of moduli (INT8) or slices (BF16), chosen on the device.
//
// Every pair (i, j) is covered, not a sample. Two exact statistics bound
// the error of the emulated C(i, j) from above and |A||B|(i, j) from below:
//  - rs_i = sum_p ceil(|a_ip| 2^(40 - E_i)) (max_p |a_ip| = f 2^E_i, f in
//    [0.5, 1)), an upper bound of 2^40 sum_p |a_ip| / 2^E_i in exact integer
//    arithmetic (order-free atomics); cs_j likewise for the columns of op(B);
//  - G_ij = sum_p q_ip r_pj with q_ip = floor(|a_ip| 2^(7 - E_i)) and
//    r_pj = floor(|b_pj| 2^(7 - F_j)), integers in [0, 127]: one INT8 GEMM,
//    exact in INT32 for k < 2^17, and
//    |A||B|(i, j) >= 2^(E_i + F_j - 14) G_ij.
// Its epilogue keeps X = max over pairs of (cs_j + rs_i) 2^-27 / G_ij (Inf
// when G_ij = 0; pairs whose row of op(A) or column of op(B) is zero are
// exact and skipped) and gmin = min G_ij over the same pairs. In units of
// |A||B|(i, j) the error bounds are then
//  - INT8 with bA + bB bits: 2^-bA X + 2^-(bA + bB) 4096 k / gmin (the
//    rounding of each scaled entry to an integer, its cross term);
//  - BF16 with s slices: 2^(1 - 8s) X + (1.004 (s - 1) 2^-8s + 2^-16s) 4 *
//    4096 k / gmin (the truncated digits, the digit pairs left out) plus the
//    FP64 roundings of the k/1024 chunk sums of s (s + 1) / 2 pairs.
// DecideKernel takes the smallest count whose bound is within 2 sqrt(k) u
// (native FP64's typical error, u = 2^-53), else the smallest whose bound
// (every term at its worst) keeps hwyBLAS's componentwise bound
// 4 (k + 2) u (|alpha||A||B| + |beta||C|) after the reconstruction's own
// roundings, else 0: the native FP64 GEMM runs instead. Operands holding
// NaN or Inf also give 0.

// Zeroes n words at z and initialises the control block.
__global__ void InitStatsKernel(uint64_t* z, int64_t n, EmuCtl* ctl) {
  for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
       i += int64_t(gridDim.x) * blockDim.x)
    z[i] = 0;
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    ctl->dec = Decision{0, 0, 0, 0};
    ctl->amax_bits = ctl->bmax_bits = ctl->xmax_bits = 0;
    ctl->gmin = 0x7FFFFFFF;
    ctl->pad = 0;
  }
}

// Coarse magnitudes of one operand, with the exact upper sums: row r (scale
// exponent s[r], bias such that the coarse value of x is
// floor(|x| 2^(bias + s[r])) in [0, 127]) into the byte plane Y (row stride
// KP, zero padding), and sum_p ceil(|x| 2^(bias + 33 + s[r])) added to
// sums[r]. A zero row stays zero everywhere.
__device__ __forceinline__ void CoarseEight(const double (&x)[8], int sc, uint8_t* dst,
                                            uint64_t* part) {
  uint64_t word = 0, sum = 0;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const double ax = fabs(x[q]);
    const double v = ldexp(ax, sc);
    // NaN or out-of-range values (operands holding NaN or Inf, which the
    // decision sends to the native kernel) become 0.
    const int c = (v >= 0.0 && v < 128.0) ? int(v) : 0;
    word |= uint64_t(uint8_t(c)) << (8 * q);
    const double w = ldexp(ax, sc + 33);
    uint64_t t = (w >= 0.0 && w <= 0x1p41) ? uint64_t(ceil(w)) : 0;
    if (ax > 0.0 && t == 0) t = 1;
    sum += t;
  }
  *reinterpret_cast<uint64_t*>(dst) = word;
  *part = sum;
}

// x(r, p) = x[r * ld + p]: 16 rows x 128 p per block, 16 threads per row.
__global__ void __launch_bounds__(256)
    CoarseRowsKernel(const double* x, int64_t ld, int64_t R, int64_t K, int64_t RP, int64_t KP,
                     const int* s, int bias, uint8_t* Y, uint64_t* sums) {
  const int tid = threadIdx.x;
  const int64_t r = int64_t(blockIdx.y) * 16 + (tid >> 4);
  const int64_t p0 = int64_t(blockIdx.x) * 128 + (tid & 15) * 8;
  if (r >= RP) return;  // uniform per 16-lane row group
  double v[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const bool ok = r < R && p0 + q < K;
    const double t = x[ok ? r * ld + p0 + q : 0];
    v[q] = ok ? t : 0.0;
  }
  uint64_t part;
  CoarseEight(v, r < R ? bias + s[r] : 0, Y + r * KP + p0, &part);
#pragma unroll
  for (int off = 8; off > 0; off >>= 1) part += __shfl_xor(part, off, 16);
  if ((tid & 15) == 0 && r < R && part != 0)
    atomicAdd(reinterpret_cast<unsigned long long*>(sums + r), (unsigned long long)part);
}

// x(r, p) = x[r + p * ld]: a 32 x 128 tile transposed through LDS.
__global__ void __launch_bounds__(256)
    CoarseColsKernel(const double* x, int64_t ld, int64_t R, int64_t K, int64_t RP, int64_t KP,
                     const int* s, int bias, uint8_t* Y, uint64_t* sums) {
  constexpr int TR = 32, TP = 128, PER = TR * TP / 256;
  __shared__ double T[TP][TR + 1];
  const int tid = threadIdx.x;
  const int64_t r0 = int64_t(blockIdx.y) * TR, p0 = int64_t(blockIdx.x) * TP;
  double tv[PER];
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int idx = tid + q * 256, rr = idx % TR, pp = idx / TR;
    const int64_t r = r0 + rr, p = p0 + pp;
    const bool ok = r < R && p < K;
    const double t = x[ok ? r + p * ld : 0];
    tv[q] = ok ? t : 0.0;
  }
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int idx = tid + q * 256;
    T[idx / TR][idx % TR] = tv[q];
  }
  __syncthreads();
  const int pq = (tid & 15) * 8;
  for (int rr = tid >> 4; rr < TR; rr += 16) {
    const int64_t r = r0 + rr;
    if (r >= RP) break;  // uniform per 16-lane row group
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = T[pq + q][rr];
    uint64_t part;
    CoarseEight(v, r < R ? bias + s[r] : 0, Y + r * KP + p0 + pq, &part);
#pragma unroll
    for (int off = 8; off > 0; off >>= 1) part += __shfl_xor(part, off, 16);
    if ((tid & 15) == 0 && r < R && part != 0)
      atomicAdd(reinterpret_cast<unsigned long long*>(sums + r), (unsigned long long)part);
  }
}

// The error bound of a count, in units of |A||B|(i, j), and the two
// targets; restated in the oracle (orc_emulated_choose).
__device__ __forceinline__ void BoundOf(const DecideArgs& a, int count, double x, double z,
                                        double* typical, double* worst) {
  const double u = 0x1p-53;
  if (a.kind == 0) {
    const int bits = a.bits[count], ba = bits / 2;
    const double t = ldexp(x, -ba) + ldexp(z, -bits);
    *typical = t;
    *worst = t * (1.0 + 0x1p-20);
  } else {
    const double s = count;
    const double trunc = ldexp(x, 1 - 8 * count);
    const double drop = (1.004 * (s - 1.0) * ldexp(1.0, -8 * count) + ldexp(1.0, -16 * count)) *
                        4.0 * z;
    const double chunks = double((a.k + 1023) / 1024) * (s * (s + 1.0) * 0.5);
    *typical = trunc + drop + sqrt(chunks) * u;
    *worst = (trunc + drop) * (1.0 + 0x1p-20) + (1.02 * chunks + 3.0) * u;
  }
}

__global__ void DecideKernel(DecideArgs a) {
  if (threadIdx.x != 0) return;
  EmuCtl* ctl = a.ctl;
  Decision d{0, 0, 0, 0};
  const bool finite = ctl->amax_bits < kInfBits && ctl->bmax_bits < kInfBits;
  const int lo = a.kind == 0 ? 2 : 1, hi = a.kind == 0 ? 24 : 16;
  if (finite && a.forced > 0) {
    d.count = min(max(a.forced, lo), hi);
  } else if (finite) {
    const double u = 0x1p-53, k = double(a.k);
    const double x = __longlong_as_double(ctl->xmax_bits);
    const double z = ctl->gmin == 0x7FFFFFFF ? 0.0 : 4096.0 * k / double(ctl->gmin);
    const double t_typ = 2.0 * sqrt(k) * u;
    const double t_bound = (4.0 * (k + 2.0) - 24.0) * u;
    for (int c = lo; c <= hi && d.count == 0; ++c) {
      double typ, worst;
      BoundOf(a, c, x, z, &typ, &worst);
      if (typ <= t_typ && worst <= t_bound) d.count = c;
    }
    for (int c = lo; c <= hi && d.count == 0; ++c) {
      double typ, worst;
      BoundOf(a, c, x, z, &typ, &worst);
      if (worst <= t_bound) d.count = c;
    }
  }
  if (a.kind == 0 && d.count > 0) {
    const int bits = a.bits[d.count];
    d.ba = bits / 2;
    d.bb = bits - d.ba;
  }
  ctl->dec = d;
}

// XCD-aware order of output tiles: consecutive workgroups go to the eight
// XCDs in turn, so workgroup b is given the tile of "local" index lid that
// keeps each XCD on its own band of 8 tile rows (column-major within the
// band), where its L2 holds the shared A and B panels.
__device__ __forceinline__ void XcdTile(int tiles_m, int tiles_n, int* tm, int* tn) {
  const int nblk = gridDim.x, bid = blockIdx.x;
  const int xcd = bid % 8, idx = bid / 8, per = nblk / 8, rem = nblk % 8;
  const int lid = xcd < rem ? xcd * (per + 1) + idx : rem * (per + 1) + (xcd - rem) * per + idx;
  const int group = 8 * tiles_n;
  const int first_m = (lid / group) * 8;
  const int gm = min(tiles_m - first_m, 8);
  *tm = first_m + (lid % group) % gm;
  *tn = (lid % group) / gm;
}

}  // namespace
}  // namespace ozaki
}  // namespace hwy_amd

#endif  // HWYBLAS_AMD_CSRC_OZAKI_SHARED_H_
