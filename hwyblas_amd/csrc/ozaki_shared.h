// Pieces shared by the two emulated DGEMMs (INT8 Ozaki scheme II in
// ozaki.hip, BF16 slices in ozaki_bf16.hip): the operand statistics kernels,
// the direct-to-LDS loads and their waits, the XCD-aware tile order, the
// grow-only workspace and the per-phase timer.
#ifndef HWYBLAS_AMD_CSRC_OZAKI_SHARED_H_
#define HWYBLAS_AMD_CSRC_OZAKI_SHARED_H_

#include <algorithm>
#include <cmath>

#include "common.h"

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

// Sampled statistics for the choice of moduli: for sample rows i of op(A)
// (blockIdx.x) and sample columns j of op(B): sum_p |a_ip| |b_pj|,
// sum_p |a_ip|, sum_p |b_pj|, max_p |a_ip| and max_p |b_pj|.
constexpr int kSamples = 32;
struct SampleOut {
  double sab[kSamples][kSamples];
  double sa[kSamples], sb[kSamples], ma[kSamples], mb[kSamples];
  uint64_t amax_bits, bmax_bits;
};

__global__ void __launch_bounds__(256)
    SampleKernel(const double* a, int64_t a_rs, int64_t a_cs, const double* b, int64_t b_rs,
                 int64_t b_cs, int64_t m, int64_t n, int64_t k, SampleOut* out) {
  const int si = blockIdx.x;
  const int64_t i = (m - 1) * si / (kSamples - 1 > 0 ? kSamples - 1 : 1);
  __shared__ double red[256];
  double s_ab[kSamples] = {};
  double s_a = 0, m_a = 0;
  double s_b[kSamples] = {}, m_b[kSamples] = {};
  for (int64_t p = threadIdx.x; p < k; p += blockDim.x) {
    const double av = fabs(a[i * a_rs + p * a_cs]);
    s_a += av;
    m_a = fmax(m_a, av);
#pragma unroll
    for (int sj = 0; sj < kSamples; ++sj) {
      const int64_t j = (n - 1) * sj / (kSamples - 1);
      const double bv = fabs(b[p * b_rs + j * b_cs]);
      s_ab[sj] += av * bv;
      if (si == 0) {
        s_b[sj] += bv;
        m_b[sj] = fmax(m_b[sj], bv);
      }
    }
  }
  auto reduce = [&](double v, bool is_max) {
    red[threadIdx.x] = v;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if (threadIdx.x < s)
        red[threadIdx.x] = is_max ? fmax(red[threadIdx.x], red[threadIdx.x + s])
                                  : red[threadIdx.x] + red[threadIdx.x + s];
      __syncthreads();
    }
    const double r = red[0];
    __syncthreads();
    return r;
  };
  for (int sj = 0; sj < kSamples; ++sj) {
    const double v = reduce(s_ab[sj], false);
    if (threadIdx.x == 0) out->sab[si][sj] = v;
  }
  const double sa = reduce(s_a, false), ma = reduce(m_a, true);
  if (threadIdx.x == 0) {
    out->sa[si] = sa;
    out->ma[si] = ma;
  }
  if (si == 0) {
    for (int sj = 0; sj < kSamples; ++sj) {
      const double sb = reduce(s_b[sj], false), mb = reduce(m_b[sj], true);
      if (threadIdx.x == 0) {
        out->sb[sj] = sb;
        out->mb[sj] = mb;
      }
    }
  }
}

// Worst ratios over the sampled (i, j): maxrow|a_i| sum_p |b_pj| / S_ij,
// maxcol|b_j| sum_p |a_ip| / S_ij and k maxrow|a_i| maxcol|b_j| / S_ij, with
// S_ij = sum_p |a_ip| |b_pj| (pairs with S_ij = 0 are skipped).
inline void WorstRatios(const SampleOut& s, int64_t k, double* wa, double* wb, double* wab) {
  *wa = *wb = *wab = 0;
  for (int i = 0; i < kSamples; ++i)
    for (int j = 0; j < kSamples; ++j) {
      const double denom = s.sab[i][j];
      if (denom <= 0) continue;
      *wa = std::max(*wa, s.ma[i] * s.sb[j] / denom);
      *wb = std::max(*wb, s.mb[j] * s.sa[i] / denom);
      *wab = std::max(*wab, double(k) * s.ma[i] * s.mb[j] / denom);
    }
}

// ---------------------------------------------------------------------------
// Direct loads into LDS and their waits

template <int N>
__device__ __forceinline__ void WaitVmcntImm() {
  // s_waitcnt with only vmcnt constrained (expcnt 7, lgkmcnt 15 = no wait);
  // vmcnt is 6 bits, split over [3:0] and [15:14].
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 0xF) | (7 << 4) | (0xF << 8) | ((N >> 4) << 14));
}

// Waits until at most `slices` later slices (PER_SLICE DMAs each) are
// outstanding.
template <int PER_SLICE>
__device__ __forceinline__ void WaitSlices(int slices) {
  switch (slices) {
    case 0: WaitVmcntImm<0>(); break;
    case 1: WaitVmcntImm<PER_SLICE>(); break;
    case 2: WaitVmcntImm<2 * PER_SLICE>(); break;
    default: WaitVmcntImm<3 * PER_SLICE>(); break;
  }
}

// One global_load_lds_dwordx4: this lane's 16 bytes at `src` land in LDS at
// `lds_wave_base` + 16 * lane. Issued as inline assembly so that the
// compiler, which would otherwise wait for every outstanding DMA before any
// LDS read, leaves the waiting to WaitVmcnt. M0 (the LDS base) must be
// written one instruction before the DMA reads it.
__device__ __forceinline__ void DmaToLds(const void* src, const void* lds_wave_base) {
  const unsigned base =
      __builtin_amdgcn_readfirstlane(unsigned(reinterpret_cast<uintptr_t>(lds_wave_base)));
  unsigned saved;
  asm volatile(
      "s_mov_b32 %0, m0\n\ts_mov_b32 m0, %1\n\ts_nop 0\n\t"
      "global_load_lds_dwordx4 %2, off\n\ts_mov_b32 m0, %0"
      : "=&s"(saved)
      : "s"(base), "v"(src)
      : "memory");
}

__device__ __forceinline__ void BarrierNoFence() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
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
