// Emulated DGEMM: FP64 in, FP64 out, computed on the INT8 matrix cores
// (v_mfma_i32_32x32x32_i8) by the Ozaki scheme II (Ozaki, Uchino, Imamura,
// "Ozaki Scheme II: A GEMM-oriented emulation of floating-point matrix
// multiplication using an integer modular technique", 2025), restated here:
//
//  1. Scaling. Row i of op(A) is scaled by 2^sa_i and column j of op(B) by
//     2^sb_j and rounded to integers A', B' with |A'| <= 2^bA, |B'| <= 2^bB,
//     bA + bB chosen so that k 2^(bA+bB) < P / 2, P = p_1 ... p_N. The exact
//     integer product C' = A' B' is then the unique integer of (-P/2, P/2)
//     with its residues modulo every p_l.
//  2. Residues. For N pairwise coprime moduli p_l <= 256, A' and B' are
//     reduced to symmetric residues in [-128, 127] (int8 "slices").
//  3. Products. C'_l = A'_l B'_l is exact in INT32 (|C'_l| <= k 128^2 < 2^31
//     for k < 131072) and is reduced mod p_l in the GEMM epilogue to one byte
//     per element and modulus.
//  4. Reconstruction (Chinese remainder theorem). With M_l = P / p_l and
//     y_l = M_l^-1 mod p_l, C'/P = sum_l c_l y_l / p_l (mod 1). The fractions
//     y_l / p_l are held as 192-bit fixed-point numbers, so the sum mod 1 is
//     an exact wrap-around integer sum (error < N 256 2^-192 of P); its
//     centred value times P gives C' to about one rounding. Finally
//     C = alpha 2^-(sa_i + sb_j) C' + beta C with hwyBLAS's Store rules.
// The number of moduli is given by the caller, or chosen on the device from
// exact statistics of every pair (i, j) (DecideKernel, ozaki_shared.h): the
// smallest N whose error bound stays at native FP64's typical level, else
// the smallest that keeps hwyBLAS's componentwise bound, else the native
// FP64 kernel runs. Inputs with NaN or Inf, which the integer scheme cannot
// carry, also run natively, as does k >= 131072. The decision never returns
// to the host, so a call is asynchronous unless it reports the count.
#include <stdio.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <mutex>

#include "common.h"
#include "ozaki_shared.h"

namespace hwy_amd {
namespace ozaki {
namespace {

constexpr int kMaxModuli = 24;
constexpr int kLimbs = 4;     // 160-bit CRT fractions in 40-bit limbs
constexpr int kLimbBits = 40;
// Pairwise coprime, largest first: 2^8, 3*5*17, 11*23, 251, 13*19, primes,
// 7*31, primes.
constexpr int kModuli[kMaxModuli] = {256, 255, 253, 251, 247, 241, 239, 233,
                                     229, 227, 223, 217, 211, 199, 197, 193,
                                     191, 181, 179, 173, 167, 163, 157, 151};

struct Crt {
  int nmod;
  int p[kMaxModuli];
  double inv_p[kMaxModuli];
  double frac[kMaxModuli][kLimbs];  // floor(2^160 y_l / p_l) in 40-bit limbs, low first
  double P;                           // prod p_l, rounded
  double pow32[kMaxModuli];           // 2^32 mod p_l
  double pow64[kMaxModuli];           // 2^64 mod p_l
};

// ---------------------------------------------------------------------------
// CRT constants for every N, built at compile time

constexpr int64_t ModInverse(int64_t a, int64_t m) {  // a^-1 mod m, gcd(a, m) = 1
  int64_t t = 0, nt = 1, r = m, nr = a % m;
  while (nr != 0) {
    const int64_t q = r / nr;
    int64_t tmp = t - q * nt;
    t = nt;
    nt = tmp;
    tmp = r - q * nr;
    r = nr;
    nr = tmp;
  }
  return t < 0 ? t + m : t;
}

constexpr Crt MakeCrt(int nmod) {
  Crt c{};
  c.nmod = nmod;
  c.P = 1;
  for (int l = 0; l < nmod; ++l) {
    c.p[l] = kModuli[l];
    c.inv_p[l] = 1.0 / kModuli[l];
    c.P *= kModuli[l];
    const uint64_t p32 = (uint64_t(1) << 32) % uint64_t(kModuli[l]);
    c.pow32[l] = double(p32);
    c.pow64[l] = double((p32 * p32) % uint64_t(kModuli[l]));
  }
  for (int l = 0; l < nmod; ++l) {
    const int64_t p = kModuli[l];
    // M_l mod p_l, then y_l = (M_l mod p_l)^-1 mod p_l.
    int64_t ml = 1;
    for (int q = 0; q < nmod; ++q)
      if (q != l) ml = (ml * kModuli[q]) % p;
    const int64_t y = ModInverse(ml, p);
    // floor(2^160 y / p) by long division in base 2^40, most significant
    // limb first (rem < p < 2^9, so rem 2^40 fits in 64 bits).
    uint64_t rem = uint64_t(y);
    for (int t = kLimbs - 1; t >= 0; --t) {
      const uint64_t num = rem << kLimbBits;
      c.frac[l][t] = double(num / uint64_t(p));
      rem = num % uint64_t(p);
    }
  }
  return c;
}

struct CrtTable {
  Crt t[kMaxModuli + 1];  // t[N]: the first N moduli
};
constexpr CrtTable MakeCrtTable() {
  CrtTable T{};
  for (int nmod = 1; nmod <= kMaxModuli; ++nmod) T.t[nmod] = MakeCrt(nmod);
  return T;
}
__device__ const CrtTable kCrtTable = MakeCrtTable();

double Log2P(int nmod) {
  double s = 0;
  for (int l = 0; l < nmod; ++l) s += std::log2(double(kModuli[l]));
  return s;
}

// Total fixed-point bits bA + bB that N moduli allow for inner dimension k:
// k 2^(bA+bB) < P / 2.
int TotalBits(int nmod, int64_t k) {
  return int(std::floor(Log2P(nmod) - 1.0 - std::log2(double(k)) - 1e-9));
}

// ---------------------------------------------------------------------------
// Device helpers

__device__ __forceinline__ int SymMod(double x, int p, double inv_p) {
  // x an integer with |x| < 2^52: x - p rint(x / p) is exact, then centred
  // into [-p/2, p/2).
  const double q = rint(x * inv_p);
  int r = int(__builtin_fma(-q, double(p), x));
  const int half = p >> 1;
  if (r >= half) r -= p;
  if (r < -half) r += p;
  return r;
}

__device__ __forceinline__ int ModU(int64_t x, int p, double inv_p) {
  // x mod p in [0, p) for |x| < 2^52.
  const double q = floor(double(x) * inv_p);
  int64_t r = x - int64_t(q) * p;
  if (r < 0) r += p;
  if (r >= p) r -= p;
  return int(r);
}

__device__ __forceinline__ int MulMod(int a, int b, int p, double inv_p) {
  return ModU(int64_t(a) * b, p, inv_p);
}

// Symmetric residue of the integer v = x 2^s (x a double; the rounded value
// when v is not an integer).
__device__ __forceinline__ int ResidueOf(double v, int p, double inv_p) {
  if (fabs(v) < 0x1p52) return SymMod(rint(v), p, inv_p);
  // v = mi 2^E exactly, mi an integer below 2^53, E >= 0.
  int e;
  const double m = frexp(v, &e);
  const double mi = ldexp(m, 53);
  int E = e - 53;
  int r = ModU(int64_t(mi), p, inv_p);
  int base = 2 % p, pw = 1;
  while (E > 0) {
    if (E & 1) pw = MulMod(pw, base, p, inv_p);
    base = MulMod(base, base, p, inv_p);
    E >>= 1;
  }
  r = MulMod(r, pw, p, inv_p);
  if (r >= (p >> 1)) r -= p;
  return r;
}

// Slices: Y[l][r][p] (row stride KP, plane stride RP * KP) = symmetric
// residue mod p_l of round(x(r, p) 2^s_r), zero for r >= R or p >= K, with
// s_r = bits - E_r (bits = bA for op(A), side 0, bB for op(B)); nothing runs
// when the decision is the native kernel. Every
// thread produces 8 consecutive bytes of one row per modulus (16 threads
// cover a row's 128 bytes: whole cache lines).
// An integer 2^52 <= |v| < 2^96 as exact 32-bit limbs l0 + l1 2^32 + l2 2^64
// of |v|: its residue mod p is that of l0 + l1 (2^32 mod p) + l2 (2^64 mod p),
// an exact FP64 integer below 2^41, reduced once per modulus.
struct Limbs {
  double l0, l1, l2;
};
__device__ __forceinline__ Limbs SplitLimbs(double v) {
  const double a = fabs(v);
  Limbs s;
  s.l2 = floor(a * 0x1p-64);
  const double a1 = a - s.l2 * 0x1p64;
  s.l1 = floor(a1 * 0x1p-32);
  s.l0 = a1 - s.l1 * 0x1p32;
  return s;
}
__device__ __forceinline__ int LimbResidue(const Limbs& s, bool neg, int p, double inv_p,
                                           double pow32, double pow64) {
  const double t = __builtin_fma(s.l2, pow64, __builtin_fma(s.l1, pow32, s.l0));
  const double q = floor(t * inv_p);
  double rr = __builtin_fma(-q, double(p), t);
  if (rr < 0.0) rr += p;
  if (rr >= p) rr -= p;
  int r = int(rr);
  if (neg && r != 0) r = p - r;
  if (r >= (p >> 1)) r -= p;
  return r;
}

__device__ __forceinline__ void SliceEight(const double (&x)[8], int sc, const Crt& crt,
                                           int8_t* dst, int64_t plane) {
  double v[8];
  bool any_big = false;
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    v[q] = ldexp(x[q], sc);
    any_big |= !(fabs(v[q]) < 0x1p52);
  }
  if (!any_big) {
    for (int l = 0; l < crt.nmod; ++l) {
      const int p = crt.p[l];
      const double ip = crt.inv_p[l];
      uint64_t word = 0;
#pragma unroll
      for (int q = 0; q < 8; ++q) {
        const int res = (v[q] == 0.0) ? 0 : SymMod(rint(v[q]), p, ip);
        word |= uint64_t(uint8_t(int8_t(res))) << (8 * q);
      }
      *reinterpret_cast<uint64_t*>(dst + l * plane) = word;
    }
    return;
  }
  // Some |v| >= 2^52 (many moduli): 32-bit limbs for 2^52 <= |v| < 2^96, the
  // general path beyond.
  int kind[8];
  Limbs limbs[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const double a = fabs(v[q]);
    kind[q] = a < 0x1p52 ? 0 : (a < 0x1p96 ? 1 : 2);
    limbs[q] = SplitLimbs(kind[q] == 1 ? v[q] : 0.0);
  }
  for (int l = 0; l < crt.nmod; ++l) {
    const int p = crt.p[l];
    const double ip = crt.inv_p[l];
    uint64_t word = 0;
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const int res = (v[q] == 0.0) ? 0
                      : kind[q] == 0 ? SymMod(rint(v[q]), p, ip)
                      : kind[q] == 1 ? LimbResidue(limbs[q], v[q] < 0.0, p, ip, crt.pow32[l],
                                                   crt.pow64[l])
                                     : ResidueOf(v[q], p, ip);
      word |= uint64_t(uint8_t(int8_t(res))) << (8 * q);
    }
    *reinterpret_cast<uint64_t*>(dst + l * plane) = word;
  }
}

// x(r, p) = x[r * ld + p] (p contiguous): each thread reads its 8 elements
// directly, all loads in flight at once (out-of-range elements read x[0]
// and become zero). Block: 16 rows x 128 p.
__global__ void __launch_bounds__(256)
    SliceRowsKernel(const double* x, int64_t ld, int64_t R, int64_t K, int64_t RP, int64_t KP,
                    const int* s0, const EmuCtl* ctl, int side, int8_t* Y) {
  const Decision d = ctl->dec;
  if (d.count == 0) return;
  const Crt& crt = kCrtTable.t[d.count];
  const int bits = side == 0 ? d.ba : d.bb;
  const int tid = threadIdx.x;
  const int64_t r = int64_t(blockIdx.y) * 16 + (tid >> 4);
  const int64_t p0 = int64_t(blockIdx.x) * 128 + (tid & 15) * 8;
  if (r >= RP) return;
  double v[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) {
    const bool ok = r < R && p0 + q < K;
    const double t = x[ok ? r * ld + p0 + q : 0];
    v[q] = ok ? t : 0.0;
  }
  SliceEight(v, r < R ? bits + s0[r] : 0, crt, Y + r * KP + p0, RP * KP);
}

// x(r, p) = x[r + p * ld] (r contiguous): a 32 x 128 tile is transposed
// through LDS, read coalesced along r.
__global__ void __launch_bounds__(256)
    SliceColsKernel(const double* x, int64_t ld, int64_t R, int64_t K, int64_t RP, int64_t KP,
                    const int* s0, const EmuCtl* ctl, int side, int8_t* Y) {
  const Decision d = ctl->dec;
  if (d.count == 0) return;
  const Crt& crt = kCrtTable.t[d.count];
  const int bits = side == 0 ? d.ba : d.bb;
  constexpr int TR = 32, TP = 128, PER = TR * TP / 256;
  __shared__ double T[TP][TR + 1];
  const int tid = threadIdx.x;
  const int64_t r0 = int64_t(blockIdx.y) * TR, p0 = int64_t(blockIdx.x) * TP;
  // Every load in flight before the first LDS store (out-of-range elements
  // read x[0] and become zero).
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
    if (r >= RP) break;
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = T[pq + q][rr];
    SliceEight(v, r < R ? bits + s0[r] : 0, crt, Y + r * KP + p0 + pq, RP * KP);
  }
}

// ---------------------------------------------------------------------------
// INT8 GEMM per modulus: R_l = (A_l B_l^T) mod p_l, with A_l (MP x KP) and
// B_l (NP x KP) row-major int8 planes (k contiguous, zero padded) and R_l a
// column-major MP x NP byte plane. blockIdx.y selects the modulus.
//
// v_mfma_i32_32x32x32_i8: lane l supplies 16 consecutive k bytes of row
// l % 32 (k = 16 (l / 32) + byte within the 32-k step) for A, likewise of
// column l % 32 for B; result element r of lane l is column l % 32, row
// 8 (r / 4) + 4 (l / 32) + r % 4.
// LDS holds each BM x BK (BN x BK) slice as rows of BK bytes in 16-byte
// chunks, chunk c of row r stored at c ^ ((r / RPB) % CH) (CH chunks per
// row, RPB rows per 256-byte bank row) so the MFMA operand reads of 16
// consecutive rows hit distinct banks.

typedef int i32x16 __attribute__((ext_vector_type(16)));

struct ModGemmArgs {
  const int8_t* a;
  const int8_t* b;
  uint8_t* r;
  int64_t a_plane, b_plane, r_plane;
  int64_t kp, ldr;
  int tiles_m, tiles_n;
  int p[kMaxModuli];
  double inv_p[kMaxModuli];
  EmuCtl* ctl;
  // The coarse product's epilogue: the exact sums of the rows of op(A)
  // and the columns of op(B), and the m x n pairs that count.
  const uint64_t* rs;
  const uint64_t* cs;
  int64_t m, n;
};

// The same product with the operand slices streamed straight into LDS
// (global_load_lds_dwordx4: each wave writes 64 consecutive 16-byte chunks,
// lane i at base + 16 i) through a STAGES-deep ring of LDS buffers, so that
// loads for STAGES - 1 slices are in flight while one is multiplied and no
// registers are spent on staging. Lane i of a wave's DMA picks the global
// chunk that belongs at its LDS position under the swizzle (the inverse of
// swz). The waves synchronise with s_barrier after waiting (vmcnt) for their
// own DMA of the slice about to be read; the slot refilled after the barrier
// is the one every wave finished reading in the previous iteration. The
// next slice's A loads go out after the first MFMA step and its B loads
// after the second, not in one burst with every wave's first LDS reads.
// EPI 0: the residue planes, reduced mod p_l (blockIdx.y = l; planes beyond
// the decided count do nothing). EPI 1: the coarse product G of the
// decision (plane 0), reduced to the maximum ratio and minimum G.
template <int BM, int BN, int BK, int WAVES_M, int WAVES_N, int STAGES, int EPI>
__global__ void __launch_bounds__(64 * WAVES_M * WAVES_N) ModGemmDmaKernel(ModGemmArgs g) {
  constexpr int NT = 64 * WAVES_M * WAVES_N;
  constexpr int NW = WAVES_M * WAVES_N;
  constexpr int WM = BM / WAVES_M, WN = BN / WAVES_N;
  constexpr int TM = WM / 32, TN = WN / 32;
  constexpr int CH = BK / 16;   // 16-byte chunks per row
  constexpr int RPB = 16 / CH;  // rows per 256-byte bank row
  constexpr int A_DMA = BM * CH / NT;  // DMA instructions per thread per slice
  constexpr int B_DMA = BN * CH / NT;
  constexpr int PER_SLICE = A_DMA + B_DMA;
  static_assert(BM * CH % NT == 0 && BN * CH % NT == 0, "load shape");
  static_assert(PER_SLICE * 3 < 64 && STAGES <= 5, "vmcnt table");
  __shared__ i32x4 As[STAGES][BM * CH];
  __shared__ i32x4 Bs[STAGES][BN * CH];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave % WAVES_M) * WM, wn0 = (wave / WAVES_M) * WN;
  const int l = blockIdx.y;
  if (EPI == 0 && l >= g.ctl->dec.count) return;  // uniform
  int tm, tn;
  XcdTile(g.tiles_m, g.tiles_n, &tm, &tn);
  const int64_t m0 = int64_t(tm) * BM, n0 = int64_t(tn) * BN;
  const int8_t* A = g.a + l * g.a_plane + m0 * g.kp;
  const int8_t* B = g.b + l * g.b_plane + n0 * g.kp;

  // LDS chunk position P -> the (row, chunk) it holds: row = P / CH,
  // c = (P % CH) ^ ((row / RPB) % CH). Lane offsets within each slice:
  int64_t a_off[A_DMA], b_off[B_DMA];
#pragma unroll
  for (int q = 0; q < A_DMA; ++q) {
    const int P = (q * NW + wave) * 64 + lane;
    const int row = P / CH, c = (P % CH) ^ ((row / RPB) % CH);
    a_off[q] = int64_t(row) * g.kp + c * 16;
  }
#pragma unroll
  for (int q = 0; q < B_DMA; ++q) {
    const int P = (q * NW + wave) * 64 + lane;
    const int row = P / CH, c = (P % CH) ^ ((row / RPB) % CH);
    b_off[q] = int64_t(row) * g.kp + c * 16;
  }
  auto issue_a = [&](int stage, int64_t k0) {
#pragma unroll
    for (int q = 0; q < A_DMA; ++q) DmaToLds(A + a_off[q] + k0, &As[stage][(q * NW + wave) * 64]);
  };
  auto issue_b = [&](int stage, int64_t k0) {
#pragma unroll
    for (int q = 0; q < B_DMA; ++q) DmaToLds(B + b_off[q] + k0, &Bs[stage][(q * NW + wave) * 64]);
  };
  auto issue = [&](int stage, int64_t k0) {
    issue_a(stage, k0);
    issue_b(stage, k0);
  };
  auto swz = [](int row, int c) { return row * CH + (c ^ ((row / RPB) % CH)); };

  i32x16 acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = i32x16{};

  const int ktiles = int(g.kp / BK);
#pragma unroll
  for (int s = 0; s < STAGES - 1; ++s)
    if (s < ktiles) issue(s, int64_t(s) * BK);
  const int lr = lane & 31, lh = lane >> 5;
  for (int kt = 0; kt < ktiles; ++kt) {
    // Slices issued after kt: min(STAGES - 2, ktiles - 1 - kt).
    const int later = min(STAGES - 2, ktiles - 1 - kt);
    WaitSlices<PER_SLICE>(later);
    BarrierNoFence();
    const int next = kt + STAGES - 1;
    const int cur = kt % STAGES;
#pragma unroll
    for (int ks = 0; ks < BK / 32; ++ks) {
      if (next < ktiles) {
        if (ks == 1) issue_a(next % STAGES, int64_t(next) * BK);
        if (ks == 2) issue_b(next % STAGES, int64_t(next) * BK);
      }
      const int c = ks * 2 + lh;
      i32x4 af[TM], bf[TN];
#pragma unroll
      for (int ti = 0; ti < TM; ++ti) af[ti] = As[cur][swz(wm0 + ti * 32 + lr, c)];
#pragma unroll
      for (int tj = 0; tj < TN; ++tj) bf[tj] = Bs[cur][swz(wn0 + tj * 32 + lr, c)];
#pragma unroll
      for (int ti = 0; ti < TM; ++ti)
#pragma unroll
        for (int tj = 0; tj < TN; ++tj)
          acc[ti][tj] = __builtin_amdgcn_mfma_i32_32x32x32_i8(af[ti], bf[tj], acc[ti][tj], 0, 0, 0);
    }
  }

  if (EPI == 1) {
    // max (cs_j + rs_i) 2^-27 / G_ij and min G_ij over the pairs whose row
    // and column are not zero; one atomic each per wave.
    uint64_t xmax = 0;
    int gmin = 0x7FFFFFFF;
#pragma unroll
    for (int ti = 0; ti < TM; ++ti)
#pragma unroll
      for (int tj = 0; tj < TN; ++tj) {
        const int64_t col = n0 + wn0 + tj * 32 + lr;
        const uint64_t cj = col < g.n ? g.cs[col] : 0;
#pragma unroll
        for (int e = 0; e < 16; ++e) {
          const int64_t row = m0 + wm0 + ti * 32 + 8 * (e / 4) + 4 * lh + e % 4;
          const uint64_t ri = row < g.m ? g.rs[row] : 0;
          if (ri == 0 || cj == 0) continue;
          const int G = acc[ti][tj][e];
          const double x = G == 0 ? __longlong_as_double(int64_t(kInfBits))
                                  : (double(cj) + double(ri)) * 0x1p-27 / double(G);
          const uint64_t xb = uint64_t(__double_as_longlong(x));
          xmax = xb > xmax ? xb : xmax;
          gmin = G < gmin ? G : gmin;
        }
      }
    for (int off = 32; off > 0; off >>= 1) {
      const uint64_t ox = __shfl_xor(xmax, off, 64);
      const int og = __shfl_xor(gmin, off, 64);
      xmax = ox > xmax ? ox : xmax;
      gmin = og < gmin ? og : gmin;
    }
    if (lane == 0) {
      if (xmax != 0) atomicMax(reinterpret_cast<unsigned long long*>(&g.ctl->xmax_bits), xmax);
      if (gmin != 0x7FFFFFFF) atomicMin(&g.ctl->gmin, gmin);
    }
    return;
  }

  const int p = g.p[l];
  const double ip = g.inv_p[l];
  uint8_t* R = g.r + l * g.r_plane;
#pragma unroll
  for (int ti = 0; ti < TM; ++ti) {
#pragma unroll
    for (int tj = 0; tj < TN; ++tj) {
      const int64_t col = n0 + wn0 + tj * 32 + lr;
#pragma unroll
      for (int grp = 0; grp < 4; ++grp) {
        const int64_t row = m0 + wm0 + ti * 32 + 8 * grp + 4 * lh;
        uint32_t word = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q)
          word |= uint32_t(ModU(acc[ti][tj][grp * 4 + q], p, ip)) << (8 * q);
        *reinterpret_cast<uint32_t*>(R + row + col * g.ldr) = word;
      }
    }
  }
}

// C(i, j) = alpha 2^-(sa_i + sb_j) C'(i, j) + beta C(i, j), C' from the
// residue planes through the fixed-point CRT sum S = sum_l c_l F_l mod 2^160,
// F_l = floor(2^160 y_l / p_l), held as four 40-bit limbs in doubles: every
// partial sum of c_l F_lt (c_l < 256, N <= 24) stays below 2^53, so the
// limb sums are exact FP64 FMAs. Carries are then propagated exactly, the
// top limb is wrapped mod 2^40 and centred, and S / 2^160 times P is C'. Each
// thread handles four consecutive rows of one column (one 32-bit residue
// word per modulus).
__device__ __forceinline__ double CrtValue(double (&S)[kLimbs], double P) {
#pragma unroll
  for (int t = 0; t + 1 < kLimbs; ++t) {
    const double carry = floor(S[t] * 0x1p-40);
    S[t] = __builtin_fma(-carry, 0x1p40, S[t]);
    S[t + 1] += carry;
  }
  S[kLimbs - 1] = __builtin_fma(-floor(S[kLimbs - 1] * 0x1p-40), 0x1p40, S[kLimbs - 1]);
  // A fraction >= 1/2 stands for a negative C': take 2^160 - S exactly (limb
  // complement plus one), so that the conversion below adds non-negative
  // terms only.
  const bool neg = S[kLimbs - 1] >= 0x1p39;
  if (neg) {
    double carry = 1;
#pragma unroll
    for (int t = 0; t < kLimbs; ++t) {
      S[t] = (0x1p40 - 1) - S[t] + carry;
      carry = (S[t] >= 0x1p40) ? 1.0 : 0.0;
      S[t] -= carry * 0x1p40;
    }
  }
  double v = S[0] * 0x1p-160;
  v = v + S[1] * 0x1p-120;
  v = v + S[2] * 0x1p-80;
  v = v + S[3] * 0x1p-40;
  v = v * P;
  return neg ? -v : v;
}

__global__ void __launch_bounds__(256)
    ReconstructKernel(const uint8_t* R, int64_t r_plane, int64_t ldr, int64_t m, int64_t n,
                      const int* s0a, const int* s0b, const EmuCtl* ctl, double alpha,
                      double beta, double* c, int64_t ldc) {
  const Decision d = ctl->dec;
  if (d.count == 0) return;
  const Crt& crt = kCrtTable.t[d.count];
  const int bits = d.ba + d.bb;
  const int64_t quads = (m + 3) / 4;
  const int64_t total = quads * n;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
       idx += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i0 = (idx % quads) * 4, j = idx / quads;
    double S[4][kLimbs] = {};
    const uint8_t* rp = R + i0 + j * ldr;
    // Every plane's word in flight at once (nmod is uniform).
    uint32_t words[kMaxModuli];
#pragma unroll
    for (int l = 0; l < kMaxModuli; ++l)
      words[l] = l < crt.nmod ? *reinterpret_cast<const uint32_t*>(rp + l * r_plane) : 0u;
#pragma unroll
    for (int l = 0; l < kMaxModuli; ++l) {
      if (l >= crt.nmod) break;
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        const double cl = double((words[l] >> (8 * q)) & 0xFFu);
#pragma unroll
        for (int t = 0; t < kLimbs; ++t) S[q][t] = __builtin_fma(cl, crt.frac[l][t], S[q][t]);
      }
    }
    const int sbj = bits + s0b[j];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      const int64_t i = i0 + q;
      if (i >= m) break;
      const double x = ldexp(CrtValue(S[q], crt.P), -(s0a[i] + sbj));
      double* cij = c + i + j * ldc;
      double v = x * alpha;
      if (beta != 0.0) v = (beta == 1.0) ? v + *cij : __builtin_fma(*cij, beta, v);
      *cij = v;
    }
  }
}

constexpr int kTileM = 256, kTileN = 256;
// k of the padded slices: a multiple of every variant's k step.
constexpr int kPadK = 128;

}  // namespace

void* GetWorkspace(size_t bytes, hipStream_t stream, bool* grown) {
  return Staging(kSlotEmulated, bytes, stream, grown);
}

struct PhaseTimer {
  bool enabled = false;
  bool recorded = false;
  hipEvent_t ev[kNumPhases + 1] = {};
};
thread_local PhaseTimer t_timer;

void Mark(int phase_boundary, hipStream_t stream) {
  PhaseTimer& t = t_timer;
  if (!t.enabled) return;
  if (!t.ev[0])
    for (auto& e : t.ev) HWY_HIP_CHECK(hipEventCreate(&e));
  HWY_HIP_CHECK(hipEventRecord(t.ev[phase_boundary], stream));
  if (phase_boundary == kNumPhases) t.recorded = true;
}

void ResetPhases() { t_timer.recorded = false; }

size_t StatsBytes(int64_t m, int64_t n) {
  return AlignUp(sizeof(EmuCtl), 256) + AlignUp(sizeof(uint64_t) * (m + n), 256) * 2 +
         AlignUp(sizeof(int) * m, 256) + AlignUp(sizeof(int) * n, 256);
}

EmuStats CarveStats(void* base, int64_t m, int64_t n) {
  char* p = static_cast<char*>(base);
  EmuStats st;
  st.ctl = reinterpret_cast<EmuCtl*>(p);
  p += AlignUp(sizeof(EmuCtl), 256);
  st.rowmax = reinterpret_cast<uint64_t*>(p);
  st.sums = st.rowmax + (m + n);  // contiguous with rowmax: zeroed together
  p += AlignUp(sizeof(uint64_t) * (m + n), 256) * 2;
  st.sa = reinterpret_cast<int*>(p);
  p += AlignUp(sizeof(int) * m, 256);
  st.sb = reinterpret_cast<int*>(p);
  return st;
}

namespace {

int TotalBitsTable(int nmod, int64_t k) { return std::max(TotalBits(nmod, k), 2); }

}  // namespace

void Decide(int kind, int forced, const EmuOperands& op, const EmuStats& st, int scale_bits,
            uint8_t* coarse_a, uint8_t* coarse_b, int64_t MPc, int64_t NPc, int64_t KPc,
            hipStream_t stream) {
  const int64_t m = op.m, n = op.n, k = op.k;
  // The sums follow the maxima directly: both are zeroed in one pass.
  const int64_t words = 2 * (m + n);
  hipLaunchKernelGGL(InitStatsKernel, dim3(unsigned(std::min<int64_t>((words + 255) / 256, 4096))),
                     dim3(256), 0, stream, st.rowmax, words, st.ctl);
  const int64_t pchunk = 512;
  auto rowmax_launch = [&](const double* x, int64_t rs, int64_t cs, int64_t R, uint64_t* out) {
    const unsigned chunks = unsigned((k + pchunk - 1) / pchunk);
    if (rs == 1)
      hipLaunchKernelGGL(RowAbsMaxKernel<true>, dim3(unsigned((R + 255) / 256), chunks),
                         dim3(256), 0, stream, x, cs, R, k, pchunk, out);
    else
      hipLaunchKernelGGL(RowAbsMaxKernel<false>, dim3(unsigned((R + 3) / 4), chunks), dim3(256),
                         0, stream, x, rs, R, k, pchunk, out);
    HWY_HIP_CHECK(hipGetLastError());
  };
  rowmax_launch(op.a, op.a_rs, op.a_cs, m, st.rowmax);
  // op(B)^T(j, p) = b[j * b_cs + p * b_rs]
  rowmax_launch(op.b, op.b_cs, op.b_rs, n, st.rowmax + m);
  hipLaunchKernelGGL(ScaleExpKernel, dim3(unsigned((m + 255) / 256)), dim3(256), 0, stream,
                     st.rowmax, m, scale_bits, st.sa, &st.ctl->amax_bits);
  hipLaunchKernelGGL(ScaleExpKernel, dim3(unsigned((n + 255) / 256)), dim3(256), 0, stream,
                     st.rowmax + m, n, scale_bits, st.sb, &st.ctl->bmax_bits);
  HWY_HIP_CHECK(hipGetLastError());
  if (forced <= 0) {
    // Coarse planes (value 2^(7 - E) |x|) and exact sums, then their product.
    const int bias = 7 - scale_bits;
    auto coarse = [&](const double* x, int64_t rs, int64_t cs, int64_t R, int64_t RP,
                      const int* s, uint8_t* y, uint64_t* sums) {
      if (rs == 1)
        hipLaunchKernelGGL(CoarseColsKernel, dim3(unsigned(KPc / 128), unsigned(RP / 32)),
                           dim3(256), 0, stream, x, cs, R, k, RP, KPc, s, bias, y, sums);
      else
        hipLaunchKernelGGL(CoarseRowsKernel, dim3(unsigned(KPc / 128), unsigned(RP / 16)),
                           dim3(256), 0, stream, x, rs, R, k, RP, KPc, s, bias, y, sums);
      HWY_HIP_CHECK(hipGetLastError());
    };
    coarse(op.a, op.a_rs, op.a_cs, m, MPc, st.sa, coarse_a, st.sums);
    coarse(op.b, op.b_cs, op.b_rs, n, NPc, st.sb, coarse_b, st.sums + m);
    ModGemmArgs g{};
    g.a = reinterpret_cast<const int8_t*>(coarse_a);
    g.b = reinterpret_cast<const int8_t*>(coarse_b);
    g.kp = KPc;
    g.tiles_m = int(MPc / kTileM);
    g.tiles_n = int(NPc / kTileN);
    g.ctl = st.ctl;
    g.rs = st.sums;
    g.cs = st.sums + m;
    g.m = m;
    g.n = n;
    hipLaunchKernelGGL((ModGemmDmaKernel<kTileM, kTileN, 128, 2, 4, 2, 1>),
                       dim3(unsigned(g.tiles_m * g.tiles_n), 1u), dim3(512), 0, stream, g);
    HWY_HIP_CHECK(hipGetLastError());
  }
  DecideArgs d{};
  d.ctl = st.ctl;
  d.kind = kind;
  d.forced = forced;
  d.k = k;
  for (int c = 0; c <= kMaxCount; ++c) d.bits[c] = c >= 1 ? TotalBitsTable(c, k) : 0;
  hipLaunchKernelGGL(DecideKernel, dim3(1), dim3(64), 0, stream, d);
  HWY_HIP_CHECK(hipGetLastError());
}

int ReportCount(const EmuStats& st, hipStream_t stream) {
  int count = 0;
  HWY_HIP_CHECK(hipMemcpyAsync(&count, &st.ctl->dec.count, sizeof(int), hipMemcpyDeviceToHost,
                               stream));
  HWY_HIP_CHECK(hipStreamSynchronize(stream));
  return count;
}

namespace {

// Workspace of one call: statistics, then kMaxModuli planes of each kind
// (the count is decided on the device).
struct EmuLayout {
  int64_t MP, NP, KP;
  size_t o_a8, o_b8, o_r, total;
};

EmuLayout LayoutOf(int64_t m, int64_t n, int64_t k) {
  EmuLayout w;
  w.MP = AlignUp(m, kTileM);
  w.NP = AlignUp(n, kTileN);
  w.KP = AlignUp(k, kPadK);
  w.o_a8 = AlignUp(StatsBytes(m, n), 256);
  w.o_b8 = w.o_a8 + AlignUp(size_t(kMaxModuli) * w.MP * w.KP, 256);
  w.o_r = w.o_b8 + AlignUp(size_t(kMaxModuli) * w.NP * w.KP, 256);
  w.total = w.o_r + size_t(kMaxModuli) * w.MP * w.NP;
  return w;
}

// The emulated DGEMM on a given workspace (at least LayoutOf(...).total
// bytes), or on the calling thread's grow-only one when `workspace` is null.
void EmulatedDgemmOn(bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha,
                     const double* a, int64_t lda, const double* b, int64_t ldb, double beta,
                     double* c, int64_t ldc, int nmod, hipStream_t stream, int* used,
                     void* workspace) {
  ResetPhases();
  if (k >= 131072) {  // beyond exact INT32 accumulation of the products
    GemmDevice<double>(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, stream);
    if (used) *used = 0;
    return;
  }
  EmuOperands op;
  op.a = a;
  op.a_rs = ta ? lda : 1;
  op.a_cs = ta ? 1 : lda;
  op.b = b;
  op.b_rs = tb ? ldb : 1;
  op.b_cs = tb ? 1 : ldb;
  op.m = m;
  op.n = n;
  op.k = k;
  const EmuLayout w = LayoutOf(m, n, k);
  const int64_t MP = w.MP, NP = w.NP, KP = w.KP;
  char* ws = static_cast<char*>(workspace ? workspace : GetWorkspace(w.total, stream));
  const EmuStats st = CarveStats(ws, m, n);
  int8_t* a8 = reinterpret_cast<int8_t*>(ws + w.o_a8);
  int8_t* b8 = reinterpret_cast<int8_t*>(ws + w.o_b8);
  uint8_t* rr = reinterpret_cast<uint8_t*>(ws + w.o_r);

  // 1. Statistics and the count of moduli (plane 0 holds the coarse
  //    magnitudes until the slices overwrite it).
  Mark(kPhaseStats, stream);
  Decide(0, nmod, op, st, 0, reinterpret_cast<uint8_t*>(a8), reinterpret_cast<uint8_t*>(b8), MP,
         NP, KP, stream);

  // 2. The int8 slices.
  Mark(kPhaseSlice, stream);
  auto slice = [&](const double* x, int64_t rs, int64_t cs, int64_t R, int64_t RP, const int* s0,
                   int side, int8_t* y) {
    if (rs == 1)
      hipLaunchKernelGGL(SliceColsKernel, dim3(unsigned(KP / 128), unsigned(RP / 32)), dim3(256),
                         0, stream, x, cs, R, k, RP, KP, s0, st.ctl, side, y);
    else
      hipLaunchKernelGGL(SliceRowsKernel, dim3(unsigned(KP / 128), unsigned(RP / 16)), dim3(256),
                         0, stream, x, rs, R, k, RP, KP, s0, st.ctl, side, y);
    HWY_HIP_CHECK(hipGetLastError());
  };
  slice(op.a, op.a_rs, op.a_cs, m, MP, st.sa, 0, a8);
  slice(op.b, op.b_cs, op.b_rs, n, NP, st.sb, 1, b8);

  // 3. One INT8 GEMM per modulus, all in one launch (planes beyond the
  //    decided count exit at once).
  ModGemmArgs g{};
  g.a = a8;
  g.b = b8;
  g.r = rr;
  g.a_plane = MP * KP;
  g.b_plane = NP * KP;
  g.r_plane = MP * NP;
  g.kp = KP;
  g.ldr = MP;
  g.tiles_m = int(MP / kTileM);
  g.tiles_n = int(NP / kTileN);
  g.ctl = st.ctl;
  for (int l = 0; l < kMaxModuli; ++l) {
    g.p[l] = kModuli[l];
    g.inv_p[l] = 1.0 / kModuli[l];
  }
  Mark(kPhaseModGemm, stream);
  hipLaunchKernelGGL((ModGemmDmaKernel<kTileM, kTileN, 128, 2, 4, 2, 0>),
                     dim3(unsigned(g.tiles_m * g.tiles_n), unsigned(kMaxModuli)), dim3(512), 0,
                     stream, g);
  HWY_HIP_CHECK(hipGetLastError());

  // 4. Reconstruction and the alpha/beta epilogue, or the native kernel.
  Mark(kPhaseReconstruct, stream);
  const int64_t quads = ((m + 3) / 4) * n;
  hipLaunchKernelGGL(ReconstructKernel,
                     dim3(unsigned(std::min<int64_t>((quads + 255) / 256, 1 << 20))), dim3(256), 0,
                     stream, rr, MP * NP, MP, m, n, st.sa, st.sb, st.ctl, alpha, beta, c, ldc);
  HWY_HIP_CHECK(hipGetLastError());
  GemmDeviceGated<double>(&st.ctl->dec.count, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c,
                          ldc, stream);
  Mark(kNumPhases, stream);
  if (used) *used = ReportCount(st, stream);
}

// The workspace of the entry that allocates nothing: a fixed array in every
// device's copy of the code object, held from load until the process ends.
constexpr size_t kScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// The device address of the current device's `scratch` (each GPU holds its
// own copy at its own address), looked up once per device. The array's name
// is a host-side stand-in and is never passed to a kernel or a copy.
void* ScratchAddress() {
  constexpr int kDevices = 64;
  static std::mutex mu;
  static void* address[kDevices] = {};
  int dev = 0;
  HWY_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kDevices) return nullptr;
  std::lock_guard<std::mutex> lock(mu);
  if (address[dev] == nullptr) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) {
      fprintf(stderr, "hwyblas_amd: hipGetSymbolAddress(scratch) failed on device %d: %s\n",
              dev, hipGetErrorString(e));
      return nullptr;
    }
    address[dev] = p;
  }
  return address[dev];
}

}  // namespace

void EmulatedDgemm(bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha,
                   const double* a, int64_t lda, const double* b, int64_t ldb, double beta,
                   double* c, int64_t ldc, int nmod, hipStream_t stream, int* used) {
  EmulatedDgemmOn(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, nmod, stream, used,
                  nullptr);
}

}  // namespace ozaki
}  // namespace hwy_amd

using namespace hwy_amd;

extern "C" {

HWYBLAS_AMD_API int hwyblas_dgemm_emulated_dev(char transa, char transb, int m, int n, int k,
                                               double alpha, const double* a, int lda,
                                               const double* b, int ldb, double beta, double* c,
                                               int ldc, int nmoduli, void* stream,
                                               int* used_moduli) {
  Op oa, ob;
  const bool va = DecodeOp(transa, &oa), vb = DecodeOp(transb, &ob);
  const int rows_a = oa == Op::kNoTrans ? m : k, rows_b = ob == Op::kNoTrans ? k : n;
  int info = 0;
  if (!va) info = 1;
  else if (!vb) info = 2;
  else if (m < 0) info = 3;
  else if (n < 0) info = 4;
  else if (k < 0) info = 5;
  else if (lda < std::max(1, rows_a)) info = 8;
  else if (ldb < std::max(1, rows_b)) info = 10;
  else if (ldc < std::max(1, m)) info = 13;
  if (info != 0) {
    Reject("DGEMM", info);
    return info;
  }
  if (used_moduli) *used_moduli = 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (m == 0 || n == 0) return 0;
  if ((k == 0 || alpha == 0.0) && beta == 1.0) return 0;
  if (k == 0 || alpha == 0.0) {
    ScaleDevice<double>(m, n, beta, c, ldc, s);
    return 0;
  }
  ozaki::EmulatedDgemm(oa == Op::kTrans, ob == Op::kTrans, m, n, k, alpha, a, lda, b, ldb, beta,
                       c, ldc, nmoduli, s, used_moduli);
  return 0;
}

HWYBLAS_AMD_API int hwyblas_dgemm_emulated_noalloc_dev(char transa, char transb, int m, int n,
                                                       int k, double alpha, const double* a,
                                                       int lda, const double* b, int ldb,
                                                       double beta, double* c, int ldc,
                                                       int nmoduli, void* stream,
                                                       int* used_moduli) {
  Op oa, ob;
  const bool va = DecodeOp(transa, &oa), vb = DecodeOp(transb, &ob);
  const int rows_a = oa == Op::kNoTrans ? m : k, rows_b = ob == Op::kNoTrans ? k : n;
  int info = 0;
  if (!va) info = 1;
  else if (!vb) info = 2;
  else if (m < 0) info = 3;
  else if (n < 0) info = 4;
  else if (k < 0) info = 5;
  else if (lda < std::max(1, rows_a)) info = 8;
  else if (ldb < std::max(1, rows_b)) info = 10;
  else if (ldc < std::max(1, m)) info = 13;
  if (info != 0) {
    Reject("DGEMM", info);
    return info;
  }
  if (used_moduli) *used_moduli = 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (m == 0 || n == 0) return 0;
  if ((k == 0 || alpha == 0.0) && beta == 1.0) return 0;
  if (k == 0 || alpha == 0.0) {
    ScaleDevice<double>(m, n, beta, c, ldc, s);
    return 0;
  }
  const bool ta = oa == Op::kTrans, tb = ob == Op::kTrans;
  void* ws = nullptr;
  if (k < 131072) {  // beyond, the native kernel runs and no workspace is used
    const size_t need = ozaki::LayoutOf(m, n, k).total;
    if (need > ozaki::kScratchBytes) {
      fprintf(stderr,
              "hwyblas_dgemm_emulated_noalloc_dev: m = %d, n = %d, k = %d needs %zu bytes of "
              "workspace, more than the %zu bytes the library holds; nothing was done\n",
              m, n, k, need, ozaki::kScratchBytes);
      return -1;
    }
    ws = ozaki::ScratchAddress();
    if (ws == nullptr) return -2;
  }
  ozaki::EmulatedDgemmOn(ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc, nmoduli, s,
                         used_moduli, ws);
  return 0;
}

HWYBLAS_AMD_API void hwyblas_amd_profile_enable(int on) {
  ozaki::t_timer.enabled = on != 0;
  ozaki::t_timer.recorded = false;
}

HWYBLAS_AMD_API float hwyblas_amd_profile_ms(int phase) {
  ozaki::PhaseTimer& t = ozaki::t_timer;
  if (!t.recorded || phase < 0 || phase >= ozaki::kNumPhases) return -1.0f;
  HWY_HIP_CHECK(hipEventSynchronize(t.ev[phase + 1]));
  float ms = 0;
  HWY_HIP_CHECK(hipEventElapsedTime(&ms, t.ev[phase], t.ev[phase + 1]));
  return ms;
}

}  // extern "C"
