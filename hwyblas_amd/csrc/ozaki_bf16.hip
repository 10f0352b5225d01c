// Emulated DGEMM on the BF16 matrix cores (v_mfma_f32_32x32x16_bf16): FP64
// in, FP64 out, by the Ozaki scheme I with BF16 slices (Ozaki, Ogita, Oishi,
// Rump, "Error-free transformations of matrix multiplication by using fast
// routines of matrix multiplication and its applications", Numer.
// Algorithms 59, 2012; with the k-chunked exact FP32 accumulation used for
// low-precision matrix cores), restated:
//
//  1. Scaling. Row i of op(A) is scaled by 2^sa_i, sa_i = -(E_i + 1) where
//     max_p |a_ip| = f 2^E_i, f in [0.5, 1), so every scaled entry has
//     |a^| < 1/2; likewise column j of op(B) by 2^sb_j.
//  2. Slices. a^ = sum_{t=1..s} d_t 2^-8t + r_s with balanced integer digits
//     d_t = rint(256 r_{t-1}) (nearest, ties to even), r_0 = a^,
//     r_t = 256 r_{t-1} - d_t, all exact in FP64 (|r_s| <= 2^-8s / 2).
//     |d_t| <= 128, so every digit is exact in BF16.
//  3. Products. For the s (s + 1) / 2 pairs (t, u) with t + u <= s + 1, taken
//     by level l = t + u from s + 1 down to 2 and t ascending within a level,
//     D_A,t D_B,u^T over k in chunks of 1024 (the last one of a pair may be
//     shorter): a chunk's sum is an integer of magnitude at most
//     1024 * 128^2 = 2^24, exact in the MFMA's FP32 accumulation whatever
//     its internal order, and is added into an FP64 accumulator as
//     acc = fma(chunk, 2^-8l, acc), one rounding per chunk in a fixed order
//     (so the result is reproducible bit for bit).
//  4. C = alpha 2^-(sa_i + sb_j) acc + beta C (hwyBLAS's Store rules), in the
//     GEMM's epilogue.
// The slice count is given by the caller or chosen on the device from exact
// statistics of every pair (i, j) (DecideKernel, ozaki_shared.h): the
// smallest count whose error bound stays at native FP64's typical level,
// else the smallest that keeps hwyBLAS's componentwise bound, else the
// native FP64 kernel runs, as it does for operands with NaN or Inf.
//
// BF16 products are exact only 8 bits at a time, so FP64 accuracy takes
// about s (s + 1) / 2 = 28 slice products (s = 7); the INT8 scheme II needs
// 14 INT8 products for the same accuracy at twice the MFMA rate.
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "common.h"
#include "ozaki_shared.h"

namespace hwy_amd {
namespace ozaki {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int kMaxSlices = 16;
constexpr int kChunk = 1024;  // k per exact FP32 partial sum
constexpr int kBK = 64;       // k per LDS stage: rows of 128 bytes
constexpr int kPadK = 256;    // k of the digit planes: a multiple of this
constexpr int kSliceTileM = 256, kSliceTileN = 128;

// The BF16 bits of an integer |d| <= 256 (exact: at most 8 significant
// bits, so the low half of its FP32 bits is zero).
__device__ __forceinline__ uint16_t Bf16OfDigit(double d) {
  return uint16_t(__float_as_uint(float(d)) >> 16);
}

// Digits of 8 consecutive entries of one row (scaled by 2^sc) into `slices`
// planes: 16 bytes per plane at dst + t * plane.
__device__ __forceinline__ void DigitsEight(const double (&x)[8], int sc, int slices,
                                            uint16_t* dst, int64_t plane) {
  double r[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) r[q] = ldexp(x[q], sc);
  for (int t = 0; t < slices; ++t) {
    uint16_t h[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) {
      const double y = r[q] * 256.0;
      const double d = rint(y);
      r[q] = y - d;
      h[q] = Bf16OfDigit(d);
    }
    i32x4 w;
#pragma unroll
    for (int q = 0; q < 4; ++q) w[q] = int(uint32_t(h[2 * q]) | (uint32_t(h[2 * q + 1]) << 16));
    *reinterpret_cast<i32x4*>(dst + t * plane) = w;
  }
}

// Y[t][r][p] (row stride KP, plane stride RP * KP), zero for r >= R or
// p >= K. x(r, p) = x[r * ld + p] (p contiguous). Block: 16 rows x 128 p.
__global__ void __launch_bounds__(256)
    DigitRowsKernel(const double* x, int64_t ld, int64_t R, int64_t K, int64_t RP, int64_t KP,
                    const int* s, const EmuCtl* ctl, uint16_t* Y) {
  const int slices = ctl->dec.count;
  if (slices == 0) return;
  const int tid = threadIdx.x;
  const int64_t r = int64_t(blockIdx.y) * 16 + (tid >> 4);
  const int64_t p0 = int64_t(blockIdx.x) * 128 + (tid & 15) * 8;
  if (r >= RP) return;
  double v[8];
#pragma unroll
  for (int q = 0; q < 8; ++q) v[q] = (r < R && p0 + q < K) ? x[r * ld + p0 + q] : 0.0;
  DigitsEight(v, r < R ? s[r] : 0, slices, Y + r * KP + p0, RP * KP);
}

// The same for x(r, p) = x[r + p * ld] (r contiguous): a 32 x 128 tile is
// transposed through LDS, read coalesced along r.
__global__ void __launch_bounds__(256)
    DigitColsKernel(const double* x, int64_t ld, int64_t R, int64_t K, int64_t RP, int64_t KP,
                    const int* s, const EmuCtl* ctl, uint16_t* Y) {
  const int slices = ctl->dec.count;
  if (slices == 0) return;
  constexpr int TR = 32, TP = 128;
  __shared__ double T[TP][TR + 1];
  const int tid = threadIdx.x;
  const int64_t r0 = int64_t(blockIdx.y) * TR, p0 = int64_t(blockIdx.x) * TP;
  for (int idx = tid; idx < TR * TP; idx += 256) {
    const int rr = idx % TR, pp = idx / TR;
    const int64_t r = r0 + rr, p = p0 + pp;
    T[pp][rr] = (r < R && p < K) ? x[r + p * ld] : 0.0;
  }
  __syncthreads();
  const int pq = (tid & 15) * 8;
  for (int rr = tid >> 4; rr < TR; rr += 16) {
    const int64_t r = r0 + rr;
    if (r >= RP) break;
    double v[8];
#pragma unroll
    for (int q = 0; q < 8; ++q) v[q] = T[pq + q][rr];
    DigitsEight(v, r < R ? s[r] : 0, slices, Y + r * KP + p0 + pq, RP * KP);
  }
}

struct SliceGemmArgs {
  const uint16_t* a;  // slices x MP x KP BF16 digits, k contiguous
  const uint16_t* b;  // slices x NP x KP
  int64_t a_plane, b_plane, kp;
  int tiles_m, tiles_n;
  const EmuCtl* ctl;  // the decided slice count (0: nothing to do)
  const int* sa;
  const int* sb;
  double alpha, beta;
  double* c;
  int64_t ldc, m, n;
};

// The slice products of one BM x BN tile of C and its epilogue. The pairs'
// k ranges are one sequence of k tiles (pair-major), streamed by
// global_load_lds_dwordx4 through a STAGES-deep ring of LDS buffers as in
// the INT8 kernel (same byte geometry: a 64-element BF16 row is 128 bytes
// in eight 16-byte chunks, chunk c of row r at c ^ ((r / 2) % 8)).
// v_mfma_f32_32x32x16_bf16: lane l supplies 8 consecutive k of row l % 32
// (k = 8 (l / 32) + element) for A, likewise of column l % 32 for B; result
// element e of lane l is column l % 32, row 8 (e / 4) + 4 (l / 32) + e % 4.
template <int BM, int BN, int WAVES_M, int WAVES_N, int STAGES>
__global__ void __launch_bounds__(64 * WAVES_M * WAVES_N) SliceGemmKernel(SliceGemmArgs g) {
  constexpr int NT = 64 * WAVES_M * WAVES_N, NW = WAVES_M * WAVES_N;
  constexpr int WM = BM / WAVES_M, WN = BN / WAVES_N;
  constexpr int TM = WM / 32, TN = WN / 32;
  constexpr int CH = kBK * 2 / 16;  // 16-byte chunks per row
  constexpr int RPB = 16 / CH;      // rows per 256-byte bank row
  constexpr int A_DMA = BM * CH / NT, B_DMA = BN * CH / NT;
  constexpr int PER_SLICE = A_DMA + B_DMA;
  constexpr int TILES_PER_CHUNK = kChunk / kBK;
  static_assert(kPadK % kBK == 0, "k tiles");
  static_assert(BM * CH % NT == 0 && BN * CH % NT == 0, "load shape");
  static_assert(WM % 32 == 0 && WN % 32 == 0, "tile shape");
  static_assert(PER_SLICE * 3 < 64 && STAGES >= 2 && STAGES <= 5, "vmcnt table");
  static_assert(STAGES * (BM + BN) * CH * 16 <= 160 * 1024, "LDS per workgroup");
  __shared__ i32x4 As[STAGES][BM * CH];
  __shared__ i32x4 Bs[STAGES][BN * CH];

  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm0 = (wave % WAVES_M) * WM, wn0 = (wave / WAVES_M) * WN;
  int tm, tn;
  XcdTile(g.tiles_m, g.tiles_n, &tm, &tn);
  const int64_t m0 = int64_t(tm) * BM, n0 = int64_t(tn) * BN;
  const uint16_t* A = g.a + m0 * g.kp;
  const uint16_t* B = g.b + n0 * g.kp;

  // LDS chunk position P holds (row P / CH, chunk (P % CH) ^ ((row / RPB) % CH)).
  // Element offsets fit 32 bits (the host checks BM * KP < 2^31).
  uint32_t a_off[A_DMA], b_off[B_DMA];
#pragma unroll
  for (int q = 0; q < A_DMA; ++q) {
    const int P = (q * NW + wave) * 64 + lane;
    const int row = P / CH, c = (P % CH) ^ ((row / RPB) % CH);
    a_off[q] = uint32_t(row) * uint32_t(g.kp) + uint32_t(c * 8);
  }
#pragma unroll
  for (int q = 0; q < B_DMA; ++q) {
    const int P = (q * NW + wave) * 64 + lane;
    const int row = P / CH, c = (P % CH) ^ ((row / RPB) % CH);
    b_off[q] = uint32_t(row) * uint32_t(g.kp) + uint32_t(c * 8);
  }
  // Pairs in order: level l = t + u from s + 1 down to 2, t ascending (1-based
  // digits). Both the prefetch and the flushes walk this order with cursors
  // of their own (wave-uniform scalars: no table lookups with a runtime
  // index, whose loads would join the DMA in the vmcnt count).
  const int KT = int(g.kp / kBK);  // k tiles per pair
  const int slices = g.ctl->dec.count;
  if (slices == 0) return;  // uniform: the native kernel runs instead
  const int total = slices * (slices + 1) / 2 * KT;
  int i_t = 0, i_u = slices - 1, i_l = slices + 1, i_kt = 0;  // next tile to fetch
  auto issue = [&](int stage) {
    const uint16_t* as = A + i_t * g.a_plane + int64_t(i_kt) * kBK;
    const uint16_t* bs = B + i_u * g.b_plane + int64_t(i_kt) * kBK;
#pragma unroll
    for (int i = 0; i < A_DMA; ++i) DmaToLds(as + a_off[i], &As[stage][(i * NW + wave) * 64]);
#pragma unroll
    for (int i = 0; i < B_DMA; ++i) DmaToLds(bs + b_off[i], &Bs[stage][(i * NW + wave) * 64]);
    if (++i_kt == KT) {
      i_kt = 0;
      ++i_t;
      if (--i_u < 0) {
        --i_l;
        i_t = 0;
        i_u = i_l - 2;
      }
    }
  };
  auto swz = [](int row, int c) { return row * CH + (c ^ ((row / RPB) % CH)); };

  f32x16 acc[TM][TN];
  double acc64[TM][TN][16];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) {
      acc[i][j] = f32x16{};
#pragma unroll
      for (int e = 0; e < 16; ++e) acc64[i][j][e] = 0.0;
    }

#pragma unroll
  for (int st = 0; st < STAGES - 1; ++st)
    if (st < total) issue(st);
  const int lr = lane & 31, lh = lane >> 5;
  int level = slices + 1, t_in_level = 0, kt = 0;  // the tile being multiplied
  for (int q = 0; q < total; ++q) {
    const int later = min(STAGES - 2, total - 1 - q);
    WaitSlices<PER_SLICE>(later);
    BarrierNoFence();
    const int next = q + STAGES - 1;
    if (next < total) issue(next % STAGES);
    const int cur = q % STAGES;
#pragma unroll
    for (int ks = 0; ks < kBK / 16; ++ks) {
      const int c = ks * 2 + lh;
      bf16x8 af[TM], bf[TN];
#pragma unroll
      for (int ti = 0; ti < TM; ++ti)
        af[ti] = __builtin_bit_cast(bf16x8, As[cur][swz(wm0 + ti * 32 + lr, c)]);
#pragma unroll
      for (int tj = 0; tj < TN; ++tj)
        bf[tj] = __builtin_bit_cast(bf16x8, Bs[cur][swz(wn0 + tj * 32 + lr, c)]);
#pragma unroll
      for (int ti = 0; ti < TM; ++ti)
#pragma unroll
        for (int tj = 0; tj < TN; ++tj)
          acc[ti][tj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(af[ti], bf[tj], acc[ti][tj], 0, 0, 0);
    }
    ++kt;
    if (kt % TILES_PER_CHUNK == 0 || kt == KT) {
      // 2^-8l, l <= 2 * kMaxSlices: a normal power of two.
      const double scale = __longlong_as_double(int64_t(1023 - 8 * level) << 52);
#pragma unroll
      for (int ti = 0; ti < TM; ++ti)
#pragma unroll
        for (int tj = 0; tj < TN; ++tj) {
#pragma unroll
          for (int e = 0; e < 16; ++e)
            acc64[ti][tj][e] = __builtin_fma(double(acc[ti][tj][e]), scale, acc64[ti][tj][e]);
          acc[ti][tj] = f32x16{};
        }
      if (kt == KT) {
        kt = 0;
        if (++t_in_level == level - 1) {
          --level;
          t_in_level = 0;
        }
      }
    }
  }

  // Epilogue: C = alpha 2^-(sa_i + sb_j) acc + beta C.
  const double alpha = g.alpha, beta = g.beta;
#pragma unroll
  for (int tj = 0; tj < TN; ++tj) {
    const int64_t col = n0 + wn0 + tj * 32 + lr;
    if (col >= g.n) continue;
    const int sbj = g.sb[col];
    double* cc = g.c + col * g.ldc;
#pragma unroll
    for (int ti = 0; ti < TM; ++ti) {
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        const int64_t row = m0 + wm0 + ti * 32 + 8 * (e / 4) + 4 * lh + e % 4;
        if (row >= g.m) continue;
        const double x = ldexp(acc64[ti][tj][e], -(g.sa[row] + sbj));
        double v = x * alpha;
        if (beta != 0.0) v = (beta == 1.0) ? v + cc[row] : __builtin_fma(cc[row], beta, v);
        cc[row] = v;
      }
    }
  }
}

}  // namespace

void EmulatedDgemmBf16(bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha,
                       const double* a, int64_t lda, const double* b, int64_t ldb, double beta,
                       double* c, int64_t ldc, int slices, hipStream_t stream, int* used) {
  ResetPhases();
  const int64_t MP = AlignUp(m, kSliceTileM), NP = AlignUp(n, kSliceTileN);
  const int64_t KP = AlignUp(k, kPadK);
  // 32-bit DMA offsets; the coarse INT8 product sums exactly for k < 2^17.
  if (int64_t(kSliceTileM) * KP >= (int64_t(1) << 31) || k >= 131072) {
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
  // The coarse planes of the decision (INT8 tile geometry) live in the
  // first digit planes until the digits overwrite them.
  const int64_t MPc = AlignUp(m, 256), NPc = AlignUp(n, 256), KPc = AlignUp(k, 128);
  const int planes = slices > 0 ? std::min(slices, kMaxSlices) : kMaxSlices;
  const size_t o_da = AlignUp(StatsBytes(m, n), 256);
  const size_t da_bytes = std::max(sizeof(uint16_t) * size_t(planes) * MP * KP, size_t(MPc * KPc));
  const size_t o_db = o_da + AlignUp(da_bytes, 256);
  const size_t db_bytes = std::max(sizeof(uint16_t) * size_t(planes) * NP * KP, size_t(NPc * KPc));
  char* ws = static_cast<char*>(GetWorkspace(o_db + db_bytes, stream));
  const EmuStats st = CarveStats(ws, m, n);
  uint16_t* da = reinterpret_cast<uint16_t*>(ws + o_da);
  uint16_t* db = reinterpret_cast<uint16_t*>(ws + o_db);

  // 1. Statistics, scale exponents -(E + 1) (|x 2^-(E+1)| < 1/2), and the
  //    slice count.
  Mark(kPhaseStats, stream);
  Decide(1, slices, op, st, -1, reinterpret_cast<uint8_t*>(da), reinterpret_cast<uint8_t*>(db),
         MPc, NPc, KPc, stream);

  // 2. The digits.
  Mark(kPhaseSlice, stream);
  auto digits = [&](const double* x, int64_t rs, int64_t cs, int64_t R, int64_t RP,
                    const int* s, uint16_t* y) {
    if (rs == 1)
      hipLaunchKernelGGL(DigitColsKernel, dim3(unsigned(KP / 128), unsigned(RP / 32)), dim3(256),
                         0, stream, x, cs, R, k, RP, KP, s, st.ctl, y);
    else
      hipLaunchKernelGGL(DigitRowsKernel, dim3(unsigned(KP / 128), unsigned(RP / 16)), dim3(256),
                         0, stream, x, rs, R, k, RP, KP, s, st.ctl, y);
    HWY_HIP_CHECK(hipGetLastError());
  };
  digits(op.a, op.a_rs, op.a_cs, m, MP, st.sa, da);
  digits(op.b, op.b_cs, op.b_rs, n, NP, st.sb, db);

  // 3. All digit pairs and the epilogue in one launch, or the native kernel.
  SliceGemmArgs g{};
  g.a = da;
  g.b = db;
  g.a_plane = MP * KP;
  g.b_plane = NP * KP;
  g.kp = KP;
  g.ctl = st.ctl;
  g.sa = st.sa;
  g.sb = st.sb;
  g.alpha = alpha;
  g.beta = beta;
  g.c = c;
  g.ldc = ldc;
  g.m = m;
  g.n = n;
  g.tiles_m = int(MP / kSliceTileM);
  g.tiles_n = int(NP / kSliceTileN);
  Mark(kPhaseModGemm, stream);
  hipLaunchKernelGGL((SliceGemmKernel<kSliceTileM, kSliceTileN, 4, 2, 3>),
                     dim3(unsigned(g.tiles_m * g.tiles_n)), dim3(512), 0, stream, g);
  HWY_HIP_CHECK(hipGetLastError());
  GemmDeviceGated<double>(&st.ctl->dec.count, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c,
                          ldc, stream);
  Mark(kPhaseReconstruct, stream);  // the epilogue is fused into the GEMM
  Mark(kNumPhases, stream);
  if (used) *used = ReportCount(st, stream);
}

}  // namespace ozaki
}  // namespace hwy_amd

using namespace hwy_amd;

extern "C" {

HWYBLAS_AMD_API int hwyblas_dgemm_emulated_bf16_dev(char transa, char transb, int m, int n,
                                                    int k, double alpha, const double* a,
                                                    int lda, const double* b, int ldb,
                                                    double beta, double* c, int ldc, int slices,
                                                    void* stream, int* used_slices) {
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
  if (used_slices) *used_slices = 0;
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (m == 0 || n == 0) return 0;
  if ((k == 0 || alpha == 0.0) && beta == 1.0) return 0;
  if (k == 0 || alpha == 0.0) {
    ScaleDevice<double>(m, n, beta, c, ldc, s);
    return 0;
  }
  ozaki::EmulatedDgemmBf16(oa == Op::kTrans, ob == Op::kTrans, m, n, k, alpha, a, lda, b, ldb,
                           beta, c, ldc, slices, s, used_slices);
  return 0;
}

}  // extern "C"
