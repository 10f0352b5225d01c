// Householder QR on the GPU: dgeqrf_ (hwyBLAS src/lapack/geqrf.cc:66-150,
// include/hwyblas/lapack.h:1273-1283) and hwyblas_dgeqrf_dev.
//
// The factorization is LAPACK's (the same reflectors, signs and storage):
// column i's reflector H(i) = I - tau_i v v^T with v(i) = 1 (implicit),
// v(i+1:m) stored below the diagonal and R on and above it, beta =
// -sign(alpha) |(alpha, x)| (xLARFG, src/lapack/householder.cc:44-104). The
// blocking is the GPU's own:
//  - panels of kPanel = 64 columns; a panel (mp x 64, mp = m - j0) is
//    factored column by column by workgroups of 256 rows each, with ONE
//    reduction per column: each workgroup publishes, for column c, the
//    scale-safe norm partial of x = P(c+1:, c) and the dots of the same
//    (unscaled, exponent-normalized) x with the later columns; then every
//    workgroup forms, identically, xLARFG's beta and tau and
//    w = v^T P = P(c, :) + scale x^T P from them, stores its rows of v,
//    updates its rows, and publishes column c + 1's partials. The panel is
//    one cooperative launch (QrPanelPersistentKernel: the rows in
//    registers, one grid barrier per column) when all its workgroups fit on
//    the device, else one launch per column (QrColumnKernel), bit for bit
//    the same. Cross-lane sums run on gfx950's lane swaps and DPP, not on
//    the LDS crossbar (one wave per SIMD here: every latency is exposed);
//  - a panel's T (I - V T V^T = H(j0) ... H(j0+63), xLARFT's forward
//    columnwise T) from G = V^T V (an MFMA GEMM): the recurrence
//    T(0:i, i) = -tau_i T(0:i, 0:i) G(0:i, i), T(i, i) = tau_i on 16 x 16
//    diagonal blocks, merged pairwise as below;
//  - blocks of four panels (256 columns): each panel's columns first get
//    the block's earlier reflectors, then the panel is factored and the
//    block's T grows as T = [T_old, -T_old (V_old^T V_p) T_p; 0, T_p];
//  - the trailing columns updated as A2 := (I - V T^T V^T) A2 = H^T A2
//    (xLARFB, side L, trans T) by three MFMA GEMMs: W = V^T A2,
//    W2 = T^T W, A2 -= V W2, with k = 256 for a whole block. The products
//    with V^T have at most 256 rows and the panel's height as k: they
//    split k (GemmSplitKDevice) to fill the device.
// Norms are computed scale-safely: each workgroup sums (x 2^-e)^2 and
// (x 2^-e) P(:, cc) with e the exponent of its largest |x|, and the sums are
// combined at the largest e.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "common.h"
#include "grid_sync.h"

namespace hwy_amd {
namespace qr {
namespace {

constexpr int kPanel = 64;   // columns per panel
constexpr int kQrInternalError = -1000;  // dgeqrf_'s info when a panel barrier timed out
constexpr int kRows = 256;   // panel rows per workgroup (one per thread)
constexpr int kWaves = kRows / 64;

// One slot of a column's records, written by every workgroup before a
// barrier and read by every workgroup after it.
struct Slot {
  double* part;  // [G][2]: {e, sum of (x 2^-e)^2}, e the exponent of the workgroup's max |x|
  double* dots;  // [G][kPanel]: sum over the workgroup's rows of (x 2^-e) P(i, cc), cc > k
  double* row;   // [kPanel]: row k of the panel, P(k, cc) for cc >= k
};

struct PanelArgs {
  double* p;     // the panel's top-left element, A(j0, j0)
  int64_t lda;
  int mp, ib;    // panel rows and columns
  int groups;    // G: workgroups per column step (mp / kRows, rounded up)
  double* tau;   // tau + j0
  double* rec;   // two record slots, by column parity
};

// Doubles of one record slot for G workgroups.
__host__ __device__ constexpr int64_t SlotDoubles(int64_t groups) {
  return groups * (2 + kPanel) + kPanel;
}

__device__ __forceinline__ Slot SlotOf(const PanelArgs& g, int k) {
  double* base = g.rec + (k & 1) * SlotDoubles(g.groups);
  return Slot{base, base + 2 * int64_t(g.groups), base + int64_t(g.groups) * (2 + kPanel)};
}

// Cross-lane exchanges without the LDS crossbar (gfx950). Lanes are paired
// across halves of groups of 2O lanes: O = 32 and 16 by v_permlane32_swap /
// v_permlane16_swap, O = 8, 4, 2, 1 within rows of 16 by DPP mirrors (lane
// 15 - l, 7 - l within 8, 3 - l within 4, and l ^ 1), so lane l's partner
// always has the other value of lane bit O.
__device__ __forceinline__ unsigned Lo32(double v) {
  return unsigned(__double_as_longlong(v));
}
__device__ __forceinline__ unsigned Hi32(double v) {
  return unsigned(__double_as_longlong(v) >> 32);
}
__device__ __forceinline__ double Join(unsigned lo, unsigned hi) {
  return __longlong_as_double(
      static_cast<long long>((static_cast<unsigned long long>(hi) << 32) | lo));
}

// Lanes below O in their group of 2O (lane bit O clear) get a + their
// partner's a, the others b + their partner's b: the swap moves the upper
// lanes' a and the lower lanes' b to the partner, so the two swapped
// registers hold exactly the pair to add.
template <int O>
__device__ __forceinline__ double SwapSum(double a, double b) {
  static_assert(O == 32 || O == 16, "swap width");
  if constexpr (O == 32) {
    const auto lo = __builtin_amdgcn_permlane32_swap(Lo32(a), Lo32(b), false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(Hi32(a), Hi32(b), false, false);
    return Join(lo[0], hi[0]) + Join(lo[1], hi[1]);
  } else {
    const auto lo = __builtin_amdgcn_permlane16_swap(Lo32(a), Lo32(b), false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(Hi32(a), Hi32(b), false, false);
    return Join(lo[0], hi[0]) + Join(lo[1], hi[1]);
  }
}

// The partner's v for O = 8, 4, 2, 1 (DPP row_mirror, row_half_mirror,
// quad_perm [3, 2, 1, 0], quad_perm [1, 0, 3, 2]).
template <int O>
__device__ __forceinline__ double Partner(double v) {
  constexpr int kCtrl = O == 8 ? 0x140 : O == 4 ? 0x141 : O == 2 ? 0x1B : 0xB1;
  static_assert(O == 8 || O == 4 || O == 2 || O == 1, "DPP width");
  const int lo = __builtin_amdgcn_update_dpp(0, int(Lo32(v)), kCtrl, 0xf, 0xf, false);
  const int hi = __builtin_amdgcn_update_dpp(0, int(Hi32(v)), kCtrl, 0xf, 0xf, false);
  return Join(unsigned(lo), unsigned(hi));
}

// Sum or maximum over the wave, in every lane.
template <bool kMax>
__device__ __forceinline__ double WaveAll(double v) {
  auto op = [](double x, double y) { return kMax ? fmax(x, y) : x + y; };
  if constexpr (kMax) {
    const auto lo = __builtin_amdgcn_permlane32_swap(Lo32(v), Lo32(v), false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(Hi32(v), Hi32(v), false, false);
    v = op(Join(lo[0], hi[0]), Join(lo[1], hi[1]));
    const auto lo16 = __builtin_amdgcn_permlane16_swap(Lo32(v), Lo32(v), false, false);
    const auto hi16 = __builtin_amdgcn_permlane16_swap(Hi32(v), Hi32(v), false, false);
    v = op(Join(lo16[0], hi16[0]), Join(lo16[1], hi16[1]));
  } else {
    v = SwapSum<32>(v, v);
    v = SwapSum<16>(v, v);
  }
  v = op(v, Partner<8>(v));
  v = op(v, Partner<4>(v));
  v = op(v, Partner<2>(v));
  v = op(v, Partner<1>(v));
  return v;
}

__device__ __forceinline__ double WaveSum(double v) { return WaveAll<false>(v); }
__device__ __forceinline__ double WaveMax(double v) { return WaveAll<true>(v); }

__device__ double Lapy2(double x, double y) {
  if (isnan(x)) return x;
  if (isnan(y)) return y;
  const double w = fmax(fabs(x), fabs(y)), z = fmin(fabs(x), fabs(y));
  if (z == 0.0 || w > 1.7976931348623157e308) return w;
  return w * sqrt(1.0 + (z / w) * (z / w));
}

// xLARFG's beta, tau and the scale of x from alpha and ||x||_2
// (src/lapack/householder.cc:44-104): beta = -sign(alpha) |(alpha, x)|,
// tau = (beta - alpha) / beta, x scaled by 1 / (alpha - beta), with the
// rescaling by 1/safmin while |beta| < safmin.
__device__ void Larfg(double alpha, double xnorm, double& tau, double& scal, double& beta) {
  tau = 0.0;
  scal = 1.0;
  beta = alpha;
  if (xnorm == 0.0) return;
  const double safmin = 2.2250738585072014e-308 / 1.1102230246251565e-16;
  double a = alpha, xn = xnorm;
  double b = -copysign(Lapy2(a, xn), a);
  double up = 1.0, down = 1.0;
  int knt = 0;
  while (fabs(b) < safmin && knt < 20) {
    ++knt;
    a *= 1.0 / safmin;
    xn *= 1.0 / safmin;
    b *= 1.0 / safmin;
    up *= 1.0 / safmin;
    down *= safmin;
  }
  if (knt > 0) b = -copysign(Lapy2(a, xn), a);
  tau = (b - a) / b;
  scal = 1.0 / (a - b);
  scal *= up;  // x is scaled by up first in xLARFG; up and scal commute
  beta = b * down;
}

// The wave's sums of 64 values per lane, scattered: on return lane l holds
// the sum over the wave of value l. x[0..31] on entry are a lane's values
// after the first halving step (offset 32) has been taken by the caller;
// each later step keeps the half of the values whose index bit matches the
// lane's bit and adds the partner's copy of that half (63 exchanges in all
// instead of 6 per value).
template <int O>
__device__ __forceinline__ void ReduceScatterStep(double (&x)[32], int lane) {
  if constexpr (O == 16) {
#pragma unroll
    for (int j = 0; j < O; ++j) x[j] = SwapSum<16>(x[j], x[j + O]);
  } else {
    const bool upper = (lane & O) != 0;
#pragma unroll
    for (int j = 0; j < O; ++j) {
      const double keep = upper ? x[j + O] : x[j];
      const double send = upper ? x[j] : x[j + O];
      x[j] = keep + Partner<O>(send);
    }
  }
  if constexpr (O > 1) ReduceScatterStep<O / 2>(x, lane);
}

__device__ __forceinline__ double WaveReduceScatter32(double (&x)[32]) {
  ReduceScatterStep<16>(x, threadIdx.x & 63);
  return x[0];
}

// Column k's records from this workgroup: with x = P(i, k) on its rows
// i > k and e the exponent of its largest |x| (-100000 when all are zero),
// the norm partial {e, sum (x 2^-e)^2} and, for each later column cc,
// sum (x 2^-e) P(i, cc); the owner of row k publishes that row. `xk` is this
// thread's P(i, k) and `pv` its row of the panel. The 64 per-column sums of
// a wave (slot k holding the norm's) are one reduce-scatter; the waves are
// added in wave order through LDS.
__device__ void ColumnPartials(const PanelArgs& g, int k, int ib, double xk,
                               const double (&pv)[kPanel], Slot out) {
  __shared__ double red[kWaves];
  __shared__ double mx_s;
  __shared__ double wsum[kWaves][kPanel];  // [w][k]: the norm sum, [w][cc > k]: the dots
  __shared__ double srow[kPanel];          // row k, staged by its owner for a wide publish
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = blockIdx.x * kRows + threadIdx.x;
  const bool below = i > k && i < g.mp;
  const double x = below ? xk : 0.0;
  const double m = WaveMax(fabs(x));
  if (lane == 0) red[wave] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    double mm = 0.0;
    for (int w = 0; w < kWaves; ++w) mm = fmax(mm, red[w]);
    mx_s = mm;
  }
  __syncthreads();
  const double mx = mx_s;
  int e = -100000;
  if (mx > 0.0) frexp(mx, &e);
  const double y = mx > 0.0 ? ldexp(x, -e) : 0.0;
  // This lane's 64 values (y^2 in slot k, y P(i, cc) for the later columns,
  // zeros elsewhere), folded at once into the first halving step.
  auto value = [&](int cc) {
    return cc == k ? y * y : ((cc > k && cc < ib && below) ? y * pv[cc] : 0.0);
  };
  double h[32];
#pragma unroll
  for (int j = 0; j < 32; ++j) h[j] = SwapSum<32>(value(j), value(j + 32));
  wsum[wave][lane] = WaveReduceScatter32(h);
  if (i == k) {
#pragma unroll
    for (int cc = 0; cc < kPanel; ++cc) srow[cc] = pv[cc];
  }
  __syncthreads();
  const int t = threadIdx.x;
  if (t >= k && t < ib) {
    double d = wsum[0][t];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) d += wsum[w][t];
    if (t == k) {
      PutRecord(out.part + 2 * blockIdx.x, double(e));
      PutRecord(out.part + 2 * blockIdx.x + 1, d);
    } else {
      PutRecord(out.dots + int64_t(blockIdx.x) * kPanel + t, d);
    }
    // The owner of row k is thread k of workgroup 0.
    if (blockIdx.x == 0) PutRecord(out.row + t, srow[t]);
  }
}

constexpr int kMaxLdsGroups = 1024;  // workgroup norm partials kept in LDS
constexpr int kDotSplit = 4;         // threads per column in the sums of the dot partials
constexpr int kDotBatch = 8;         // dot partials in flight per thread

// Column c from its records. Every workgroup forms, identically: ||x|| from
// the norm partials (the largest exponent emax, then the sums rescaled to
// it, one lane per partial, reduced over a wave), alpha = P(c, c) from the
// row record, xLARFG's tau, beta and scale, and for each later column cc
// w(cc) = v^T P(:, cc) = P(c, cc) + scale 2^emax sum_q 2^(e_q - emax) d_q(cc)
// (the dots of the unscaled x, so that the norm and the dots need one
// reduction between them; kDotSplit threads per column each sum every
// kDotSplit-th partial, the sums then added in order). Then its rows:
// v = scale x stored below the diagonal (beta on it), P(i, cc) -= tau v(i)
// w(cc), and column c + 1's records. Every record load of the step is
// issued at its start (one memory round trip) when G <= kDotSplit *
// kDotBatch. `cur` is this thread's P(i, c) on entry, P(i, c + 1) on return.
__device__ void ColumnStep(const PanelArgs& g, int c, int ib, double (&pv)[kPanel], double& cur) {
  __shared__ double coef[4];  // tau, scale, scale 2^emax, beta
  __shared__ double w[kPanel];
  __shared__ double wpart[kDotSplit][kPanel];
  __shared__ int se[kMaxLdsGroups];
  __shared__ double ss[kMaxLdsGroups];
  __shared__ int emax_s;
  const Slot in = SlotOf(g, c);
  const int G = g.groups;
  const int t = threadIdx.x, lane = t & 63;
  // Every load first: the norm partials (to LDS), this thread's row-record
  // entry, and its first kDotBatch dot partials.
  const double rowv = (t > c && t < ib) ? GetRecord(in.row + t) : 0.0;
  const double alpha = t == 0 ? GetRecord(in.row + c) : 0.0;
  const int cci = t / kDotSplit, part = t % kDotSplit;
  const int cc = c + 1 + cci;
  const bool summer = cc < ib;
  double dv[kDotBatch];
#pragma unroll
  for (int u = 0; u < kDotBatch; ++u) {
    const int q = part + kDotSplit * u;
    dv[u] = (summer && q < G) ? GetRecord(in.dots + int64_t(q) * kPanel + cc) : 0.0;
  }
  for (int q = t; q < G && q < kMaxLdsGroups; q += kRows) {
    se[q] = int(GetRecord(in.part + 2 * q));
    ss[q] = GetRecord(in.part + 2 * q + 1);
  }
  __syncthreads();
  auto exp_of = [&](int q) { return q < kMaxLdsGroups ? se[q] : int(GetRecord(in.part + 2 * q)); };
  if (t < 64) {
    int emax = -100000;
    for (int q = lane; q < G; q += 64) emax = max(emax, exp_of(q));
    emax = int(WaveMax(double(emax)));
    double sum = 0.0;
    if (emax != -100000) {
      for (int q = lane; q < G; q += 64) {
        const int e = exp_of(q);
        const double sq = q < kMaxLdsGroups ? ss[q] : GetRecord(in.part + 2 * q + 1);
        if (e != -100000) sum += ldexp(sq, 2 * (e - emax));
      }
    }
    sum = WaveSum(sum);
    if (t == 0) {
      const double xnorm = emax == -100000 ? 0.0 : ldexp(sqrt(sum), emax);
      double tau, scal, beta;
      Larfg(alpha, xnorm, tau, scal, beta);
      coef[0] = tau;
      coef[1] = scal;
      coef[2] = ldexp(scal, emax);
      coef[3] = beta;
      emax_s = emax;
      if (blockIdx.x == 0) g.tau[c] = tau;
    }
  }
  __syncthreads();
  const double tau = coef[0], scal = coef[1], f = coef[2];
  const int emax = emax_s;
  if (summer) {
    double d = 0.0;
#pragma unroll
    for (int u = 0; u < kDotBatch; ++u) {
      const int q = part + kDotSplit * u;
      if (q < G) {
        const int e = exp_of(q);
        if (e != -100000) d += ldexp(dv[u], e - emax);
      }
    }
    for (int q0 = part + kDotSplit * kDotBatch; q0 < G; q0 += kDotSplit * kDotBatch) {
#pragma unroll
      for (int u = 0; u < kDotBatch; ++u) {
        const int q = q0 + kDotSplit * u;
        dv[u] = q < G ? GetRecord(in.dots + int64_t(q) * kPanel + cc) : 0.0;
      }
#pragma unroll
      for (int u = 0; u < kDotBatch; ++u) {
        const int q = q0 + kDotSplit * u;
        if (q < G) {
          const int e = exp_of(q);
          if (e != -100000) d += ldexp(dv[u], e - emax);
        }
      }
    }
    wpart[part][cc] = d;
  }
  __syncthreads();
  if (t > c && t < ib) {
    double d = wpart[0][t];
#pragma unroll
    for (int p = 1; p < kDotSplit; ++p) d += wpart[p][t];
    w[t] = rowv + f * d;
  } else if (t < kPanel) {
    w[t] = 0.0;
  }
  __syncthreads();
  const int i = blockIdx.x * kRows + threadIdx.x;
  const bool inr = i < g.mp;
  double v = 0.0;
  if (i == c) {
    v = 1.0;
    g.p[i + int64_t(c) * g.lda] = coef[3];
  } else if (i > c && inr) {
    v = cur * scal;
    g.p[i + int64_t(c) * g.lda] = v;
  }
  // Every entry of the row, without per-column branches (w is zero outside
  // (c, ib), so the others keep their values; they are not read again),
  // so that the reads of w are issued together.
  double next = 0.0;
  if (i >= c && inr) {
#pragma unroll
    for (int k = 0; k < kPanel; ++k) {
      const double nv = pv[k] - tau * v * w[k];
      pv[k] = nv;
      next = k == c + 1 ? nv : next;
    }
  }
  if (c + 1 < ib) ColumnPartials(g, c + 1, ib, next, pv, SlotOf(g, c + 1));
  cur = next;
}

// This thread's row of the panel (zeros outside it).
__device__ __forceinline__ void LoadRow(const PanelArgs& g, double (&pv)[kPanel]) {
  const int i = blockIdx.x * kRows + threadIdx.x;
#pragma unroll
  for (int cc = 0; cc < kPanel; ++cc)
    pv[cc] = (i < g.mp && cc < g.ib) ? g.p[i + int64_t(cc) * g.lda] : 0.0;
}

// Column 0's records (the path of one launch per column).
__global__ void __launch_bounds__(kRows) QrFirstColumnKernel(PanelArgs g) {
  double pv[kPanel];
  LoadRow(g, pv);
  ColumnPartials(g, 0, g.ib, pv[0], pv, SlotOf(g, 0));
}

// Column c (the path of one launch per column): the rows loaded, the step,
// the updated entries stored.
__global__ void __launch_bounds__(kRows) QrColumnKernel(PanelArgs g, int c) {
  double pv[kPanel];
  LoadRow(g, pv);
  double cur = 0.0;
#pragma unroll
  for (int cc = 0; cc < kPanel; ++cc)
    if (cc == c) cur = pv[cc];
  ColumnStep(g, c, g.ib, pv, cur);
  const int i = blockIdx.x * kRows + threadIdx.x;
  if (i >= c && i < g.mp) {
#pragma unroll
    for (int cc = 0; cc < kPanel; ++cc)
      if (cc > c && cc < g.ib) g.p[i + int64_t(cc) * g.lda] = pv[cc];
  }
}

// The whole panel in one cooperative launch: each thread keeps its row of
// the panel in registers and the kernel boundary between column steps
// becomes one grid-wide barrier (GridSync); the steps are those of
// QrColumnKernel, so the bits are the same. Records alternate between two
// slots: a slot is rewritten only after the barrier that follows its last
// read. R's entries (rows < ib, right of the diagonal) are stored at the
// end. After a timed-out wait every workgroup returns and workgroup 0 sets
// the panel's tau to NaN.
struct PersistArgs {
  PanelArgs g;
  unsigned* bar;    // arrivals (bar_base at launch: the count runs on over panels)
  unsigned* failed;
  unsigned bar_base;
};

__global__ void __launch_bounds__(kRows) QrPanelPersistentKernel(PersistArgs a) {
  const PanelArgs& g = a.g;
  double pv[kPanel];
  LoadRow(g, pv);
  ColumnPartials(g, 0, g.ib, pv[0], pv, SlotOf(g, 0));
  double cur = pv[0];
  unsigned arrivals = a.bar_base;
  for (int c = 0; c < g.ib; ++c) {
    // An opaque copy of the column count: the per-column tests are then
    // formed where they are used, not hoisted out of the loop into
    // (spilled) scalar registers.
    int ib = g.ib;
    asm volatile("" : "+s"(ib));
    arrivals += unsigned(g.groups);
    if (!GridSync(a.bar, arrivals, a.failed)) break;
    ColumnStep(g, c, ib, pv, cur);
  }
  if (__hip_atomic_load(a.failed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
    if (blockIdx.x == 0 && int(threadIdx.x) < g.ib) g.tau[threadIdx.x] = __builtin_nan("");
    return;
  }
  // R right of the diagonal: rows i < ib, final since column i's update.
  const int i = blockIdx.x * kRows + threadIdx.x;
  int ib = g.ib;
  asm volatile("" : "+s"(ib));
  if (i < ib) {
#pragma unroll
    for (int cc = 0; cc < kPanel; ++cc)
      if (cc > i && cc < ib) g.p[i + int64_t(cc) * g.lda] = pv[cc];
  }
}

// Columns of V (rows x ib, leading dimension ldv): a panel's reflectors
// with the unit diagonal and zeros above written out, the panel's first row
// at row d of V (rows above it zero).
__global__ void ExplicitVKernel(const double* p, int64_t lda, int rows, int ib, int d, double* v,
                                int64_t ldv) {
  const int64_t total = int64_t(rows) * ib;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int i = int(t % rows), j = int(t / rows), r = i - d;
    v[i + int64_t(j) * ldv] = r == j ? 1.0 : (r > j ? p[r + int64_t(j) * lda] : 0.0);
  }
}

// T (ib x ib, upper, leading dimension ldt) of the block reflector
// I - V T V^T = H(0) ... H(ib-1) from G = V^T V and tau (xLARFT, forward,
// columnwise; zeros below the diagonal and outside ib), by blocks: the four
// 16 x 16 diagonal blocks by the column recurrence T(0:i, i) = -tau_i
// T(0:i, 0:i) G(0:i, i), T(i, i) = tau_i (sixteen threads each, one row per
// thread), then merged pairwise as the T of a product of two block
// reflectors, T = [T1, -T1 G12 T2; 0, T2], first into 32 x 32 blocks and
// then into the whole. G, T and the products' scratch live in LDS.
constexpr int kLarftThreads = 256;
constexpr int kLarftBlock = 16;

// out(r0.., c0..) = -T(r0.., r0..) (G(r0.., c0..) T(c0.., c0..)), blocks of
// order h at (r0, c0) for `pairs` pairs side by side (pair q at r0 + 2hq).
__device__ void LarftMerge(double (*sg)[kPanel + 1], double (*st)[kPanel + 1],
                           double (*sw)[kPanel + 1], int h, int pairs) {
  const int tid = threadIdx.x, outs = pairs * h * h;
  for (int o = tid; o < outs; o += kLarftThreads) {
    const int q = o / (h * h), e = o % (h * h), r = e % h, c = e / h;
    const int r0 = 2 * h * q, c0 = r0 + h;
    double s = 0.0;
    for (int u = 0; u < h; ++u) s += sg[r0 + r][c0 + u] * st[c0 + u][c0 + c];
    sw[r0 + r][c0 + c] = s;
  }
  __syncthreads();
  for (int o = tid; o < outs; o += kLarftThreads) {
    const int q = o / (h * h), e = o % (h * h), r = e % h, c = e / h;
    const int r0 = 2 * h * q, c0 = r0 + h;
    double s = 0.0;
    for (int u = r; u < h; ++u) s += st[r0 + r][r0 + u] * sw[r0 + u][c0 + c];
    st[r0 + r][c0 + c] = -s;
  }
  __syncthreads();
}

__global__ void __launch_bounds__(kLarftThreads) LarftKernel(const double* gm, int ib,
                                                             const double* tau, double* t,
                                                             int ldt) {
  __shared__ double sg[kPanel][kPanel + 1];  // sg[q][i] = G(q, i)
  __shared__ double st[kPanel][kPanel + 1];  // st[r][q] = T(r, q)
  __shared__ double sw[kPanel][kPanel + 1];  // products
  __shared__ double stau[kPanel];
  const int tid = threadIdx.x;
  for (int idx = tid; idx < kPanel * kPanel; idx += kLarftThreads) {
    const int r = idx % kPanel, c = idx / kPanel;
    sg[r][c] = (r < ib && c < ib) ? gm[r + int64_t(c) * ib] : 0.0;
    st[r][c] = 0.0;
  }
  if (tid < kPanel) stau[tid] = tid < ib ? tau[tid] : 0.0;
  __syncthreads();
  // The diagonal blocks: thread R < 64 owns row R of its block b = R / 16.
  const int b = tid / kLarftBlock, r = tid % kLarftBlock, R = tid;
  for (int i = 0; i < kLarftBlock; ++i) {
    if (tid < kPanel) {
      const int I = b * kLarftBlock + i;
      if (r < i) {
        double s = 0.0;
        for (int q = r; q < i; ++q) s += st[R][b * kLarftBlock + q] * sg[b * kLarftBlock + q][I];
        st[R][I] = -stau[I] * s;
      }
      if (r == i) st[R][I] = stau[I];
    }
    __syncthreads();
  }
  LarftMerge(sg, st, sw, kLarftBlock, 2);      // 16 -> 32
  LarftMerge(sg, st, sw, 2 * kLarftBlock, 1);  // 32 -> 64
  for (int idx = tid; idx < kPanel * kPanel; idx += kLarftThreads) {
    const int rr = idx % kPanel, c = idx / kPanel;
    t[rr + int64_t(c) * ldt] = st[rr][c];
  }
}

__global__ void ZeroKernel(double* p, int64_t n) {
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < n;
       t += int64_t(gridDim.x) * blockDim.x)
    p[t] = 0.0;
}

constexpr int kBlock = 4 * kPanel;  // columns per block reflector (four panels)

struct Work {
  unsigned* sync; // [0]: barrier arrivals, [32]: failed (64 words)
  double* v;      // m x VCols(n)
  double* t;      // kBlock x kBlock
  double* gm;     // kBlock x kBlock: V_p^T V_p, V_old^T V_p and a product
  double* w;      // kBlock x n
  double* w2;     // kBlock x n
  double* rec;    // the panel's two record slots
  double* split;  // split-k partials of V^T V and V^T A2
  int64_t split_cap;
};

int64_t SplitCap(int64_t n) { return 8 * int64_t(kBlock) * std::max<int64_t>(n, 4096); }

// Columns of the explicit V: a block's width at most.
int64_t VCols(int64_t n) { return std::min<int64_t>(kBlock, std::max<int64_t>(n, 1)); }

size_t WorkBytes(int64_t m, int64_t n) {
  const int64_t blocks = (m + kRows - 1) / kRows;
  return 8 * (32 + size_t(m) * VCols(n) + 2 * kBlock * kBlock + 2 * size_t(kBlock) * n +
              2 * size_t(SlotDoubles(blocks)) + size_t(SplitCap(n))) + 256;
}

Work Carve(void* base, int64_t m, int64_t n) {
  const int64_t blocks = (m + kRows - 1) / kRows;
  double* p = static_cast<double*>(base);
  Work w;
  w.sync = reinterpret_cast<unsigned*>(p);
  p += 32;
  w.v = p;
  p += m * VCols(n);
  w.t = p;
  p += kBlock * kBlock;
  w.gm = p;
  p += kBlock * kBlock;
  w.w = p;
  p += int64_t(kBlock) * n;
  w.w2 = p;
  p += int64_t(kBlock) * n;
  w.rec = p;
  p += 2 * SlotDoubles(blocks);
  w.split = p;
  w.split_cap = SplitCap(n);
  return w;
}

}  // namespace

__global__ void FillWords(unsigned* p, int n, unsigned v) {
  for (int t = threadIdx.x; t < n; t += blockDim.x) p[t] = v;
}

// How many workgroups of QrPanelPersistentKernel can be resident at once on
// the current device (0: one launch per column step; also with
// HWYBLAS_AMD_QR_PANEL=steps).
int PersistentPanelCapacity() {
  const char* e = getenv("HWYBLAS_AMD_QR_PANEL");
  if (e && std::string(e) == "steps") return 0;
  int dev = 0, cus = 0, coop = 0, per_cu = 0;
  HWY_HIP_CHECK(hipGetDevice(&dev));
  HWY_HIP_CHECK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, dev));
  HWY_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  HWY_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(
      &per_cu, reinterpret_cast<const void*>(QrPanelPersistentKernel), kRows, 0));
  return coop ? per_cu * cus : 0;
}

// A = QR in place (m x n, lda), tau (device) of min(m, n); asynchronous.
// With `failed` (host memory), the cooperative panel's failure flag is
// copied there on the stream (nonzero: a panel barrier timed out, A and tau
// are undefined and the panel's tau NaN).
void Geqrf(int64_t m, int64_t n, double* a, int64_t lda, double* tau, hipStream_t s,
           unsigned* failed = nullptr) {
  const int64_t k = std::min(m, n);
  if (k == 0) return;
  const Work w = Carve(Staging(kSlotPackA, WorkBytes(m, n), s), m, n);
  // The cooperative panel kernel serves every panel whose workgroups are all
  // resident at once; the barrier's arrival count runs on over the panels.
  int cap = PersistentPanelCapacity();
  const bool coop_possible = cap > 0;
  unsigned arrivals = 0;
  if (cap > 0) {
    hipLaunchKernelGGL(FillWords, dim3(1), dim3(64), 0, s, w.sync, 64, 0u);
    // Test hook: HWYBLAS_AMD_QR_FAIL_BARRIER=1 raises the failure flag first,
    // so that every bounded wait gives up at its first poll and every
    // workgroup returns: the path a timed-out barrier takes, without waiting
    // for one.
    const char* fail = getenv("HWYBLAS_AMD_QR_FAIL_BARRIER");
    if (fail && *fail == '1')
      hipLaunchKernelGGL(FillWords, dim3(1), dim3(64), 0, s, w.sync + 32, 1, 1u);
  }
  // Panel (column j0, ib columns, mp = m - j0 rows): one cooperative launch,
  // or the column launches.
  auto factor_panel = [&](int64_t j0, int ib) {
    const int mp = int(m - j0);
    const int blocks = (mp + kRows - 1) / kRows;
    PanelArgs g{a + j0 + j0 * lda, lda, mp, ib, blocks, tau + j0, w.rec};
    bool coop = false;
    if (blocks <= cap) {
      PersistArgs pa{g, w.sync, w.sync + 32, arrivals};
      void* args[] = {&pa};
      // A cooperative launch the runtime refuses (another process holding
      // the device's cooperative queue, for one) leaves this panel and the
      // later ones to the column launches, which give the same bits.
      const hipError_t e = hipLaunchCooperativeKernel(
          reinterpret_cast<const void*>(QrPanelPersistentKernel), dim3(blocks), dim3(kRows), args,
          0, s);
      if (e == hipSuccess) {
        coop = true;
        arrivals += unsigned(ib) * unsigned(blocks);
      } else {
        (void)hipGetLastError();
        cap = 0;
      }
    }
    if (!coop) {
      hipLaunchKernelGGL(QrFirstColumnKernel, dim3(blocks), dim3(kRows), 0, s, g);
      for (int c = 0; c < ib; ++c)
        hipLaunchKernelGGL(QrColumnKernel, dim3(blocks), dim3(kRows), 0, s, g, c);
    }
    HWY_HIP_CHECK(hipGetLastError());
  };
  // Columns of the explicit V (leading dimension mp) from the panel at
  // column j0 + d (its first row at row d of V).
  auto explicit_v = [&](int64_t j0, int d, int ib, int mp, double* v) {
    const int64_t vt = int64_t(mp) * ib;
    hipLaunchKernelGGL(ExplicitVKernel, dim3(unsigned(std::min<int64_t>((vt + 255) / 256, 2048))),
                       dim3(256), 0, s, a + (j0 + d) + (j0 + d) * lda, lda, mp, ib, d, v,
                       int64_t(mp));
  };
  double* gpp = w.gm;                          // V_p^T V_p (kPanel x kPanel)
  double* gop = w.gm + kPanel * kPanel;         // V_old^T V_p (up to kBlock x kPanel)
  double* tmp = gop + int64_t(kBlock) * kPanel;  // (V_old^T V_p) T_p
  // Blocks of kBlock columns (kBlock / kPanel panels). Panel p's columns
  // first get the block's earlier reflectors, H_old^T = I - V_old T_old^T
  // V_old^T; then the panel is factored and T grows as xLARFT's T of the
  // product: T = [T_old, -T_old (V_old^T V_p) T_p; 0, T_p]. The trailing
  // columns get H^T = I - V T^T V^T of the whole block, k = kBlock: a
  // pass over them per kBlock columns instead of per panel.
  for (int64_t j0 = 0; j0 < k; j0 += kBlock) {
    const int kb = int(std::min<int64_t>(kBlock, k - j0));
    const int mp = int(m - j0);
    const int64_t n2 = n - j0 - kb;
    double* v = w.v;
    double* t = w.t;
    if (kb <= kPanel && n2 <= 0) {  // one panel and nothing right of it
      factor_panel(j0, kb);
      continue;
    }
    hipLaunchKernelGGL(ZeroKernel, dim3(16), dim3(256), 0, s, t, int64_t(kBlock) * kBlock);
    for (int p0 = 0; p0 < kb; p0 += kPanel) {
      const int ib = std::min(kPanel, kb - p0);
      if (p0 > 0) {
        // H_old^T on this panel's columns.
        double* ap = a + j0 + (j0 + p0) * lda;
        GemmSplitKDevice(true, false, p0, ib, mp, 1.0, v, mp, ap, lda, 0.0, w.w, p0, w.split,
                         w.split_cap, s);
        GemmDevice<double>(true, false, p0, ib, p0, 1.0, t, kBlock, w.w, p0, 0.0, w.w2, p0, s);
        GemmDevice<double>(false, false, mp, ib, p0, -1.0, v, mp, w.w2, p0, 1.0, ap, lda, s);
      }
      factor_panel(j0 + p0, ib);
      if (p0 + ib >= kb && n2 <= 0) break;  // no later panel or column needs V and T
      double* vp = v + int64_t(mp) * p0;
      explicit_v(j0, p0, ib, mp, vp);
      const int mq = mp - p0;  // V_p's rows below its zeros
      GemmSplitKDevice(true, false, ib, ib, mq, 1.0, vp + p0, mp, vp + p0, mp, 0.0, gpp, ib,
                       w.split, w.split_cap, s);
      double* tp = t + p0 + int64_t(p0) * kBlock;
      hipLaunchKernelGGL(LarftKernel, dim3(1), dim3(kLarftThreads), 0, s, gpp, ib,
                         tau + j0 + p0, tp, kBlock);
      HWY_HIP_CHECK(hipGetLastError());
      if (p0 > 0) {
        GemmSplitKDevice(true, false, p0, ib, mq, 1.0, v + p0, mp, vp + p0, mp, 0.0, gop, p0,
                         w.split, w.split_cap, s);
        GemmDevice<double>(false, false, p0, ib, ib, 1.0, gop, p0, tp, kBlock, 0.0, tmp, p0, s);
        GemmDevice<double>(false, false, p0, ib, p0, -1.0, t, kBlock, tmp, p0, 0.0,
                           t + int64_t(p0) * kBlock, kBlock, s);
      }
    }
    if (n2 <= 0) continue;
    // H^T applied to the trailing columns.
    double* a3 = a + j0 + (j0 + kb) * lda;
    GemmSplitKDevice(true, false, kb, n2, mp, 1.0, v, mp, a3, lda, 0.0, w.w, kb, w.split,
                     w.split_cap, s);
    GemmDevice<double>(true, false, kb, n2, kb, 1.0, t, kBlock, w.w, kb, 0.0, w.w2, kb, s);
    GemmDevice<double>(false, false, mp, n2, kb, -1.0, v, mp, w.w2, kb, 1.0, a3, lda, s);
  }
  if (failed) {
    *failed = 0;
    if (coop_possible)
      HWY_HIP_CHECK(hipMemcpyAsync(failed, w.sync + 32, sizeof(unsigned), hipMemcpyDeviceToHost,
                                   s));
  }
}

}  // namespace qr
}  // namespace hwy_amd

using namespace hwy_amd;

namespace {

hipStream_t StreamOr(void* stream) {
  return stream ? static_cast<hipStream_t>(stream) : Context().stream;
}

}  // namespace

extern "C" {

HWYBLAS_AMD_API int hwyblas_dgeqrf_dev(int m, int n, double* a, int lda, double* tau,
                                       void* stream) {
  int info = 0;
  if (m < 0) info = -1;
  else if (n < 0) info = -2;
  else if (lda < std::max(1, m)) info = -4;
  if (info != 0) {
    Reject("DGEQRF", -info);
    return info;
  }
  qr::Geqrf(m, n, a, lda, tau, StreamOr(stream));
  return 0;
}

// The Fortran entry: A and tau host or device memory; lwork as xGEQRF's
// (>= max(1, n), or -1 for the query: work[0] = n * 32, the reference's
// optimal size, though the GPU path keeps its work arrays on the device).
HWYBLAS_AMD_API void dgeqrf_(const hwyblas_int* m, const hwyblas_int* n, double* a,
                             const hwyblas_int* lda, double* tau, double* work,
                             const hwyblas_int* lwork, hwyblas_int* info) {
  const int M = *m, N = *n, K = std::min(M, N);
  const bool query = *lwork == -1;
  int e = 0;
  if (M < 0) e = -1;
  else if (N < 0) e = -2;
  else if (*lda < std::max(1, M)) e = -4;
  else if (!query && (*lwork <= 0 || (M > 0 && *lwork < std::max(1, N)))) e = -7;
  *info = e;
  if (e != 0) {
    Reject("DGEQRF", -e);
    return;
  }
  if (query) {
    work[0] = K == 0 ? 1.0 : double(N) * 32.0;
    return;
  }
  if (K == 0) {
    work[0] = 1.0;
    return;
  }
  hipStream_t s = Context().stream;
  double* da = a;
  double* dtau = tau;
  const bool host_a = !IsDevicePointer(a), host_tau = !IsDevicePointer(tau);
  if (host_a) {
    da = static_cast<double*>(Staging(kSlotA, size_t(M) * N * 8));
    HWY_HIP_CHECK(hwy_amd::Memcpy2DAsync(da, size_t(M) * 8, a, size_t(*lda) * 8, size_t(M) * 8, N,
                                         hipMemcpyHostToDevice, s));
  }
  if (host_tau) dtau = static_cast<double*>(Staging(kSlotB, size_t(K) * 8));
  unsigned failed = 0;
  qr::Geqrf(M, N, da, host_a ? M : *lda, dtau, s, &failed);
  if (host_a)
    HWY_HIP_CHECK(hwy_amd::Memcpy2DAsync(a, size_t(*lda) * 8, da, size_t(M) * 8, size_t(M) * 8, N,
                                         hipMemcpyDeviceToHost, s));
  if (host_tau)
    HWY_HIP_CHECK(hipMemcpyAsync(tau, dtau, size_t(K) * 8, hipMemcpyDeviceToHost, s));
  HWY_HIP_CHECK(hipStreamSynchronize(s));
  work[0] = double(N) * 32.0;
  if (failed) {
    fprintf(stderr, "hwyblas_amd: dgeqrf's panel kernel timed out at a grid barrier\n");
    *info = qr::kQrInternalError;
    ReportInternal("DGEQRF", qr::kQrInternalError);
  }
}

}  // extern "C"
