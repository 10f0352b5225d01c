// Householder QR on the GPU: dgeqrf_ (hwyBLAS src/lapack/geqrf.cc:66-150,
// include/hwyblas/lapack.h:1273-1283) and hwyblas_dgeqrf_dev.
//
// The factorization is LAPACK's (the same reflectors, signs and storage):
// column i's reflector H(i) = I - tau_i v v^T with v(i) = 1 (implicit),
// v(i+1:m) stored below the diagonal and R on and above it, beta =
// -sign(alpha) |(alpha, x)| (xLARFG, src/lapack/householder.cc:44-104). The
// blocking is the GPU's own:
//  - panels of kPanel = 64 columns; a panel (mp x 64, mp = m - j0) is
//    factored column by column by workgroups of 256 rows each: in one
//    cooperative launch (QrPanelPersistentKernel, two grid barriers per
//    column) when all its workgroups fit on the device, else in two
//    launches per column: QrReflectKernel (the norm from the previous launch's
//    per-workgroup partial sums, beta and tau computed redundantly and
//    identically by every workgroup, its rows of v scaled, and one partial
//    dot product v^T P(:, c') per workgroup for the panel's later columns)
//    and QrUpdateKernel (the dots summed in workgroup order, the rank-1
//    update of its rows, and the next column's norm partials); each thread
//    issues all of its row's loads before the first sum;
//  - the block reflector's T (I - V T V^T = H(j0) ... H(j0+63), xLARFT's
//    forward columnwise T) from G = V^T V (an MFMA GEMM) and the recurrence
//    T(0:i, i) = -tau_i T(0:i, 0:i) G(0:i, i), T(i, i) = tau_i;
//  - the trailing columns updated as A2 := (I - V T^T V^T) A2 = H^T A2
//    (xLARFB, side L, trans T) by three MFMA GEMMs: W = V^T A2,
//    W2 = T^T W, A2 -= V W2. V^T V and V^T A2 have 64 rows and the panel's
//    height as k: they split k (GemmSplitKDevice) to fill the device.
// Norms are computed scale-safely: each workgroup sums (x 2^-e)^2 with e the
// exponent of its largest |x|, and the sums are combined at the largest e.
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

struct PanelArgs {
  double* p;       // the panel's top-left element, A(j0, j0)
  int64_t lda;
  int mp, ib;      // panel rows and columns
  double* tau;     // tau + j0
  double* beta;    // ib betas (the new diagonal), written by workgroup 0
  double* part;    // norm partials: [blocks][2] = {e, sum of (x 2^-e)^2}
  double* dots;    // dot partials: [blocks][kPanel]
};

__device__ __forceinline__ double WaveSum(double v) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

__device__ __forceinline__ double WaveMax(double v) {
  for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
  return v;
}

// This workgroup's norm partial of x (its threads' values, zero outside the
// column's range): {e, sum (x 2^-e)^2} with e the exponent of max |x| (the
// sum is 0 and e = -100000 when all are zero).
__device__ void NormPartial(double x, double* part_out) {
  __shared__ double red[kWaves];
  __shared__ double mx_s;
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  double m = WaveMax(fabs(x));
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
  double s = WaveSum(y * y);
  __syncthreads();
  if (lane == 0) red[wave] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int w = 0; w < kWaves; ++w) t += red[w];
    PutRecord(part_out, double(e));
    PutRecord(part_out + 1, t);
  }
}

// ||x||_2 from the partials of `blocks` workgroups, computed by one whole
// wave (every lane gets it): lane l takes partials l, l + 64, ..., then the
// largest exponent and the scaled sum are reduced over the wave. Every
// workgroup computes it from the same partials in the same order, so all
// of them get the same bits.
__device__ double NormOfWave(const double* part, int blocks) {
  const int lane = threadIdx.x & 63;
  int emax = -100000;
  for (int b = lane; b < blocks; b += 64) emax = max(emax, int(GetRecord(part + 2 * b)));
  emax = int(WaveMax(double(emax)));
  if (emax == -100000) return 0.0;
  double t = 0.0;
  for (int b = lane; b < blocks; b += 64) {
    const int e = int(GetRecord(part + 2 * b));
    if (e != -100000) t += ldexp(GetRecord(part + 2 * b + 1), 2 * (e - emax));
  }
  t = WaveSum(t);
  return ldexp(sqrt(t), emax);
}

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

// Norm partials of panel column 0 (rows 1..mp-1).
__global__ void __launch_bounds__(kRows) QrNormKernel(PanelArgs g) {
  const int i = blockIdx.x * kRows + threadIdx.x;
  const double x = (i > 0 && i < g.mp) ? g.p[i] : 0.0;
  NormPartial(x, g.part + 2 * blockIdx.x);
}

// Column c: the reflector (xLARFG) and the partial dots for columns c' > c.
__global__ void __launch_bounds__(kRows) QrReflectKernel(PanelArgs g, int c) {
  __shared__ double coef[2];  // scal, and the workgroup-0 copy of tau
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = blockIdx.x * kRows + threadIdx.x;
  double xnorm_w = 0.0;
  if (threadIdx.x < 64) xnorm_w = NormOfWave(g.part, gridDim.x);
  if (threadIdx.x == 0) {
    double tau, scal, beta;
    Larfg(g.p[c + int64_t(c) * g.lda], xnorm_w, tau, scal, beta);
    coef[0] = scal;
    coef[1] = tau;
    if (blockIdx.x == 0) {
      g.tau[c] = tau;
      g.beta[c] = beta;
    }
  }
  __syncthreads();
  const double scal = coef[0];
  // v(i) for this thread's row: 1 on the diagonal row, the scaled x below.
  double v = 0.0;
  if (i == c) {
    v = 1.0;
  } else if (i > c && i < g.mp) {
    double* pic = g.p + i + int64_t(c) * g.lda;
    v = *pic * scal;
    *pic = v;
  }
  // Partial dots with the later columns (row c contributes P(c, c') * 1):
  // every load first, then one sum per column over the wave, then over the
  // workgroup's waves (in wave order) through LDS.
  __shared__ double wsum[kWaves][kPanel];
  const bool mine = i >= c && i < g.mp;
  double pv[kPanel];
#pragma unroll
  for (int cc = 0; cc < kPanel; ++cc)
    pv[cc] = (mine && cc > c && cc < g.ib) ? g.p[i + int64_t(cc) * g.lda] : 0.0;
#pragma unroll
  for (int cc = 0; cc < kPanel; ++cc) {
    if (cc > c && cc < g.ib) {
      const double d = WaveSum(v * pv[cc]);
      if (lane == 0) wsum[wave][cc] = d;
    }
  }
  __syncthreads();
  const int cc = threadIdx.x;
  if (cc > c && cc < g.ib) {
    double d = wsum[0][cc];
#pragma unroll
    for (int w = 1; w < kWaves; ++w) d += wsum[w][cc];
    g.dots[int64_t(blockIdx.x) * kPanel + cc] = d;
  }
}

// Column c: w(c') = sum of the dot partials, P(i, c') -= tau v(i) w(c') for
// this workgroup's rows i >= c, the diagonal set to beta, and the norm
// partials of column c + 1 (rows c + 2 ..).
__global__ void __launch_bounds__(kRows) QrUpdateKernel(PanelArgs g, int c) {
  __shared__ double w[kPanel];
  const int i = blockIdx.x * kRows + threadIdx.x;
  const int nparts = gridDim.x;
  const bool mine = i >= c && i < g.mp;
  // This row's entries of the later columns, loaded while w is summed.
  double pv[kPanel];
#pragma unroll
  for (int cc = 0; cc < kPanel; ++cc)
    pv[cc] = (mine && cc > c && cc < g.ib) ? g.p[i + int64_t(cc) * g.lda] : 0.0;
  const double v = !mine ? 0.0 : (i == c ? 1.0 : g.p[i + int64_t(c) * g.lda]);
  const double tau = g.tau[c];
  // w(c') = the workgroups' dot partials in workgroup order, loads in
  // batches of 8.
  for (int cc = c + 1 + threadIdx.x; cc < g.ib; cc += blockDim.x) {
    double s = 0.0;
    int q = 0;
    for (; q + 8 <= nparts; q += 8) {
      double t[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) t[u] = g.dots[int64_t(q + u) * kPanel + cc];
#pragma unroll
      for (int u = 0; u < 8; ++u) s += t[u];
    }
    for (; q < nparts; ++q) s += g.dots[int64_t(q) * kPanel + cc];
    w[cc] = s;
  }
  __syncthreads();
  double next = 0.0;  // this row's entry of column c + 1 after the update
  if (mine) {
#pragma unroll
    for (int cc = 0; cc < kPanel; ++cc) {
      if (cc > c && cc < g.ib) {
        const double nv = pv[cc] - tau * v * w[cc];
        g.p[i + int64_t(cc) * g.lda] = nv;
        if (cc == c + 1) next = nv;
      }
    }
    if (i == c) g.p[i + int64_t(c) * g.lda] = g.beta[c];
  }
  if (c + 1 < g.ib) NormPartial(i > c + 1 && i < g.mp ? next : 0.0, g.part + 2 * blockIdx.x);
}

// The whole panel in one cooperative launch: each thread keeps its row of
// the panel in registers, and the two kernel boundaries of a column step
// become two grid-wide barriers (GridSync). Per column c, as QrReflectKernel
// and QrUpdateKernel with the same operations in the same order (so the
// bits are theirs): after barrier A every workgroup forms the norm from the
// partials and alpha from the record published by the owner of row c, and
// xLARFG's tau, beta and scale; it scales its rows of v (stored to the
// panel), and publishes its partial dots with the later columns; after
// barrier B it sums the partials in workgroup order, updates its rows, and
// publishes the next column's norm partial (and, on row c + 1, alpha).
// Records are reused every column: a record is rewritten only after the
// barrier that follows its last read. R's entries (rows < ib, right of the
// diagonal) are stored at the end. After a timed-out wait every workgroup
// returns and workgroup 0 sets the panel's tau to NaN.
struct PersistArgs {
  PanelArgs g;
  double* alpha;    // the record of P(c, c) before column c's reflector
  unsigned* bar;    // arrivals (bar_base at launch: the count runs on over panels)
  unsigned* failed;
  unsigned bar_base;
};

__global__ void __launch_bounds__(kRows) QrPanelPersistentKernel(PersistArgs a) {
  const PanelArgs& g = a.g;
  __shared__ double coef[3];  // scal, tau, beta
  __shared__ double wsum[kWaves][kPanel];
  __shared__ double w[kPanel];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int i = blockIdx.x * kRows + threadIdx.x;
  const int G = gridDim.x;
  const bool in = i < g.mp;
  double pv[kPanel];
#pragma unroll
  for (int cc = 0; cc < kPanel; ++cc)
    pv[cc] = (in && cc < g.ib) ? g.p[i + int64_t(cc) * g.lda] : 0.0;
  NormPartial(i > 0 && in ? pv[0] : 0.0, g.part + 2 * blockIdx.x);
  if (i == 0) PutRecord(a.alpha, pv[0]);
  double cur = pv[0];  // this row's entry of column c
  unsigned arrivals = a.bar_base;
  for (int c = 0; c < g.ib; ++c) {
    // An opaque copy of the column count: the per-column tests below are
    // then formed where they are used, not hoisted out of the column loop
    // into (spilled) scalar registers.
    int ib = g.ib;
    asm volatile("" : "+s"(ib));
    arrivals += unsigned(G);
    if (!GridSync(a.bar, arrivals, a.failed)) break;
    double xnorm_w = 0.0;
    if (threadIdx.x < 64) xnorm_w = NormOfWave(g.part, G);
    if (threadIdx.x == 0) {
      double tau, scal, beta;
      Larfg(GetRecord(a.alpha), xnorm_w, tau, scal, beta);
      coef[0] = scal;
      coef[1] = tau;
      coef[2] = beta;
      if (blockIdx.x == 0) {
        g.tau[c] = tau;
        g.beta[c] = beta;
      }
    }
    __syncthreads();
    const double scal = coef[0], tau = coef[1];
    double v = 0.0;
    if (i == c) {
      v = 1.0;
      g.p[i + int64_t(c) * g.lda] = coef[2];
    } else if (i > c && in) {
      v = cur * scal;
      g.p[i + int64_t(c) * g.lda] = v;
    }
    const bool mine = i >= c && in;
#pragma unroll
    for (int cc = 0; cc < kPanel; ++cc) {
      if (cc > c && cc < ib) {
        const double d = WaveSum(v * (mine ? pv[cc] : 0.0));
        if (lane == 0) wsum[wave][cc] = d;
      }
    }
    __syncthreads();
    {
      const int cc = threadIdx.x;
      if (cc > c && cc < ib) {
        double d = wsum[0][cc];
#pragma unroll
        for (int q = 1; q < kWaves; ++q) d += wsum[q][cc];
        PutRecord(g.dots + int64_t(blockIdx.x) * kPanel + cc, d);
      }
    }
    arrivals += unsigned(G);
    if (!GridSync(a.bar, arrivals, a.failed)) break;
    for (int cc = c + 1 + int(threadIdx.x); cc < ib; cc += kRows) {
      double sum = 0.0;
      int q = 0;
      for (; q + 8 <= G; q += 8) {
        double t[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) t[u] = GetRecord(g.dots + int64_t(q + u) * kPanel + cc);
#pragma unroll
        for (int u = 0; u < 8; ++u) sum += t[u];
      }
      for (; q < G; ++q) sum += GetRecord(g.dots + int64_t(q) * kPanel + cc);
      w[cc] = sum;
    }
    __syncthreads();
    double next = 0.0;
    if (mine) {
#pragma unroll
      for (int cc = 0; cc < kPanel; ++cc) {
        if (cc > c && cc < ib) {
          const double nv = pv[cc] - tau * v * w[cc];
          pv[cc] = nv;
          if (cc == c + 1) next = nv;
        }
      }
    }
    if (c + 1 < ib) {
      NormPartial(i > c + 1 && in ? next : 0.0, g.part + 2 * blockIdx.x);
      if (i == c + 1) PutRecord(a.alpha, next);
    }
    cur = next;
  }
  if (__hip_atomic_load(a.failed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
    if (blockIdx.x == 0 && int(threadIdx.x) < g.ib) g.tau[threadIdx.x] = __builtin_nan("");
    return;
  }
  // R right of the diagonal: rows i < ib, final since column i's update.
  int ib = g.ib;
  asm volatile("" : "+s"(ib));
  if (i < ib) {
#pragma unroll
    for (int cc = 0; cc < kPanel; ++cc)
      if (cc > i && cc < ib) g.p[i + int64_t(cc) * g.lda] = pv[cc];
  }
}

// V (mp x ib, leading dimension mp): the panel's reflectors with the unit
// diagonal and zeros above written out.
__global__ void ExplicitVKernel(const double* p, int64_t lda, int mp, int ib, double* v) {
  const int64_t total = int64_t(mp) * ib;
  for (int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; t < total;
       t += int64_t(gridDim.x) * blockDim.x) {
    const int i = int(t % mp), j = int(t / mp);
    v[t] = i == j ? 1.0 : (i > j ? p[i + int64_t(j) * lda] : 0.0);
  }
}

// T (ib x ib, upper, leading dimension kPanel) from G = V^T V and tau:
// T(0:i, i) = -tau_i T(0:i, 0:i) G(0:i, i), T(i, i) = tau_i; zeros below.
// One workgroup of kPanel threads, thread r owning row r of T in LDS (G and
// tau loaded into LDS first, every load in flight at once).
__global__ void __launch_bounds__(kPanel) LarftKernel(const double* gm, int ib,
                                                       const double* tau, double* t) {
  __shared__ double sg[kPanel][kPanel + 1];  // sg[q][i] = G(q, i)
  __shared__ double st[kPanel][kPanel + 1];  // st[r][q] = T(r, q)
  __shared__ double stau[kPanel];
  const int r = threadIdx.x;
  double gcol[kPanel];
#pragma unroll
  for (int i = 0; i < kPanel; ++i) gcol[i] = (r < ib && i < ib) ? gm[r + int64_t(i) * ib] : 0.0;
#pragma unroll
  for (int i = 0; i < kPanel; ++i) {
    sg[r][i] = gcol[i];
    st[r][i] = 0.0;
  }
  stau[r] = r < ib ? tau[r] : 0.0;
  __syncthreads();
  for (int i = 0; i < ib; ++i) {
    // T(r, i) = -tau_i sum_{q = r}^{i - 1} T(r, q) G(q, i), ascending q.
    if (r < i) {
      double s = 0.0;
      for (int q = r; q < i; ++q) s += st[r][q] * sg[q][i];
      st[r][i] = -stau[i] * s;
    }
    if (r == i) st[r][i] = stau[i];
    __syncthreads();
  }
#pragma unroll
  for (int j = 0; j < kPanel; ++j) t[r + j * kPanel] = st[r][j];
}

struct Work {
  double* alpha;  // the persistent panel's alpha record (8 doubles)
  unsigned* sync; // [0]: barrier arrivals, [32]: failed (64 words)
  double* v;      // m x kPanel
  double* t;      // kPanel x kPanel
  double* gm;     // kPanel x kPanel
  double* w;      // kPanel x n
  double* w2;     // kPanel x n
  double* part;   // [blocks][2]
  double* dots;   // [blocks][kPanel]
  double* beta;   // kPanel
  double* split;  // split-k partials of V^T V and V^T A2
  int64_t split_cap;
};

int64_t SplitCap(int64_t n) { return 8 * int64_t(kPanel) * std::max<int64_t>(n, 4096); }

size_t WorkBytes(int64_t m, int64_t n) {
  const int64_t blocks = (m + kRows - 1) / kRows;
  return 8 * (8 + 32 + size_t(m) * kPanel + 2 * kPanel * kPanel + 2 * size_t(kPanel) * n +
              2 * blocks + size_t(blocks) * kPanel + kPanel + size_t(SplitCap(n))) + 256;
}

Work Carve(void* base, int64_t m, int64_t n) {
  const int64_t blocks = (m + kRows - 1) / kRows;
  double* p = static_cast<double*>(base);
  Work w;
  w.alpha = p;
  p += 8;
  w.sync = reinterpret_cast<unsigned*>(p);
  p += 32;
  w.v = p;
  p += m * kPanel;
  w.t = p;
  p += kPanel * kPanel;
  w.gm = p;
  p += kPanel * kPanel;
  w.w = p;
  p += int64_t(kPanel) * n;
  w.w2 = p;
  p += int64_t(kPanel) * n;
  w.part = p;
  p += 2 * blocks;
  w.dots = p;
  p += blocks * kPanel;
  w.beta = p;
  p += kPanel;
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
  for (int64_t j0 = 0; j0 < k; j0 += kPanel) {
    const int ib = int(std::min<int64_t>(kPanel, k - j0));
    const int mp = int(m - j0);
    const int blocks = (mp + kRows - 1) / kRows;
    PanelArgs g{a + j0 + j0 * lda, lda, mp, ib, tau + j0, w.beta, w.part, w.dots};
    bool coop = false;
    if (blocks <= cap) {
      PersistArgs pa{g, w.alpha, w.sync, w.sync + 32, arrivals};
      void* args[] = {&pa};
      // A cooperative launch the runtime refuses (another process holding
      // the device's cooperative queue, for one) leaves this panel and the
      // later ones to the column launches, which give the same bits.
      const hipError_t e = hipLaunchCooperativeKernel(
          reinterpret_cast<const void*>(QrPanelPersistentKernel), dim3(blocks), dim3(kRows), args,
          0, s);
      if (e == hipSuccess) {
        coop = true;
        arrivals += 2u * unsigned(ib) * unsigned(blocks);
      } else {
        (void)hipGetLastError();
        cap = 0;
      }
    }
    if (!coop) {
      hipLaunchKernelGGL(QrNormKernel, dim3(blocks), dim3(kRows), 0, s, g);
      for (int c = 0; c < ib; ++c) {
        hipLaunchKernelGGL(QrReflectKernel, dim3(blocks), dim3(kRows), 0, s, g, c);
        hipLaunchKernelGGL(QrUpdateKernel, dim3(blocks), dim3(kRows), 0, s, g, c);
      }
    }
    HWY_HIP_CHECK(hipGetLastError());
    const int64_t n2 = n - j0 - ib;
    if (n2 <= 0) continue;
    // The block reflector, then H^T applied to the trailing columns.
    const int64_t vt = int64_t(mp) * ib;
    hipLaunchKernelGGL(ExplicitVKernel, dim3(unsigned(std::min<int64_t>((vt + 255) / 256, 2048))),
                       dim3(256), 0, s, g.p, lda, mp, ib, w.v);
    GemmSplitKDevice(true, false, ib, ib, mp, 1.0, w.v, mp, w.v, mp, 0.0, w.gm, ib, w.split,
                     w.split_cap, s);
    hipLaunchKernelGGL(LarftKernel, dim3(1), dim3(kPanel), 0, s, w.gm, ib, g.tau, w.t);
    HWY_HIP_CHECK(hipGetLastError());
    double* a2 = a + j0 + (j0 + ib) * lda;
    GemmSplitKDevice(true, false, ib, n2, mp, 1.0, w.v, mp, a2, lda, 0.0, w.w, ib, w.split,
                     w.split_cap, s);
    GemmDevice<double>(true, false, ib, n2, ib, 1.0, w.t, kPanel, w.w, ib, 0.0, w.w2, ib, s);
    GemmDevice<double>(false, false, mp, n2, ib, -1.0, w.v, mp, w.w2, ib, 1.0, a2, lda, s);
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
