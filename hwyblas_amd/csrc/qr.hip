// Householder QR on the GPU: dgeqrf_ (hwyBLAS src/lapack/geqrf.cc:66-150,
// include/hwyblas/lapack.h:1273-1283) and hwyblas_dgeqrf_dev.
//
// The factorization is LAPACK's (the same reflectors, signs and storage):
// column i's reflector H(i) = I - tau_i v v^T with v(i) = 1 (implicit),
// v(i+1:m) stored below the diagonal and R on and above it, beta =
// -sign(alpha) |(alpha, x)| (xLARFG, src/lapack/householder.cc:44-104). The
// blocking is the GPU's own:
//  - panels of kPanel = 64 columns; a panel (mp x 64, mp = m - j0) is
//    factored column by column by workgroups of 256 rows each, two launches
//    per column: QrReflectKernel (the norm from the previous launch's
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

#include "common.h"

namespace hwy_amd {
namespace qr {
namespace {

constexpr int kPanel = 64;   // columns per panel
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
    part_out[0] = double(e);
    part_out[1] = t;
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
  for (int b = lane; b < blocks; b += 64) emax = max(emax, int(part[2 * b]));
  emax = int(WaveMax(double(emax)));
  if (emax == -100000) return 0.0;
  double t = 0.0;
  for (int b = lane; b < blocks; b += 64) {
    const int e = int(part[2 * b]);
    if (e != -100000) t += ldexp(part[2 * b + 1], 2 * (e - emax));
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
    const double xnorm = xnorm_w;
    const double alpha = g.p[c + int64_t(c) * g.lda];
    double tau = 0.0, scal = 1.0, beta = alpha;
    if (xnorm != 0.0) {
      const double safmin = 2.2250738585072014e-308 / 1.1102230246251565e-16;
      double a = alpha, xn = xnorm;
      double b = -copysign(Lapy2(a, xn), a);
      double up = 1.0, down = 1.0;
      int knt = 0;
      while (fabs(b) < safmin && knt < 20) {  // as xLARFG: rescale by 1/safmin
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
  return 8 * (size_t(m) * kPanel + 2 * kPanel * kPanel + 2 * size_t(kPanel) * n +
              2 * blocks + size_t(blocks) * kPanel + kPanel + size_t(SplitCap(n))) + 256;
}

Work Carve(void* base, int64_t m, int64_t n) {
  const int64_t blocks = (m + kRows - 1) / kRows;
  double* p = static_cast<double*>(base);
  Work w;
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

// A = QR in place (m x n, lda), tau (device) of min(m, n); asynchronous.
void Geqrf(int64_t m, int64_t n, double* a, int64_t lda, double* tau, hipStream_t s) {
  const int64_t k = std::min(m, n);
  if (k == 0) return;
  const Work w = Carve(Staging(kSlotPackA, WorkBytes(m, n), s), m, n);
  for (int64_t j0 = 0; j0 < k; j0 += kPanel) {
    const int ib = int(std::min<int64_t>(kPanel, k - j0));
    const int mp = int(m - j0);
    const int blocks = (mp + kRows - 1) / kRows;
    PanelArgs g{a + j0 + j0 * lda, lda, mp, ib, tau + j0, w.beta, w.part, w.dots};
    hipLaunchKernelGGL(QrNormKernel, dim3(blocks), dim3(kRows), 0, s, g);
    for (int c = 0; c < ib; ++c) {
      hipLaunchKernelGGL(QrReflectKernel, dim3(blocks), dim3(kRows), 0, s, g, c);
      hipLaunchKernelGGL(QrUpdateKernel, dim3(blocks), dim3(kRows), 0, s, g, c);
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
  qr::Geqrf(M, N, da, host_a ? M : *lda, dtau, s);
  if (host_a)
    HWY_HIP_CHECK(hwy_amd::Memcpy2DAsync(a, size_t(*lda) * 8, da, size_t(M) * 8, size_t(M) * 8, N,
                                         hipMemcpyDeviceToHost, s));
  if (host_tau)
    HWY_HIP_CHECK(hipMemcpyAsync(tau, dtau, size_t(K) * 8, hipMemcpyDeviceToHost, s));
  HWY_HIP_CHECK(hipStreamSynchronize(s));
  work[0] = double(N) * 32.0;
}

}  // extern "C"
