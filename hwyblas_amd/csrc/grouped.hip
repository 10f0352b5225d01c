// Batched ("grouped") GEMM and TRSM over many independent, uneven problems
// in one persistent launch.
//
// A plan uploads the problem descriptors and the work lists once. GEMM work
// comes in two kinds: workgroup tiles (the four waves of a workgroup compute
// one large tile with the plain kernel's LDS-staged macro-tile) and wave
// tiles (one wave computes one small tile straight from memory). TRSM work
// items are blocks of 64 right-hand sides of one problem. The launch is a
// tiny kernel that resets the work counters followed by the persistent
// kernel, whose grid fills the GPU. Workgroups first claim workgroup tiles
// from one atomic counter (largest first) until none are left; then every
// wave claims units of wave tiles from a second counter. A unit is a run of
// consecutive small tiles of similar total cost, so one atomic serves many
// tiny tiles, while the largest tiles stay units of their own and keep the
// load balanced. No wave waits on another wave's work. The two launches
// touch no host memory and use no memset, so a plan can be captured in a
// hipGraph and replayed.
//
// Both tile kinds sum k in the order of the single-call kernel (MFMA steps
// of 4, ascending k), and the epilogue applies hwyBLAS's Store rules
// (src/gemm_kernel-inl.h:181-224), so every tile configuration gives the
// bits hwyblas_?gemm_dev gives. Per problem, (k == 0 or alpha == 0) with
// beta == 1 adds no work (C untouched); alpha == 0 never reads A or B.
//
// TRSM chunks solve op(A) X = alpha B (left) or X op(A) = alpha B (right)
// for up to 64 right-hand sides, one per lane: B := alpha B, then blocks of
// 16 rows are solved in LDS, each element as (b - sum_q l_iq x_q) *
// (1 / l_ii) (hwyBLAS's base kernel, src/trsm-inl.h:107-174), and each
// solved block updates the rows below it on the MFMA units.
#include <algorithm>
#include <cstring>
#include <vector>

#include "common.h"
#include "gemm_kernel.h"
#include "grouped.h"
#include "trsm_chunk.h"

namespace hwy_amd {
namespace grouped {

using trsm_chunk::TrsmChunk;
using trsm_chunk::TrsmDesc;
using trsm_chunk::TrsmWaveLds;

namespace {

using gemm_detail::GemmParams;
using gemm_detail::MacroTile;
using gemm_detail::Mfma;
using gemm_detail::TileLds;

template <typename T>
struct GemmDesc {
  const T* a;
  const T* b;
  T* c;
  int64_t lda, ldb, ldc;
  int m, n, k;   // k = 0 when alpha == 0 (A and B are not read)
  int ta, tb;
  T alpha, beta;
};

struct WorkItem {
  int problem;
  int i0, j0;  // first row / column (GEMM) or first right-hand side (TRSM)
  int cfg;
};

// One wave computes the BM x BN tile of C at (i0, j0).
template <typename T, int BM, int BN, int UK>
__device__ __forceinline__ void GemmTile(const GemmDesc<T>& d, int i0, int j0) {
  using M = Mfma<T>;
  using Acc = typename M::Acc;
  constexpr int TM = BM / 16, TN = BN / 16;
  const int lane = threadIdx.x & 63;
  const int li = lane & 15, lk = lane >> 4;
  Acc acc[TM][TN];
#pragma unroll
  for (int i = 0; i < TM; ++i)
#pragma unroll
    for (int j = 0; j < TN; ++j) acc[i][j] = Acc{};

  // op(A)(i, p) and op(B)(p, j) element offsets.
  const int64_t a_is = d.ta ? d.lda : 1, a_ps = d.ta ? 1 : d.lda;
  const int64_t b_ps = d.tb ? d.ldb : 1, b_js = d.tb ? 1 : d.ldb;
  int cols[TN];
  bool rok[TM], cok[TN];
  const T* ap[TM];  // this lane's row of op(A) at k = lk
  const T* bp[TN];  // this lane's column of op(B) at k = lk
#pragma unroll
  for (int ti = 0; ti < TM; ++ti) {
    const int row = i0 + ti * 16 + li;
    rok[ti] = row < d.m;
    ap[ti] = d.a + (rok[ti] ? int64_t(row) * a_is : 0) + int64_t(lk) * a_ps;
  }
#pragma unroll
  for (int tj = 0; tj < TN; ++tj) {
    cols[tj] = j0 + tj * 16 + li;
    cok[tj] = cols[tj] < d.n;
    bp[tj] = d.b + (cok[tj] ? int64_t(cols[tj]) * b_js : 0) + int64_t(lk) * b_ps;
  }
  // Loads of one step (k in [p0, p0 + 4 UK)); out-of-range entries are 0.
  auto load = [&](int p0, T (&af)[UK][TM], T (&bf)[UK][TN]) {
#pragma unroll
    for (int u = 0; u < UK; ++u) {
      const int q = p0 + 4 * u;
      const bool pok = q + lk < d.k;
#pragma unroll
      for (int ti = 0; ti < TM; ++ti)
        af[u][ti] = (rok[ti] && pok) ? ap[ti][int64_t(q) * a_ps] : T(0);
#pragma unroll
      for (int tj = 0; tj < TN; ++tj)
        bf[u][tj] = (cok[tj] && pok) ? bp[tj][int64_t(q) * b_ps] : T(0);
    }
  };
  auto mma = [&](const T (&af)[UK][TM], const T (&bf)[UK][TN]) {
#pragma unroll
    for (int u = 0; u < UK; ++u)
#pragma unroll
      for (int ti = 0; ti < TM; ++ti)
#pragma unroll
        for (int tj = 0; tj < TN; ++tj) acc[ti][tj] = M::Run(af[u][ti], bf[u][tj], acc[ti][tj]);
  };
  constexpr int kStep = 4 * UK;
  // 32-bit registers of one stage of operands and of the accumulators.
  constexpr int kWords = int(sizeof(T)) / 4;
  constexpr int kStageRegs = UK * (TM + TN) * kWords, kAccRegs = TM * TN * 4 * kWords;
  if constexpr (2 * kStageRegs + kAccRegs <= 72) {
    // Two register stages: the next step's loads are in flight while the
    // current step multiplies. k still advances in ascending steps of 4.
    T af0[UK][TM], bf0[UK][TN], af1[UK][TM], bf1[UK][TN];
    if (d.k > 0) load(0, af0, bf0);
    for (int p0 = 0; p0 < d.k; p0 += 2 * kStep) {
      if (p0 + kStep < d.k) load(p0 + kStep, af1, bf1);
      mma(af0, bf0);
      if (p0 + kStep >= d.k) break;
      if (p0 + 2 * kStep < d.k) load(p0 + 2 * kStep, af0, bf0);
      mma(af1, bf1);
    }
  } else {
    // Larger tiles keep one stage (two would spill at two waves per SIMD).
    for (int p0 = 0; p0 < d.k; p0 += kStep) {
      T af[UK][TM], bf[UK][TN];
      load(p0, af, bf);
      mma(af, bf);
    }
  }
  const T alpha = d.alpha, beta = d.beta;
  const bool beta_zero = beta == T(0), beta_one = beta == T(1);
#pragma unroll
  for (int ti = 0; ti < TM; ++ti) {
#pragma unroll
    for (int tj = 0; tj < TN; ++tj) {
      if (!cok[tj]) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int row = i0 + ti * 16 + M::Row(lk, r);
        if (row < d.m) {
          T* cij = d.c + row + cols[tj] * d.ldc;
          T v = d.k ? acc[ti][tj][r] * alpha : T(0);
          if (!beta_zero) v = beta_one ? v + *cij : __builtin_fma(*cij, beta, v);
          *cij = v;
        }
      }
    }
  }
}

template <typename T>
__device__ __forceinline__ void RunGemmItem(const GemmDesc<T>& d, const WorkItem& w) {
  switch (w.cfg) {
#define HWY_GROUPED_CASE(ID, BM, BN, UK) \
  case ID:                               \
    GemmTile<T, BM, BN, UK>(d, w.i0, w.j0); break;
    HWY_GROUPED_WAVE_CONFIGS(HWY_GROUPED_CASE)
#undef HWY_GROUPED_CASE
    default: break;
  }
}

// Workgroup tile `w` of problem `d` (all threads of the workgroup call it).
template <typename T, bool TA, bool TB>
__device__ __forceinline__ void BlockTile(const GemmParams<T>& p, const WorkItem& w, T* lds) {
  switch (w.cfg) {
#define HWY_GROUPED_BLOCK_CASE(ID, BM, BN, BK) \
  case ID:                                     \
    MacroTile<T, TA, TB, BM, BN, BK, 2, 2>(p, w.i0, w.j0, lds); break;
    HWY_GROUPED_BLOCK_CONFIGS(HWY_GROUPED_BLOCK_CASE)
#undef HWY_GROUPED_BLOCK_CASE
    default: break;
  }
}

template <typename T>
__device__ __forceinline__ void RunBlockItem(const GemmDesc<T>& d, const WorkItem& w, T* lds) {
  GemmParams<T> p;
  p.a = d.a;
  p.b = d.b;
  p.c = d.c;
  p.m = d.m;
  p.n = d.n;
  p.k = d.k;
  p.lda = d.lda;
  p.ldb = d.ldb;
  p.ldc = d.ldc;
  p.alpha = d.alpha;
  p.beta = d.beta;
  p.tiles_m = p.tiles_n = 0;
  if (!d.ta && !d.tb) BlockTile<T, false, false>(p, w, lds);
  else if (!d.ta && d.tb) BlockTile<T, false, true>(p, w, lds);
  else if (d.ta && !d.tb) BlockTile<T, true, false>(p, w, lds);
  else BlockTile<T, true, true>(p, w, lds);
}

constexpr int CMax(int a, int b) { return a > b ? a : b; }

// LDS elements of the largest workgroup tile over all transpositions.
template <typename T>
constexpr int BlockLdsElems() {
  int mx = 1;
#define HWY_GROUPED_LDS(ID, BM, BN, BK)                                \
  mx = CMax(mx, TileLds<T, false, false, BM, BN, BK>::kElems);         \
  mx = CMax(mx, TileLds<T, false, true, BM, BN, BK>::kElems);          \
  mx = CMax(mx, TileLds<T, true, false, BM, BN, BK>::kElems);          \
  mx = CMax(mx, TileLds<T, true, true, BM, BN, BK>::kElems);
  HWY_GROUPED_BLOCK_CONFIGS(HWY_GROUPED_LDS)
#undef HWY_GROUPED_LDS
  return mx;
}

// Counters live 128 bytes apart so their atomics do not share a line.
constexpr int kCounterStride = 32;
constexpr int kNumCounters = 2;

__global__ void ResetCounters(int* counters) {
  if (threadIdx.x < kNumCounters && blockIdx.x == 0) counters[threadIdx.x * kCounterStride] = 0;
}

// Claims the next work index for this wave (the same index in every lane).
__device__ __forceinline__ int ClaimItem(int* counter) {
  int t = 0;
  if ((threadIdx.x & 63) == 0) t = atomicAdd(counter, 1);
  return __builtin_amdgcn_readfirstlane(__shfl(t, 0, 64));
}

template <typename T>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2)))
    GroupedGemmKernel(const GemmDesc<T>* __restrict__ descs,
                      const WorkItem* __restrict__ block_items, int n_block,
                      const WorkItem* __restrict__ wave_items,
                      const int* __restrict__ unit_start, int n_units, int* counters) {
  __shared__ T lds[BlockLdsElems<T>()];
  __shared__ int claimed;
  // Phase 1: workgroup tiles, claimed by the workgroup as a whole.
  if (n_block > 0) {
    for (;;) {
      if (threadIdx.x == 0) claimed = atomicAdd(&counters[0], 1);
      __syncthreads();
      const int t = __builtin_amdgcn_readfirstlane(claimed);
      __syncthreads();
      if (t >= n_block) break;
      const WorkItem w = block_items[t];
      const GemmDesc<T> d = descs[w.problem];
      RunBlockItem<T>(d, w, lds);
    }
  }
  // Phase 2: units of wave tiles, claimed by each wave.
  for (;;) {
    const int u = ClaimItem(&counters[kCounterStride]);
    if (u >= n_units) break;
    const int t1 = unit_start[u + 1];
    for (int t = unit_start[u]; t < t1; ++t) {
      const WorkItem w = wave_items[t];
      const GemmDesc<T> d = descs[w.problem];
      RunGemmItem<T>(d, w);
    }
  }
}

// ---------------------------------------------------------------------------
// TRSM

template <typename T>
__global__ void __launch_bounds__(256)
    GroupedTrsmKernel(const TrsmDesc<T>* __restrict__ descs, const WorkItem* __restrict__ items,
                      const int* __restrict__ unit_start, int n_units, int* counters) {
  __shared__ TrsmWaveLds<T> lds[4];
  TrsmWaveLds<T>& mine = lds[threadIdx.x >> 6];
  for (;;) {
    const int u = ClaimItem(&counters[kCounterStride]);
    if (u >= n_units) break;
    const int t1 = unit_start[u + 1];
    for (int t = unit_start[u]; t < t1; ++t) {
      const WorkItem w = items[t];
      const TrsmDesc<T> d = descs[w.problem];
      TrsmChunk<T>(d, w.i0, mine);
    }
  }
}

// ---------------------------------------------------------------------------
// Plans

int DeviceCUs() {
  int dev = 0, cus = 0;
  HWY_HIP_CHECK(hipGetDevice(&dev));
  HWY_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
  return cus;
}

}  // namespace

struct Plan {
  int kind;  // 0 GEMM, 1 TRSM
  int dtype_bytes;
  void* descs = nullptr;
  WorkItem* items = nullptr;  // workgroup items, then wave items
  int* units = nullptr;       // n_units + 1 starts into the wave items
  int* counters = nullptr;
  int n_block = 0, n_wave = 0, n_units = 0;
  int blocks = 0;
  int device = -1;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  std::vector<int> cfg_of_problem;
};

namespace {

// Estimated cost of a tile: its area times k plus a fixed cost per tile
// (claim, descriptor load, epilogue) worth about 16 steps of k.
double TileCost(const GemmConfig& c, int k) { return double(c.bm) * c.bn * double(k + 16); }

// Workgroups of the persistent grid: two 256-thread workgroups per CU
// (the kernels run at two waves per SIMD).
constexpr int kBlocksPerCu = 2;

// Largest first (longest-processing-time order).
std::vector<int> ByCostDescending(const std::vector<double>& cost) {
  std::vector<int> order(cost.size());
  for (size_t t = 0; t < order.size(); ++t) order[t] = int(t);
  std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return cost[x] > cost[y]; });
  return order;
}

// Splits items (already in descending cost) into units of consecutive items:
// a unit closes once it holds at least `target` cost, so large items stay
// alone and runs of small ones share one claim.
std::vector<int> BuildUnits(const std::vector<double>& sorted_cost, double target) {
  std::vector<int> starts;
  double acc = 0;
  for (size_t t = 0; t < sorted_cost.size(); ++t) {
    if (acc == 0) starts.push_back(int(t));
    acc += sorted_cost[t];
    if (acc >= target) acc = 0;
  }
  starts.push_back(int(sorted_cost.size()));
  return starts;
}

// Uploads the work lists and sizes the grid.
void UploadWork(Plan* plan, const void* descs, size_t desc_bytes, const std::vector<WorkItem>& items,
                const std::vector<int>& units) {
  HWY_HIP_CHECK(hipGetDevice(&plan->device));
  HWY_HIP_CHECK(hipMalloc(&plan->descs, std::max<size_t>(1, desc_bytes)));
  HWY_HIP_CHECK(hipMalloc(&plan->items, std::max<size_t>(1, items.size()) * sizeof(WorkItem)));
  HWY_HIP_CHECK(hipMalloc(&plan->units, units.size() * sizeof(int)));
  HWY_HIP_CHECK(hipMalloc(&plan->counters, kNumCounters * kCounterStride * sizeof(int)));
  if (desc_bytes) HWY_HIP_CHECK(hipMemcpy(plan->descs, descs, desc_bytes, hipMemcpyHostToDevice));
  if (!items.empty())
    HWY_HIP_CHECK(hipMemcpy(plan->items, items.data(), items.size() * sizeof(WorkItem),
                            hipMemcpyHostToDevice));
  HWY_HIP_CHECK(hipMemcpy(plan->units, units.data(), units.size() * sizeof(int),
                          hipMemcpyHostToDevice));
  const int want = plan->n_block + (plan->n_units + 3) / 4;
  plan->blocks = std::max(1, std::min(DeviceCUs() * kBlocksPerCu, want));
}

template <typename T>
Plan* MakeGemmPlan(int count, const hwyblas_gemm_problem* probs, const int* cfgs, int* info,
                   int* bad) {
  *info = 0;
  *bad = -1;
  std::vector<GemmDesc<T>> descs(count);
  std::vector<WorkItem> block_items, wave_items;
  std::vector<double> block_cost, wave_cost;
  std::vector<int> chosen(count, -1);
  for (int i = 0; i < count; ++i) {
    const hwyblas_gemm_problem& p = probs[i];
    Op oa, ob;
    const bool va = DecodeOp(p.transa, &oa), vb = DecodeOp(p.transb, &ob);
    const int rows_a = oa == Op::kNoTrans ? p.m : p.k, rows_b = ob == Op::kNoTrans ? p.k : p.n;
    int e = 0;
    if (!va) e = 1;
    else if (!vb) e = 2;
    else if (p.m < 0) e = 3;
    else if (p.n < 0) e = 4;
    else if (p.k < 0) e = 5;
    else if (p.lda < std::max(1, rows_a)) e = 8;
    else if (p.ldb < std::max(1, rows_b)) e = 10;
    else if (p.ldc < std::max(1, p.m)) e = 13;
    if (e != 0) {
      *info = e;
      *bad = i;
      Reject(sizeof(T) == 8 ? "DGEMM" : "SGEMM", e);
      return nullptr;
    }
    GemmDesc<T>& d = descs[i];
    d.a = static_cast<const T*>(p.a);
    d.b = static_cast<const T*>(p.b);
    d.c = static_cast<T*>(p.c);
    d.lda = p.lda;
    d.ldb = p.ldb;
    d.ldc = p.ldc;
    d.m = p.m;
    d.n = p.n;
    d.ta = oa == Op::kTrans;
    d.tb = ob == Op::kTrans;
    d.alpha = T(p.alpha);
    d.beta = T(p.beta);
    const bool no_product = p.k == 0 || d.alpha == T(0);
    d.k = no_product ? 0 : p.k;
    if (p.m == 0 || p.n == 0) continue;
    if (no_product && d.beta == T(1)) continue;
    int cfg = cfgs ? cfgs[i] : -1;
    if (cfg < 0 || cfg >= kNumGroupedGemmConfigs)
      cfg = DefaultGroupedConfig(int(sizeof(T)), p.m, p.n, d.k);
    // Without a product only the epilogue runs (C = beta C): a wave tile.
    if (no_product && kGroupedGemmConfigs[cfg].block) cfg = 9;
    chosen[i] = cfg;
    const GemmConfig& gc = kGroupedGemmConfigs[cfg];
    auto& list = gc.block ? block_items : wave_items;
    auto& cost = gc.block ? block_cost : wave_cost;
    for (int j0 = 0; j0 < p.n; j0 += gc.bn)
      for (int i0 = 0; i0 < p.m; i0 += gc.bm) {
        list.push_back({i, i0, j0, cfg});
        cost.push_back(TileCost(gc, d.k));
      }
  }
  std::vector<WorkItem> items;
  items.reserve(block_items.size() + wave_items.size());
  for (int t : ByCostDescending(block_cost)) items.push_back(block_items[t]);
  std::vector<double> sorted_wave_cost;
  sorted_wave_cost.reserve(wave_items.size());
  double total = 0;
  for (int t : ByCostDescending(wave_cost)) {
    items.push_back(wave_items[t]);
    sorted_wave_cost.push_back(wave_cost[t]);
    total += wave_cost[t];
  }
  const int resident_waves = DeviceCUs() * kBlocksPerCu * 4;
  const std::vector<int> units = BuildUnits(sorted_wave_cost, total / (8.0 * resident_waves));

  Plan* plan = new Plan();
  plan->kind = 0;
  plan->dtype_bytes = int(sizeof(T));
  plan->n_block = int(block_items.size());
  plan->n_wave = int(wave_items.size());
  plan->n_units = int(units.size()) - 1;
  plan->cfg_of_problem = chosen;
  UploadWork(plan, descs.data(), descs.size() * sizeof(GemmDesc<T>), items, units);
  return plan;
}

template <typename T>
Plan* MakeTrsmPlan(int count, const hwyblas_trsm_problem* probs, int* info, int* bad) {
  *info = 0;
  *bad = -1;
  std::vector<TrsmDesc<T>> descs(count);
  std::vector<WorkItem> items;
  std::vector<double> cost;
  for (int i = 0; i < count; ++i) {
    const hwyblas_trsm_problem& p = probs[i];
    bool left, lower, unit;
    Op op;
    const bool vs = DecodeSide(p.side, &left), vu = DecodeUplo(p.uplo, &lower);
    const bool vt = DecodeOp(p.transa, &op), vd = DecodeDiag(p.diag, &unit);
    const int rows_a = left ? p.m : p.n;
    int e = 0;
    if (!vs) e = 1;
    else if (!vu) e = 2;
    else if (!vt) e = 3;
    else if (!vd) e = 4;
    else if (p.m < 0) e = 5;
    else if (p.n < 0) e = 6;
    else if (p.lda < std::max(1, rows_a)) e = 9;
    else if (p.ldb < std::max(1, p.m)) e = 11;
    if (e != 0) {
      *info = e;
      *bad = i;
      Reject(sizeof(T) == 8 ? "DTRSM" : "STRSM", e);
      return nullptr;
    }
    TrsmDesc<T>& d = descs[i];
    const bool trans = op == Op::kTrans;
    d.a = static_cast<const T*>(p.a);
    d.b = static_cast<T*>(p.b);
    d.rs = trans ? p.lda : 1;
    d.cs = trans ? 1 : p.lda;
    d.ldb = p.ldb;
    d.m = p.m;
    d.n = p.n;
    d.left = left;
    d.unit = unit;
    // op(A) is lower iff (uplo lower) xor trans; L = op(A) (left) or op(A)^T.
    const bool op_lower = lower != trans;
    d.lower_solve = left ? op_lower : !op_lower;
    d.alpha = T(p.alpha);
    if (p.m == 0 || p.n == 0) continue;
    const int len = left ? p.m : p.n, nv = left ? p.n : p.m;
    for (int v0 = 0; v0 < nv; v0 += 64) {
      items.push_back({i, v0, 0, 0});
      cost.push_back(double(len) * len);
    }
  }
  std::vector<WorkItem> sorted;
  std::vector<double> sorted_cost;
  for (int t : ByCostDescending(cost)) {
    sorted.push_back(items[t]);
    sorted_cost.push_back(cost[t]);
  }
  // Every chunk is a unit of its own (a chunk is already a long solve).
  const std::vector<int> units = BuildUnits(sorted_cost, 0.0);
  Plan* plan = new Plan();
  plan->kind = 1;
  plan->dtype_bytes = int(sizeof(T));
  plan->n_wave = int(sorted.size());
  plan->n_units = int(units.size()) - 1;
  UploadWork(plan, descs.data(), descs.size() * sizeof(TrsmDesc<T>), sorted, units);
  return plan;
}

}  // namespace

void Enqueue(const Plan* plan, hipStream_t stream) {
  if (plan->n_block + plan->n_wave == 0) return;
  hipLaunchKernelGGL(ResetCounters, dim3(1), dim3(64), 0, stream, plan->counters);
  const dim3 grid(unsigned(plan->blocks)), block(256);
  const WorkItem* wave_items = plan->items + plan->n_block;
  if (plan->kind == 0) {
    if (plan->dtype_bytes == 8)
      hipLaunchKernelGGL(GroupedGemmKernel<double>, grid, block, 0, stream,
                         static_cast<const GemmDesc<double>*>(plan->descs), plan->items,
                         plan->n_block, wave_items, plan->units, plan->n_units, plan->counters);
    else
      hipLaunchKernelGGL(GroupedGemmKernel<float>, grid, block, 0, stream,
                         static_cast<const GemmDesc<float>*>(plan->descs), plan->items,
                         plan->n_block, wave_items, plan->units, plan->n_units, plan->counters);
  } else {
    if (plan->dtype_bytes == 8)
      hipLaunchKernelGGL(GroupedTrsmKernel<double>, grid, block, 0, stream,
                         static_cast<const TrsmDesc<double>*>(plan->descs), wave_items,
                         plan->units, plan->n_units, plan->counters);
    else
      hipLaunchKernelGGL(GroupedTrsmKernel<float>, grid, block, 0, stream,
                         static_cast<const TrsmDesc<float>*>(plan->descs), wave_items,
                         plan->units, plan->n_units, plan->counters);
  }
  HWY_HIP_CHECK(hipGetLastError());
}

void LaunchGraph(Plan* plan, hipStream_t stream) {
  if (plan->n_block + plan->n_wave == 0) return;
  if (!plan->exec) {
    // Capture on a private stream (the caller's may already be capturing).
    hipStream_t cap;
    HWY_HIP_CHECK(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
    HWY_HIP_CHECK(hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal));
    Enqueue(plan, cap);
    HWY_HIP_CHECK(hipStreamEndCapture(cap, &plan->graph));
    HWY_HIP_CHECK(hipGraphInstantiate(&plan->exec, plan->graph, nullptr, nullptr, 0));
    HWY_HIP_CHECK(hipStreamDestroy(cap));
  }
  HWY_HIP_CHECK(hipGraphLaunch(plan->exec, stream));
}

void Destroy(Plan* plan) {
  if (!plan) return;
  if (plan->exec) (void)hipGraphExecDestroy(plan->exec);
  if (plan->graph) (void)hipGraphDestroy(plan->graph);
  (void)hipFree(plan->descs);
  (void)hipFree(plan->items);
  (void)hipFree(plan->units);
  (void)hipFree(plan->counters);
  delete plan;
}

// Without tuning data: a 64 x 64 workgroup tile for large problems, else
// the 64 x 64 wave tile (tiles of fewer rows cost more claims and loads than
// they save in padding on the mixed batches measured).
int DefaultGroupedConfig(int dtype_bytes, int m, int n, int k) {
  (void)dtype_bytes;
  if (m >= 96 && n >= 96 && k >= 32) return 10;
  return 9;
}

}  // namespace grouped
}  // namespace hwy_amd

using namespace hwy_amd;

extern "C" {

HWYBLAS_AMD_API void* hwyblas_gemm_plan_create(int dtype_bytes, int count,
                                               const hwyblas_gemm_problem* problems,
                                               const int* configs, int* info, int* bad_index) {
  int i = 0, b = -1;
  grouped::Plan* p = (dtype_bytes == 8)
                         ? grouped::MakeGemmPlan<double>(count, problems, configs, &i, &b)
                         : grouped::MakeGemmPlan<float>(count, problems, configs, &i, &b);
  if (info) *info = i;
  if (bad_index) *bad_index = b;
  return p;
}

HWYBLAS_AMD_API void* hwyblas_trsm_plan_create(int dtype_bytes, int count,
                                               const hwyblas_trsm_problem* problems, int* info,
                                               int* bad_index) {
  int i = 0, b = -1;
  grouped::Plan* p = (dtype_bytes == 8)
                         ? grouped::MakeTrsmPlan<double>(count, problems, &i, &b)
                         : grouped::MakeTrsmPlan<float>(count, problems, &i, &b);
  if (info) *info = i;
  if (bad_index) *bad_index = b;
  return p;
}

HWYBLAS_AMD_API void hwyblas_plan_launch(void* plan, void* stream) {
  grouped::Enqueue(static_cast<grouped::Plan*>(plan), static_cast<hipStream_t>(stream));
}

HWYBLAS_AMD_API void hwyblas_plan_launch_graph(void* plan, void* stream) {
  grouped::LaunchGraph(static_cast<grouped::Plan*>(plan), static_cast<hipStream_t>(stream));
}

HWYBLAS_AMD_API int hwyblas_plan_work_items(void* plan) {
  if (!plan) return 0;
  const auto* p = static_cast<grouped::Plan*>(plan);
  return p->n_block + p->n_wave;
}

HWYBLAS_AMD_API int hwyblas_plan_problem_config(void* plan, int problem) {
  auto* p = static_cast<grouped::Plan*>(plan);
  if (!p || problem < 0 || problem >= int(p->cfg_of_problem.size())) return -1;
  return p->cfg_of_problem[problem];
}

HWYBLAS_AMD_API void hwyblas_plan_destroy(void* plan) {
  grouped::Destroy(static_cast<grouped::Plan*>(plan));
}

HWYBLAS_AMD_API int hwyblas_grouped_config_count(void) { return grouped::kNumGroupedGemmConfigs; }

HWYBLAS_AMD_API int hwyblas_grouped_config_info(int cfg, int* bm, int* bn, int* unroll,
                                                int* block) {
  if (cfg < 0 || cfg >= grouped::kNumGroupedGemmConfigs) return -1;
  *bm = grouped::kGroupedGemmConfigs[cfg].bm;
  *bn = grouped::kGroupedGemmConfigs[cfg].bn;
  *unroll = grouped::kGroupedGemmConfigs[cfg].uk;
  if (block) *block = grouped::kGroupedGemmConfigs[cfg].block;
  return 0;
}

HWYBLAS_AMD_API int hwyblas_grouped_default_config(int dtype_bytes, int m, int n, int k) {
  return grouped::DefaultGroupedConfig(dtype_bytes, m, n, k);
}

}  // extern "C"
