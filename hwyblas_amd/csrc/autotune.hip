// Autotuner for the grouped GEMM tile configurations, and the choice of a
// configuration per problem.
//
// A tuning sample is the time per problem of one tile configuration on one
// shape (dtype, transa, transb, m, n, k): a plan of `copies` problems of
// that shape, each with its own operands (so they stream from HBM as in a
// large uneven batch, instead of sitting in L2), is launched `reps` times
// after a warm-up; the time per problem is the event-timed total over
// reps x copies. `copies` grows until the output covers 2^25 elements, so
// every configuration fills the GPU, bounded by a 1.5 GiB operand budget.
//
// Choosing configurations for a batch:
//  - policy 0: the built-in rule (DefaultGroupedConfig);
//  - policy 1: the installed cost model (a trained MLP, mlp.hip) predicts
//    the time of every configuration and the fastest is taken; without an
//    installed model this falls back to policy 2;
//  - policy 2: exhaustive tuning: every distinct shape of the batch is timed
//    under every configuration (once per process: results are kept in a
//    table that later calls reuse).
#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>
#include <tuple>
#include <vector>

#include "common.h"
#include "grouped.h"
#include "mlp.h"

extern "C" void* hwyblas_gemm_plan_create(int, int, const hwyblas_gemm_problem*, const int*, int*,
                                          int*);
extern "C" void hwyblas_plan_destroy(void*);
extern "C" void hwyblas_amd_fill_uniform(void*, int, int64_t, int64_t, int64_t, uint64_t, double,
                                         double, void*);

namespace hwy_amd {
namespace grouped {
namespace {

constexpr int64_t kTuneOutputElems = int64_t(1) << 25;
constexpr size_t kTuneBudgetBytes = size_t(3) << 29;  // 1.5 GiB
constexpr int kMaxCopies = 8192;

// Operand memory for a tuning run, filled with uniform[-1, 1] values when
// (re)allocated (timing does not depend on them, but NaN or subnormal data
// could).
void* EnsureWorkspace(size_t bytes, hipStream_t s) {
  const size_t rounded = (bytes + (size_t(64) << 20) - 1) & ~((size_t(64) << 20) - 1);
  bool grown = false;
  void* p = Staging(kSlotTune, rounded, s, &grown);
  if (grown)
    hwyblas_amd_fill_uniform(p, 4, int64_t(rounded / 4), 1, int64_t(rounded / 4), 12345, -1.0,
                             1.0, s);
  return p;
}

int CopiesFor(int dtype_bytes, int m, int n, int k) {
  const int64_t per = (int64_t(m) * k + int64_t(k) * n + int64_t(m) * n) * dtype_bytes + 3 * 256;
  int64_t copies = (kTuneOutputElems + int64_t(m) * n - 1) / (int64_t(m) * n);
  copies = std::min<int64_t>(copies, kMaxCopies);
  copies = std::min<int64_t>(copies, int64_t(kTuneBudgetBytes) / per);
  return int(std::max<int64_t>(copies, 1));
}

size_t Align256(size_t x) { return (x + 255) & ~size_t(255); }

}  // namespace

// Times every configuration on one shape; us[cfg] = microseconds per problem.
void TuneShape(int dtype_bytes, bool ta, bool tb, int m, int n, int k, int reps, float* us,
               hipStream_t s) {
  const int copies = CopiesFor(dtype_bytes, m, n, k);
  const size_t a_bytes = Align256(size_t(m) * k * dtype_bytes);
  const size_t b_bytes = Align256(size_t(k) * n * dtype_bytes);
  const size_t c_bytes = Align256(size_t(m) * n * dtype_bytes);
  char* base = static_cast<char*>(
      EnsureWorkspace(size_t(copies) * (a_bytes + b_bytes + c_bytes), s));
  std::vector<hwyblas_gemm_problem> probs(copies);
  for (int i = 0; i < copies; ++i) {
    hwyblas_gemm_problem& p = probs[i];
    p.transa = ta ? 'T' : 'N';
    p.transb = tb ? 'T' : 'N';
    p.m = m;
    p.n = n;
    p.k = k;
    p.alpha = 1.0;
    p.beta = 0.0;
    char* slot = base + size_t(i) * (a_bytes + b_bytes + c_bytes);
    p.a = slot;
    p.lda = std::max(1, ta ? k : m);
    p.b = slot + a_bytes;
    p.ldb = std::max(1, tb ? n : k);
    p.c = slot + a_bytes + b_bytes;
    p.ldc = std::max(1, m);
  }
  hipEvent_t e0, e1;
  HWY_HIP_CHECK(hipEventCreate(&e0));
  HWY_HIP_CHECK(hipEventCreate(&e1));
  std::vector<int> cfgs(copies);
  for (int c = 0; c < kNumGroupedGemmConfigs; ++c) {
    std::fill(cfgs.begin(), cfgs.end(), c);
    int info = 0, bad = -1;
    Plan* plan = static_cast<Plan*>(
        hwyblas_gemm_plan_create(dtype_bytes, copies, probs.data(), cfgs.data(), &info, &bad));
    Enqueue(plan, s);
    HWY_HIP_CHECK(hipEventRecord(e0, s));
    for (int r = 0; r < reps; ++r) Enqueue(plan, s);
    HWY_HIP_CHECK(hipEventRecord(e1, s));
    HWY_HIP_CHECK(hipEventSynchronize(e1));
    float ms = 0;
    HWY_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
    us[c] = ms * 1000.0f / float(reps * copies);
    hwyblas_plan_destroy(plan);
  }
  HWY_HIP_CHECK(hipEventDestroy(e0));
  HWY_HIP_CHECK(hipEventDestroy(e1));
}

namespace {

using ShapeKey = std::tuple<int, int, int, int, int, int>;  // bytes, ta, tb, m, n, k

std::mutex& TableMutex() {
  static std::mutex mu;
  return mu;
}
std::map<ShapeKey, int>& TunedTable() {
  static std::map<ShapeKey, int> table;
  return table;
}

int Argmin(const float* v, int n) {
  int best = 0;
  for (int i = 1; i < n; ++i)
    if (v[i] < v[best]) best = i;
  return best;
}

}  // namespace

// Features of a shape for the cost model: log2 of m, n, k scaled to about
// [0, 1] over the batched range, the transpositions, and the precision.
void ShapeFeatures(int dtype_bytes, bool ta, bool tb, int m, int n, int k, float* f) {
  f[0] = std::log2(float(std::max(m, 1))) / 9.0f;
  f[1] = std::log2(float(std::max(n, 1))) / 9.0f;
  f[2] = std::log2(float(std::max(k, 1))) / 9.0f;
  f[3] = ta ? 1.0f : 0.0f;
  f[4] = tb ? 1.0f : 0.0f;
  f[5] = dtype_bytes == 8 ? 1.0f : 0.0f;
}

int ChooseConfigs(int dtype_bytes, int count, const hwyblas_gemm_problem* probs, int policy,
                  int* cfgs, hipStream_t s) {
  std::vector<int> todo;  // problems that need a choice
  std::vector<char> ta(count), tb(count);
  for (int i = 0; i < count; ++i) {
    const hwyblas_gemm_problem& p = probs[i];
    Op oa, ob;
    if (!DecodeOp(p.transa, &oa) || !DecodeOp(p.transb, &ob) || p.m < 0 || p.n < 0 || p.k < 0)
      return -1;
    ta[i] = oa == Op::kTrans;
    tb[i] = ob == Op::kTrans;
    cfgs[i] = DefaultGroupedConfig(dtype_bytes, p.m, p.n, p.alpha == 0.0 ? 0 : p.k);
    if (p.m > 0 && p.n > 0 && p.k > 0 && p.alpha != 0.0) todo.push_back(i);
  }
  if (policy == 0 || todo.empty()) return 0;
  const mlp::Model* model = mlp::InstalledModel();
  if (policy == 1 && model) {
    std::vector<float> feats(todo.size() * mlp::kInputs);
    for (size_t t = 0; t < todo.size(); ++t) {
      const hwyblas_gemm_problem& p = probs[todo[t]];
      ShapeFeatures(dtype_bytes, ta[todo[t]], tb[todo[t]], p.m, p.n, p.k,
                    &feats[t * mlp::kInputs]);
    }
    std::vector<float> pred(todo.size() * model->outputs);
    mlp::Predict(*model, int(todo.size()), feats.data(), pred.data(), s);
    for (size_t t = 0; t < todo.size(); ++t)
      cfgs[todo[t]] = Argmin(&pred[t * model->outputs], std::min(model->outputs,
                                                                 kNumGroupedGemmConfigs));
    return 0;
  }
  // Exhaustive tuning, shape by shape, through the process-wide table.
  std::lock_guard<std::mutex> lock(TableMutex());
  std::vector<float> us(kNumGroupedGemmConfigs);
  for (int i : todo) {
    const hwyblas_gemm_problem& p = probs[i];
    const ShapeKey key{dtype_bytes, ta[i], tb[i], p.m, p.n, p.k};
    auto it = TunedTable().find(key);
    if (it == TunedTable().end()) {
      TuneShape(dtype_bytes, ta[i], tb[i], p.m, p.n, p.k, 3, us.data(), s);
      it = TunedTable().emplace(key, Argmin(us.data(), kNumGroupedGemmConfigs)).first;
    }
    cfgs[i] = it->second;
  }
  return 0;
}

}  // namespace grouped
}  // namespace hwy_amd

using namespace hwy_amd;

extern "C" {

HWYBLAS_AMD_API int hwyblas_grouped_tune(int dtype_bytes, int nshapes, const int* shapes, int reps,
                                         float* us, void* stream) {
  if ((dtype_bytes != 4 && dtype_bytes != 8) || nshapes < 0 || reps < 1 || !shapes || !us)
    return -1;
  for (int i = 0; i < nshapes; ++i) {
    const int* sh = shapes + 5 * i;
    if (sh[2] < 1 || sh[3] < 1 || sh[4] < 1) return -1;
  }
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : Context().stream;
  for (int i = 0; i < nshapes; ++i) {
    const int* sh = shapes + 5 * i;
    grouped::TuneShape(dtype_bytes, sh[0] != 0, sh[1] != 0, sh[2], sh[3], sh[4], reps,
                       us + size_t(i) * grouped::kNumGroupedGemmConfigs, s);
  }
  return 0;
}

HWYBLAS_AMD_API int hwyblas_grouped_tune_copies(int dtype_bytes, int m, int n, int k) {
  if (m < 1 || n < 1 || k < 1) return 0;
  return grouped::CopiesFor(dtype_bytes, m, n, k);
}

HWYBLAS_AMD_API void hwyblas_grouped_features(int dtype_bytes, int ta, int tb, int m, int n,
                                              int k, float* features) {
  grouped::ShapeFeatures(dtype_bytes, ta != 0, tb != 0, m, n, k, features);
}

HWYBLAS_AMD_API int hwyblas_grouped_choose_configs(int dtype_bytes, int count,
                                                   const hwyblas_gemm_problem* problems,
                                                   int policy, int* configs, void* stream) {
  if ((dtype_bytes != 4 && dtype_bytes != 8) || count < 0 || policy < 0 || policy > 2 ||
      (count > 0 && (!problems || !configs)))
    return -1;
  hipStream_t s = stream ? static_cast<hipStream_t>(stream) : Context().stream;
  return grouped::ChooseConfigs(dtype_bytes, count, problems, policy, configs, s);
}

}  // extern "C"
