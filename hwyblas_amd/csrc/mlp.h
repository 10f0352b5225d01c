// The grouped-GEMM cost model: a small MLP that predicts, from a problem's
// shape, the time per problem of every tile configuration. Four dense
// layers, width 64: inputs (log m, log n, log k, transa, transb, precision)
// -> 64 -> 64 -> 64 (ReLU) -> one output per configuration (log2 of the
// microseconds per problem). Parameters are FP32, row-major W[out][in]:
// W1 b1 W2 b2 W3 b3 W4 b4.
#ifndef HWYBLAS_AMD_CSRC_MLP_H_
#define HWYBLAS_AMD_CSRC_MLP_H_

#include "common.h"

namespace hwy_amd {
namespace mlp {

constexpr int kInputs = 6;
constexpr int kWidth = 64;
constexpr int kMaxOutputs = 32;

constexpr int OffW1() { return 0; }
constexpr int OffB1() { return kWidth * kInputs; }
constexpr int OffW2() { return OffB1() + kWidth; }
constexpr int OffB2() { return OffW2() + kWidth * kWidth; }
constexpr int OffW3() { return OffB2() + kWidth; }
constexpr int OffB3() { return OffW3() + kWidth * kWidth; }
constexpr int OffW4() { return OffB3() + kWidth; }
constexpr int OffB4(int outputs) { return OffW4() + outputs * kWidth; }
constexpr int NumParams(int outputs) { return OffB4(outputs) + outputs; }

struct Model {
  int outputs = 0;
  int device = -1;
  float* params = nullptr;  // device
  float* grads = nullptr;   // device
  float* adam_m = nullptr;  // device
  float* adam_v = nullptr;  // device
  int step = 0;
};

// The model policy 1 of the configuration choice uses (nullptr: none).
const Model* InstalledModel();

// Host features in, host predictions out (n x outputs).
void Predict(const Model& model, int n, const float* features, float* out, hipStream_t s);

}  // namespace mlp
}  // namespace hwy_amd

#endif  // HWYBLAS_AMD_CSRC_MLP_H_
