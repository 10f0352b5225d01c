// Per-thread HIP context, device staging buffers, pointer classification and
// the exported device utilities (memory, streams, events, the seeded fill).
//
// Each calling thread uses, per device, a stream of default priority and
// grow-only staging buffers, the device counterpart of hwyBLAS's
// thread-local scratch slots (src/common.cc:46-64): reused across calls,
// freed first when they must grow (to keep peak memory low), never shrunk.
// The Fortran/CBLAS routines are therefore reentrant across threads without
// locks on the call path.
//
// The HIP resources are never released from a destructor: a thread-exit or
// process-exit destructor may run after the HIP runtime (or a profiler
// wrapped around it) has begun tearing down, and a HIP call from there can
// crash the exit. Instead, a thread that ends hands its set of contexts to a
// pool under a mutex, and the next thread that needs a set takes one from
// the pool. Device memory and streams thus stay bounded by the largest
// number of threads that called the library at once, and whatever is held
// at process exit is reclaimed with the process.
#include <stdio.h>
#include <stdlib.h>

#include <mutex>
#include <vector>

#include "common.h"

namespace hwy_amd {

namespace {

constexpr int kMaxDevices = 64;

struct ContextSet {
  ThreadContext ctx[kMaxDevices];
};

// Heap-allocated and never destroyed, so that threads ending during static
// destruction still find them.
std::mutex& PoolMutex() {
  static std::mutex* mu = new std::mutex;
  return *mu;
}
std::vector<ContextSet*>& Pool() {
  static std::vector<ContextSet*>* pool = new std::vector<ContextSet*>;
  return *pool;
}

// The calling thread's set, taken from the pool (or created) on first use
// and returned to the pool when the thread ends. Work a returned stream
// still holds stays ordered before whatever the next owner enqueues on it.
class ThreadSlot {
 public:
  ContextSet& Get() {
    if (set_ == nullptr) {
      std::lock_guard<std::mutex> lock(PoolMutex());
      std::vector<ContextSet*>& pool = Pool();
      if (pool.empty()) {
        set_ = new ContextSet;
      } else {
        set_ = pool.back();
        pool.pop_back();
      }
    }
    return *set_;
  }
  ~ThreadSlot() {
    if (set_ == nullptr) return;
    std::lock_guard<std::mutex> lock(PoolMutex());
    Pool().push_back(set_);
  }

 private:
  ContextSet* set_ = nullptr;
};

thread_local ThreadSlot t_slot;

}  // namespace

ThreadContext& Context() {
  int dev = 0;
  HWY_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    fprintf(stderr, "hwyblas_amd: device %d out of range\n", dev);
    abort();
  }
  ThreadContext& c = t_slot.Get().ctx[dev];
  if (c.device < 0) {
    HWY_HIP_CHECK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    c.device = dev;
  }
  return c;
}

void* Staging(int slot, size_t bytes, hipStream_t user_stream, bool* grown) {
  ThreadContext& c = Context();
  if (bytes == 0) bytes = 1;
  if (grown != nullptr) *grown = c.cap[slot] < bytes;
  if (c.cap[slot] < bytes) {
    if (c.buf[slot]) {
      // The old buffer may still be read by work queued on either stream.
      HWY_HIP_CHECK(hipStreamSynchronize(c.stream));
      if (user_stream != nullptr) HWY_HIP_CHECK(hipStreamSynchronize(user_stream));
      HWY_HIP_CHECK(hipFree(c.buf[slot]));
      c.buf[slot] = nullptr;
      c.cap[slot] = 0;
    }
    // Round up to 2 MiB so that slowly growing sizes do not reallocate often.
    const size_t rounded = (bytes + (size_t(2) << 20) - 1) & ~((size_t(2) << 20) - 1);
    HWY_HIP_CHECK(hipMalloc(&c.buf[slot], rounded));
    c.cap[slot] = rounded;
  }
  return c.buf[slot];
}

void* SplitPartials(size_t bytes, hipStream_t stream) {
  ThreadContext& c = Context();
  if (c.split_done != nullptr) {
    // Growing frees the old buffer, which the last user may still read.
    if (c.cap[kSlotSplit] < bytes) HWY_HIP_CHECK(hipEventSynchronize(c.split_done));
    else if (c.split_stream != stream) HWY_HIP_CHECK(hipStreamWaitEvent(stream, c.split_done, 0));
  }
  return Staging(kSlotSplit, bytes, stream);
}

void SplitPartialsDone(hipStream_t stream) {
  ThreadContext& c = Context();
  if (c.split_done == nullptr)
    HWY_HIP_CHECK(hipEventCreateWithFlags(&c.split_done, hipEventDisableTiming));
  HWY_HIP_CHECK(hipEventRecord(c.split_done, stream));
  c.split_stream = stream;
}

hipError_t Memcpy2DAsync(void* dst, size_t dpitch, const void* src, size_t spitch, size_t width,
                         size_t height, hipMemcpyKind kind, hipStream_t stream) {
  if (width == 0 || height == 0) return hipSuccess;
  if (height == 1 || (dpitch == width && spitch == width))
    return hipMemcpyAsync(dst, src, width * height, kind, stream);
  return hipMemcpy2DAsync(dst, dpitch, src, spitch, width, height, kind, stream);
}

bool IsDevicePointer(const void* p) {
  if (p == nullptr) return false;
  hipPointerAttribute_t attr;
  const hipError_t e = hipPointerGetAttributes(&attr, p);
  if (e != hipSuccess) {
    (void)hipGetLastError();  // unknown (pageable host) memory: clear the error
    return false;
  }
  return attr.type == hipMemoryTypeDevice || attr.type == hipMemoryTypeManaged;
}

namespace {

__device__ __forceinline__ uint64_t SplitMix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

uint64_t SplitMix64Host(uint64_t x) {
  x += 0x9E3779B97F4A7C15ull;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
  return x ^ (x >> 31);
}

template <typename T>
__global__ void FillUniformKernel(T* x, int64_t rows, int64_t cols, int64_t ld,
                                  uint64_t stream_key, double lo, double width) {
  const int64_t total = rows * cols;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
       idx += int64_t(gridDim.x) * blockDim.x) {
    const double u = double(SplitMix64(stream_key + uint64_t(idx)) >> 11) * 0x1.0p-53;
    const int64_t i = idx % rows, j = idx / rows;
    x[i + j * ld] = T(__builtin_fma(u, width, lo));
  }
}

// x(i, j) = +-u 2^e: u in [1, 2) with 52 random bits and the sign from one
// hash of the compact index, e uniform in [emin, emax] from a second.
__global__ void FillPow2Kernel(double* x, int64_t rows, int64_t cols, int64_t ld, uint64_t key,
                               uint64_t key2, int emin, int span) {
  const int64_t total = rows * cols;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
       idx += int64_t(gridDim.x) * blockDim.x) {
    const uint64_t w = SplitMix64(key + uint64_t(idx));
    const uint64_t w2 = SplitMix64(key2 + uint64_t(idx));
    const double u = 1.0 + double(w >> 12) * 0x1.0p-52;
    const int e = emin + int((w2 >> 32) % uint64_t(span));
    const double v = ldexp(u, e);
    const int64_t i = idx % rows, j = idx / rows;
    x[i + j * ld] = (w & 1) ? -v : v;
  }
}

// x(i, j) = (x(i, j) * rs[i]) * cs[j] (either scale may be absent).
__global__ void ScaleRowsColsKernel(double* x, int64_t rows, int64_t cols, int64_t ld,
                                    const double* rs, const double* cs) {
  const int64_t total = rows * cols;
  for (int64_t idx = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; idx < total;
       idx += int64_t(gridDim.x) * blockDim.x) {
    const int64_t i = idx % rows, j = idx / rows;
    double v = x[i + j * ld];
    if (rs) v = v * rs[i];
    if (cs) v = v * cs[j];
    x[i + j * ld] = v;
  }
}

}  // namespace
}  // namespace hwy_amd

using namespace hwy_amd;

extern "C" {

HWYBLAS_AMD_API int hwyblas_amd_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

HWYBLAS_AMD_API void* hwyblas_amd_malloc(size_t bytes) {
  void* p = nullptr;
  HWY_HIP_CHECK(hipMalloc(&p, bytes ? bytes : 1));
  return p;
}

HWYBLAS_AMD_API void hwyblas_amd_free(void* p) {
  if (p) HWY_HIP_CHECK(hipFree(p));
}

HWYBLAS_AMD_API void hwyblas_amd_memcpy(void* dst, const void* src, size_t bytes,
                                        int kind) {
  const hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice
                          : kind == 2 ? hipMemcpyDeviceToHost
                                      : hipMemcpyDeviceToDevice;
  HWY_HIP_CHECK(hipMemcpy(dst, src, bytes, k));
  // A device-to-device hipMemcpy may return before the copy ends; this call
  // returns only once the bytes are in place.
  HWY_HIP_CHECK(hipStreamSynchronize(nullptr));
}

HWYBLAS_AMD_API void hwyblas_amd_memset(void* dst, int value, size_t bytes) {
  HWY_HIP_CHECK(hipMemset(dst, value, bytes));
}

HWYBLAS_AMD_API void hwyblas_amd_synchronize(void) {
  HWY_HIP_CHECK(hipDeviceSynchronize());
}

HWYBLAS_AMD_API void* hwyblas_amd_thread_stream(void) { return Context().stream; }

HWYBLAS_AMD_API void* hwyblas_amd_stream_create(void) {
  hipStream_t s = nullptr;
  HWY_HIP_CHECK(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  return s;
}

HWYBLAS_AMD_API void hwyblas_amd_stream_destroy(void* stream) {
  if (stream) HWY_HIP_CHECK(hipStreamDestroy(static_cast<hipStream_t>(stream)));
}

HWYBLAS_AMD_API void hwyblas_amd_stream_synchronize(void* stream) {
  HWY_HIP_CHECK(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
}

HWYBLAS_AMD_API void* hwyblas_amd_event_create(void) {
  hipEvent_t e = nullptr;
  HWY_HIP_CHECK(hipEventCreate(&e));
  return e;
}

HWYBLAS_AMD_API void hwyblas_amd_event_destroy(void* event) {
  if (event) HWY_HIP_CHECK(hipEventDestroy(static_cast<hipEvent_t>(event)));
}

HWYBLAS_AMD_API void hwyblas_amd_event_record(void* event, void* stream) {
  HWY_HIP_CHECK(hipEventRecord(static_cast<hipEvent_t>(event),
                               static_cast<hipStream_t>(stream)));
}

HWYBLAS_AMD_API float hwyblas_amd_event_elapsed_ms(void* start, void* stop) {
  HWY_HIP_CHECK(hipEventSynchronize(static_cast<hipEvent_t>(stop)));
  float ms = 0;
  HWY_HIP_CHECK(hipEventElapsedTime(&ms, static_cast<hipEvent_t>(start),
                                    static_cast<hipEvent_t>(stop)));
  return ms;
}

HWYBLAS_AMD_API void hwyblas_amd_fill_uniform(void* x, int dtype_bytes,
                                              int64_t rows, int64_t cols,
                                              int64_t ld, uint64_t seed,
                                              double lo, double hi,
                                              void* stream) {
  if (rows <= 0 || cols <= 0) return;
  const uint64_t key = SplitMix64Host(seed ^ 0x5851F42D4C957F2Dull);
  const int64_t total = rows * cols;
  const unsigned blocks = unsigned(std::min<int64_t>((total + 255) / 256, 1 << 16));
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype_bytes == 8) {
    hipLaunchKernelGGL(FillUniformKernel<double>, dim3(blocks), dim3(256), 0, s,
                       static_cast<double*>(x), rows, cols, ld, key, lo, hi - lo);
  } else {
    hipLaunchKernelGGL(FillUniformKernel<float>, dim3(blocks), dim3(256), 0, s,
                       static_cast<float*>(x), rows, cols, ld, key, lo, hi - lo);
  }
  HWY_HIP_CHECK(hipGetLastError());
}

HWYBLAS_AMD_API void hwyblas_amd_fill_pow2(double* x, int64_t rows, int64_t cols, int64_t ld,
                                           uint64_t seed, int emin, int emax, void* stream) {
  if (rows <= 0 || cols <= 0 || emax < emin) return;
  const uint64_t key = SplitMix64Host(seed ^ 0x5851F42D4C957F2Dull);
  const uint64_t key2 = SplitMix64Host(seed ^ 0x2545F4914F6CDD1Dull);
  const unsigned blocks = unsigned(std::min<int64_t>((rows * cols + 255) / 256, 1 << 16));
  hipLaunchKernelGGL(FillPow2Kernel, dim3(blocks), dim3(256), 0, static_cast<hipStream_t>(stream),
                     x, rows, cols, ld, key, key2, emin, emax - emin + 1);
  HWY_HIP_CHECK(hipGetLastError());
}

HWYBLAS_AMD_API void hwyblas_amd_scale_rows_cols(double* x, int64_t rows, int64_t cols,
                                                 int64_t ld, const double* row_scale,
                                                 const double* col_scale, void* stream) {
  if (rows <= 0 || cols <= 0) return;
  const unsigned blocks = unsigned(std::min<int64_t>((rows * cols + 255) / 256, 1 << 16));
  hipLaunchKernelGGL(ScaleRowsColsKernel, dim3(blocks), dim3(256), 0,
                     static_cast<hipStream_t>(stream), x, rows, cols, ld, row_scale, col_scale);
  HWY_HIP_CHECK(hipGetLastError());
}

HWYBLAS_AMD_API void hwyblas_amd_capture_begin(void* stream) {
  HWY_HIP_CHECK(hipStreamBeginCapture(static_cast<hipStream_t>(stream),
                                      hipStreamCaptureModeThreadLocal));
}

HWYBLAS_AMD_API void* hwyblas_amd_capture_end(void* stream) {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  HWY_HIP_CHECK(hipStreamEndCapture(static_cast<hipStream_t>(stream), &graph));
  HWY_HIP_CHECK(hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0));
  HWY_HIP_CHECK(hipGraphDestroy(graph));
  return exec;
}

HWYBLAS_AMD_API void hwyblas_amd_graph_launch(void* exec, void* stream) {
  HWY_HIP_CHECK(hipGraphLaunch(static_cast<hipGraphExec_t>(exec), static_cast<hipStream_t>(stream)));
}

HWYBLAS_AMD_API void hwyblas_amd_graph_destroy(void* exec) {
  if (exec) HWY_HIP_CHECK(hipGraphExecDestroy(static_cast<hipGraphExec_t>(exec)));
}

HWYBLAS_AMD_API const char* hwyblas_amd_version(void) { return "hwyblas_amd 0.1.0 (gfx950)"; }

}  // extern "C"
