// Grid-wide synchronization for cooperative (all workgroups resident)
// launches: a bounded arrival barrier and coherent record stores/loads.
//
// Records, the arrival counter and the polls are agent-scope atomic
// accesses: coherent across the XCDs' L2 caches without cache-wide
// write-back or invalidation fences, which would dominate a step that
// synchronizes once per matrix column. Each wait is bounded (kSpinLimit
// polls); a timed-out wait sets *failed, and every later wait, in any
// workgroup, gives up at its first poll, so that every workgroup returns.
// In a cooperative launch every workgroup is resident and the bound is
// never reached in a working run.
#pragma once

#include <hip/hip_runtime.h>

namespace hwy_amd {

constexpr unsigned kSpinLimit = 1u << 22;  // polls (a few seconds at most)

// Arrive, then wait until `target` arrivals in all. Every thread's record
// stores have completed (vmcnt) before thread 0 arrives; records are read
// after the wait with GetRecord. False when the wait timed out (here or in
// another workgroup).
__device__ inline bool GridSync(unsigned* bar, unsigned target, unsigned* failed) {
  __shared__ int ok;
  asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
  __syncthreads();
  if (threadIdx.x == 0) {
    __hip_atomic_fetch_add(bar, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int good = 1;
    unsigned polls = 0;
    while (__hip_atomic_load(bar, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(failed, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
          ++polls > kSpinLimit) {
        __hip_atomic_exchange(failed, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        good = 0;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    ok = good;
  }
  asm volatile("" ::: "memory");
  __syncthreads();
  return ok != 0;
}

__device__ __forceinline__ void PutRecord(double* x, double v) {
  __hip_atomic_store(x, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double GetRecord(const double* x) {
  return __hip_atomic_load(const_cast<double*>(x), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

}  // namespace hwy_amd
