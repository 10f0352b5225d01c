// Direct global -> LDS loads (global_load_lds_dwordx4) and the waits and
// barriers of the pipelines built on them, shared by the INT8 / BF16
// emulation kernels and the DMA-fed FP64 / FP32 GEMM.
#ifndef HWYBLAS_AMD_CSRC_DMA_H_
#define HWYBLAS_AMD_CSRC_DMA_H_

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace hwy_amd {

template <int N>
__device__ __forceinline__ void WaitVmcntImm() {
  // s_waitcnt with only vmcnt constrained (expcnt 7, lgkmcnt 15 = no wait);
  // vmcnt is 6 bits, split over [3:0] and [15:14].
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 0xF) | (7 << 4) | (0xF << 8) | ((N >> 4) << 14));
}

// Waits until at most `slices` later slices (PER_SLICE DMAs each) are
// outstanding.
template <int PER_SLICE>
__device__ __forceinline__ void WaitSlices(int slices) {
  switch (slices) {
    case 0: WaitVmcntImm<0>(); break;
    case 1: WaitVmcntImm<PER_SLICE>(); break;
    case 2: WaitVmcntImm<2 * PER_SLICE>(); break;
    default: WaitVmcntImm<3 * PER_SLICE>(); break;
  }
}

// One global_load_lds_dwordx4: this lane's 16 bytes at `src` land in LDS at
// `lds_wave_base` + 16 * lane. Issued as inline assembly so that the
// compiler, which would otherwise wait for every outstanding DMA before any
// LDS read, leaves the waiting to WaitVmcnt. M0 (the LDS base) must be
// written one instruction before the DMA reads it.
__device__ __forceinline__ void DmaToLds(const void* src, const void* lds_wave_base) {
  const unsigned base =
      __builtin_amdgcn_readfirstlane(unsigned(reinterpret_cast<uintptr_t>(lds_wave_base)));
  unsigned saved;
  asm volatile(
      "s_mov_b32 %0, m0\n\ts_mov_b32 m0, %1\n\ts_nop 0\n\t"
      "global_load_lds_dwordx4 %2, off\n\ts_mov_b32 m0, %0"
      : "=&s"(saved)
      : "s"(base), "v"(src)
      : "memory");
}

__device__ __forceinline__ void BarrierNoFence() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

}  // namespace hwy_amd

#endif  // HWYBLAS_AMD_CSRC_DMA_H_
