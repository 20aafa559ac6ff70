// Direct global-to-LDS loads (global_load_lds_dwordx4) and the waits and
// barriers that go with them, for kernels that stage tiles through a ring of
// LDS buffers without passing the data through registers.
//
// The loads are issued from asm, so the compiler does not count them: its own
// vmcnt waits for other loads only grow more conservative, and it adds no
// vmcnt(0) before reading a stage; the kernel waits with wait_vmcnt for the
// loads it issued itself and makes the other waves' loads visible with
// stage_barrier.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace tups {

typedef __attribute__((address_space(3))) int8_t lds_i8_t;

// LDS byte address of a __shared__ object, uniform across the wave
template <class T>
__device__ __forceinline__ unsigned lds_addr(const T* p) {
  return __builtin_amdgcn_readfirstlane((unsigned)(uintptr_t)(const lds_i8_t*)(const void*)p);
}

// M0 carries the LDS base; nothing else in these kernels reads M0
#pragma clang diagnostic push
#pragma clang diagnostic ignored "-Winline-asm"
// 16 bytes from gbase + voff (uniform 64-bit base, per-lane 32-bit offset) to
// LDS at lds_base + 16 lane
__device__ __forceinline__ void dma16(const void* gbase, unsigned voff, unsigned lds_base) {
  asm volatile(
      "s_mov_b32 m0, %2\n\t"
      "global_load_lds_dwordx4 %0, %1"
      :
      : "v"(voff), "s"(gbase), "s"(lds_base)
      : "m0");
}
#pragma clang diagnostic pop

template <int N>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(N >= 0 && N < 16, "vmcnt immediate");
  __builtin_amdgcn_s_waitcnt(0x0F70 | N);  // vmcnt(N); expcnt, lgkmcnt not waited
}

// a workgroup barrier that orders the compiler's LDS accesses but adds no wait
// for outstanding loads (a fence would wait for the LDS loads in flight)
__device__ __forceinline__ void stage_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

}  // namespace tups
