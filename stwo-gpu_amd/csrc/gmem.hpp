// Global-memory access helpers for kernels that read columns through device pointer
// tables.  A pointer loaded from a table is generic; casting it to the global address
// space lets the compiler emit global_* instead of flat_* accesses, and forming the byte
// offset in 32 bits (the caller guarantees offsets below 2^32: M31 columns hold at most
// 2^30 values) gives one load with a scalar base and a 32-bit vector offset instead of
// 64-bit address arithmetic per access.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#if defined(__HIP_DEVICE_COMPILE__)
#define SH_GLOBAL __attribute__((address_space(1)))
#else
#define SH_GLOBAL
#endif

namespace sh {

template <class T>
__device__ __forceinline__ SH_GLOBAL T* as_global(T* p) {
  return (SH_GLOBAL T*)p;
}

// element i of base; i * sizeof(T) must be below 2^32
template <class T>
__device__ __forceinline__ SH_GLOBAL T& at(SH_GLOBAL T* base, uint32_t i) {
  return *(SH_GLOBAL T*)((SH_GLOBAL char*)base + i * (uint32_t)sizeof(T));
}
template <class T>
__device__ __forceinline__ const SH_GLOBAL T& at(const SH_GLOBAL T* base, uint32_t i) {
  return *(const SH_GLOBAL T*)((const SH_GLOBAL char*)base + i * (uint32_t)sizeof(T));
}

}  // namespace sh
