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
#define SH_CONST __attribute__((address_space(4)))
#else
#define SH_GLOBAL
#define SH_CONST
#endif

namespace sh {

template <class T>
__device__ __forceinline__ SH_GLOBAL T* as_global(T* p) {
  return (SH_GLOBAL T*)p;
}

// A pointer that is the same in every lane, passed through readfirstlane: where the compiler
// cannot prove it uniform (a value carried through a loop, say) the result still lives in
// SGPRs; where it can, the readfirstlane folds away.
template <class T>
__device__ __forceinline__ T* uniform(T* p) {
  const uint64_t a = (uint64_t)p;
  return (T*)(((uint64_t)(uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)(a >> 32)) << 32) |
              (uint32_t)__builtin_amdgcn_readfirstlane((uint32_t)a));
}

// p[i] read through the scalar cache into SGPRs: p and i uniform, and nothing this kernel
// writes (tables uploaded before the launch).  A generic pointer read in a kernel that also
// stores is otherwise a vector (flat) load, even at a uniform address.
template <class T>
__device__ __forceinline__ T ld_uniform(const T* p, uint32_t i) {
  return ((const SH_CONST T*)p)[i];
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

// One column as a raw buffer (a scalar resource descriptor): loads and stores take a
// vector byte offset and a scalar one, so accesses of a thread that differ by uniform
// strides share one vector offset and cost no vector address arithmetic (global_* accesses
// take only a scalar base plus a vector offset, and per-access uniform offsets become
// 64-bit vector adds).  The base must be uniform.  The descriptor's range is 2^32 - 1
// bytes and an access must end inside it: word offsets from the base stay below 2^30 - 1,
// so callers put the base at a block's first word rather than at a column's (the last
// word of a 2^30-word column is out of range from the column's start).
struct BufCol {
  __amdgpu_buffer_rsrc_t rs;
  // num_records 0xffffffff: no range clamp below 4 GiB; dword data format (word 3).  The
  // base goes through readfirstlane so the descriptor stays in SGPRs even where the
  // compiler cannot see that the pointer is uniform (else every access becomes a loop).
  __device__ __forceinline__ explicit BufCol(const void* p) : rs(make(p)) {}
  static __device__ __forceinline__ __amdgpu_buffer_rsrc_t make(const void* p) {
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(uniform(p)), (short)0, -1, 0x00020000);
  }
  // word i + s (i per lane, s uniform)
  __device__ __forceinline__ uint32_t load(uint32_t i, uint32_t s) const {
    return __builtin_amdgcn_raw_buffer_load_b32(rs, i * 4u, s * 4u, 0);
  }
  __device__ __forceinline__ void store(uint32_t v, uint32_t i, uint32_t s) const {
    __builtin_amdgcn_raw_buffer_store_b32(v, rs, i * 4u, s * 4u, 0);
  }
};

}  // namespace sh
