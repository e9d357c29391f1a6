// gu_wave.hpp -- wave64 helpers used by the engines.
//
// The primitives that the reference exposes to kernel authors (lane and wave
// ids, bit extract / insert, race_and_resolve, the wave-aggregated
// increment, the wave and workgroup scans) live in the public device header
// include/graph_utils_amd_device.hpp, which this file includes, so the
// engines and callers' kernels run the same code.  This file adds the
// engine-only helpers: uniform reads of another lane's value and a 64-bit
// OR over the wave.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/graph_utils_amd_device.hpp"

namespace gu {

// Uniform read of a 32/64-bit value from lane k (k must be wave-uniform).
__device__ __forceinline__ int readlane(int v, int k) { return __builtin_amdgcn_readlane(v, k); }
__device__ __forceinline__ uint64_t readlane64(uint64_t v, int k) {
  uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, k);
  uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), k);
  return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t readfirst64(uint64_t v) {
  uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
  uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// The same pointer, declared global.  A row pointer read from a device
// table (or chosen between such pointers) is generic to the compiler, which
// then emits flat loads, stores and atomics; flat operations also count
// against the LDS wait counter, so every wait for an LDS result would wait
// for them too.  Every row, queue and stack of the engines lives in global
// memory, and this cast lets the compiler emit global_* instructions.
// Functions that take such pointers take them as gp<T> (a cast to a
// generic pointer and back would be folded away and the knowledge lost).
template <class T>
using gp = __attribute__((address_space(1))) T *;
template <class T>
__device__ __forceinline__ gp<T> glob(T *p) {
  return (gp<T>)p;
}
// Atomics on such pointers (the generic forms, whose addresses the
// compiler then knows to be global).
__device__ __forceinline__ int cas_g(gp<int> p, int cmp, int val) { return atomicCAS((int *)p, cmp, val); }
__device__ __forceinline__ unsigned long long add_g(gp<unsigned long long> p, unsigned long long v) {
  return atomicAdd((unsigned long long *)p, v);
}

// First of `k` consecutive positions for this lane in a list whose length
// counter is *cnt, with one atomic per wave instead of one per lane (a
// wave prefix sum of the lanes' k places them).  Every lane of a full wave
// must call it (wave-uniform control flow).  Up to 1024 lanes of a
// workgroup taking positions from one LDS counter serialise on its
// address; 16 wave atomics do not.
__device__ __forceinline__ unsigned wave_reserve(unsigned *cnt, unsigned k) {
  const int incl = wave_inclusive_sum((int)k);
  const int tot = __builtin_amdgcn_readlane(incl, kWave - 1);
  unsigned base = 0;
  if (tot) {
    if (lane_id() == kWave - 1) base = atomicAdd(cnt, (unsigned)tot);
    base = (unsigned)__builtin_amdgcn_readlane((int)base, kWave - 1);
  }
  return base + (unsigned)(incl - (int)k);
}

// 64-bit OR over the wave (xor-butterfly of shuffles).
__device__ __forceinline__ uint64_t wave_or(uint64_t x) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) x |= (uint64_t)__shfl_xor((unsigned long long)x, o, kWave);
  return x;
}

}  // namespace gu
