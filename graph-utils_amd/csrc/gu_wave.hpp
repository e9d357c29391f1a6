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
