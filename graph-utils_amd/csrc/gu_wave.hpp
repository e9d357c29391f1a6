// gu_wave.hpp -- wave64 primitives for gfx950 (CDNA4).
//
// Replaces graph-utils' warp-32 helpers with 64-lane forms:
//   lane / wave ids           kernel_utils.cuh:8-19 (getWarpId, getLaneId)
//   bit extract / insert      kernel_utils.cuh:22-35 (getBit, setBit)
//   race_and_resolve_wave     race_and_resolve.cuh:13-25 (+ __bfind, .cu:5-10)
//   wave-aggregated increment shuffle_based.cuh:58-76 (atomicAggInc)
//   wave / workgroup scans    the CUB WarpScan / BlockScan the reference uses
//                             (load_balanced_search.cu:65,125; scan_based.cu:107)
// Ballots are 64-bit; ranks come from mbcnt; scans use DPP row shifts plus
// row broadcasts, so no LDS round trip is needed inside a wave.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace gu {

constexpr int kWave = 64;

__device__ __forceinline__ int lane_id() {
  return __builtin_amdgcn_mbcnt_hi(~0u, __builtin_amdgcn_mbcnt_lo(~0u, 0u));
}
__device__ __forceinline__ int wave_id() { return threadIdx.x / kWave; }

__device__ __forceinline__ uint64_t ballot(bool p) { return __ballot(p); }

// Number of set bits of `mask` in lanes below this one.
__device__ __forceinline__ int rank_in(uint64_t mask) {
  return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                   __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

__device__ __forceinline__ int get_bit(uint32_t v, int pos) {
  return (int)__builtin_amdgcn_ubfe(v, (uint32_t)pos, 1u);
}
__device__ __forceinline__ uint32_t set_bit(uint32_t v, uint32_t bit, int pos) {
  return (v & ~(1u << pos)) | ((bit & 1u) << pos);
}

// Uniform read of a 32/64-bit value from lane k (k must be wave-uniform).
__device__ __forceinline__ int readlane(int v, int k) {
  return __builtin_amdgcn_readlane(v, k);
}
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

// Highest lane whose predicate holds, the same for every lane; a lane gets
// its own id when no predicate holds (race_and_resolve.cuh:13-25 semantics,
// with the 32-bit bfind replaced by a 64-bit count-leading-zeros).
__device__ __forceinline__ int race_and_resolve_wave(bool pred) {
  uint64_t vote = ballot(pred);
  if (vote == 0) return lane_id();
  return 63 - __clzll((long long)vote);
}

// One atomic per wave; every participating lane gets a distinct slot.
// Participation is given by `p`; lanes with p == false get -1.
__device__ __forceinline__ int agg_inc(unsigned int *ctr, bool p) {
  uint64_t m = ballot(p);
  if (m == 0) return -1;
  int leader = __ffsll((unsigned long long)m) - 1;
  unsigned int base = 0;
  if (lane_id() == leader) base = atomicAdd(ctr, (unsigned int)__popcll(m));
  base = (unsigned int)__builtin_amdgcn_readlane((int)base, leader);
  return p ? (int)(base + rank_in(m)) : -1;
}

// 64-bit OR over the wave (butterfly through ds_swizzle-free xor shuffles).
__device__ __forceinline__ uint64_t wave_or(uint64_t x) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) x |= (uint64_t)__shfl_xor((unsigned long long)x, o, kWave);
  return x;
}

// DPP helpers: row_shr:k (0x110+k), row_bcast:15 (0x142), row_bcast:31 (0x143).
template <int CTRL, int ROWMASK, bool BOUND0>
__device__ __forceinline__ int dpp_i32(int old, int src) {
  return __builtin_amdgcn_update_dpp(old, src, CTRL, ROWMASK, 0xf, BOUND0);
}

// Inclusive prefix sum over the 64 lanes (int32).
__device__ __forceinline__ int wave_inclusive_sum(int x) {
  // Within each row of 16 lanes: shifts of 1, 2, 4, 8; lanes shifted in from
  // outside the row read 0 (bound_ctrl).
  x += dpp_i32<0x111, 0xf, true>(0, x);
  x += dpp_i32<0x112, 0xf, true>(0, x);
  x += dpp_i32<0x114, 0xf, true>(0, x);
  x += dpp_i32<0x118, 0xf, true>(0, x);
  // Carry row totals: lane 15 of row r into row r+1 (rows 1 and 3), then
  // lane 31 into rows 2 and 3.
  x += dpp_i32<0x142, 0xa, false>(0, x);
  x += dpp_i32<0x143, 0xc, false>(0, x);
  return x;
}

__device__ __forceinline__ int wave_total(int inclusive) {
  return __builtin_amdgcn_readlane(inclusive, kWave - 1);
}

// Exclusive workgroup scan (blockDim.x a multiple of 64, at most 1024).
// `lds` needs blockDim.x/64 + 1 ints.  Returns the exclusive prefix; *total
// receives the workgroup sum.  Contains __syncthreads().
__device__ __forceinline__ int block_exclusive_sum(int x, int *lds, int *total) {
  const int lane = lane_id();
  const int wid = wave_id();
  const int nw = blockDim.x / kWave;
  int inc = wave_inclusive_sum(x);
  if (lane == kWave - 1) lds[wid] = inc;
  __syncthreads();
  if (wid == 0) {
    int t = lane < nw ? lds[lane] : 0;
    int ti = wave_inclusive_sum(t);
    if (lane < nw) lds[lane] = ti - t;  // exclusive wave offsets
    if (lane == kWave - 1) lds[nw] = ti;
  }
  __syncthreads();
  int res = lds[wid] + inc - x;
  *total = lds[nw];
  __syncthreads();
  return res;
}

}  // namespace gu
