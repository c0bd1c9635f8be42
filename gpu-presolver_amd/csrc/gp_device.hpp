// Device-side helpers shared by the HIP kernel files (gfx950, wave64).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace gpdev {

constexpr int kBlock = 256;  // 4 waves of 64 lanes

__host__ inline unsigned grid_for(int64_t n, int block = kBlock) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > 0x7fffffff) g = 0x7fffffff;
  return static_cast<unsigned>(g);
}

__device__ inline int lane_id() { return static_cast<int>(threadIdx.x & 63); }

__device__ inline uint64_t lanemask_lt(int lane) {
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// Flag and counter updates aggregated per wave: the calling lanes (those for
// which the event happened) elect their lowest lane, which issues a single
// atomic for the whole wave.  Safe in divergent code: only the lanes that
// reach the call take part.
__device__ inline bool wave_leader(uint64_t* active) {
  const uint64_t m = __ballot(1);
  *active = m;
  return lane_id() == __ffsll(static_cast<unsigned long long>(m)) - 1;
}
// A yes/no flag word (0 until some thread sets it to 1, read only after the
// kernel): the wave's leader stores 1 with a plain store, which lands in its
// XCD's L2 and is written back at the end of the kernel.  No read-modify-
// write and no device-coherent load: when most waves of a large grid set the
// same flag, those would all queue on one address at one memory channel.
__device__ inline void wave_set_flag(int32_t* addr) {
  uint64_t m;
  if (wave_leader(&m)) __hip_atomic_store(addr, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
// Max only ever grows the word, so a leader that reads a value at least its
// own skips the atomic.
__device__ inline void wave_max(int32_t* addr, int32_t v) {
  uint64_t m;
  if (wave_leader(&m) && __hip_atomic_load(addr, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) < v)
    atomicMax(addr, v);
}
__device__ inline void wave_add1(int32_t* addr) {
  uint64_t m;
  if (wave_leader(&m)) atomicAdd(addr, static_cast<int32_t>(__popcll(m)));
}

// Inclusive wave64 prefix sum of an int32.
__device__ inline int32_t wave_inclusive_sum(int32_t v) {
  const int lane = lane_id();
#pragma unroll
  for (int d = 1; d < 64; d <<= 1) {
    const int32_t o = __shfl_up(v, d, 64);
    if (lane >= d) v += o;
  }
  return v;
}

// Exclusive block prefix sum of one int32 per thread (blockDim.x == kBlock).
// Returns the thread's exclusive prefix; *total gets the block sum.
__device__ inline int32_t block_exclusive_sum(int32_t v, int32_t* total) {
  __shared__ int32_t s_wave[kBlock / 64];
  const int lane = lane_id();
  const int wave = static_cast<int>(threadIdx.x >> 6);
  const int32_t inc = wave_inclusive_sum(v);
  if (lane == 63) s_wave[wave] = inc;
  __syncthreads();
  int32_t wave_base = 0;
  int32_t sum = 0;
#pragma unroll
  for (int w = 0; w < kBlock / 64; ++w) {
    const int32_t t = s_wave[w];
    if (w < wave) wave_base += t;
    sum += t;
  }
  __syncthreads();
  *total = sum;
  return wave_base + inc - v;
}

// Replace-unless-dominated semantics of the reference's double CAS loops
// (rules/rule_singleton_rows.cu:18-50): slot := v unless slot >= v (max) or
// slot <= v (min).
__device__ inline void atomic_max_f64(double* addr, double v) {
  unsigned long long* a = reinterpret_cast<unsigned long long*>(addr);
  unsigned long long old = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (true) {
    const double cur = __longlong_as_double(static_cast<long long>(old));
    if (cur >= v) return;
    const unsigned long long want = static_cast<unsigned long long>(__double_as_longlong(v));
    const unsigned long long prev = atomicCAS(a, old, want);
    if (prev == old) return;
    old = prev;
  }
}

__device__ inline void atomic_min_f64(double* addr, double v) {
  unsigned long long* a = reinterpret_cast<unsigned long long*>(addr);
  unsigned long long old = __hip_atomic_load(a, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  while (true) {
    const double cur = __longlong_as_double(static_cast<long long>(old));
    if (cur <= v) return;
    const unsigned long long want = static_cast<unsigned long long>(__double_as_longlong(v));
    const unsigned long long prev = atomicCAS(a, old, want);
    if (prev == old) return;
    old = prev;
  }
}

}  // namespace gpdev
