// Kernels of the serving runtime: the seeded reduced-space point that stands
// in for an LP solver between presolve and postsolve, and an order-free
// checksum of postsolved vectors (bit patterns, so "the same result" means
// bit-identical).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "gp_device.hpp"
#include "gp_kernels.h"

using namespace gpdev;

namespace {

// splitmix64 (Steele, Lea, Flood: "Fast splittable pseudorandom number
// generators", OOPSLA 2014): counter-based, so element i's draw depends only
// on (seed, i).
__device__ inline uint64_t splitmix64(uint64_t z) {
  z += 0x9e3779b97f4a7c15ull;
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
  return z ^ (z >> 31);
}

// Uniform [0, 1) on the 2^-10 grid.
__device__ inline double unit_dyadic(uint64_t seed, uint64_t i) {
  return static_cast<double>(splitmix64(seed ^ splitmix64(i)) >> 54) * (1.0 / 1024.0);
}

__global__ void k_synthetic_point(uint64_t seed, int32_t n, int32_t m, const double* l, const double* u, double* x,
                                  double* y, double* z) {
  const int32_t t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) {
    const double r = unit_dyadic(seed, 3ull * static_cast<uint64_t>(t));
    const double lj = l[t], uj = u[t];
    const bool fl = isfinite(lj), fu = isfinite(uj);
    double v;
    if (fl && fu) v = lj + r * (uj - lj);
    else if (fl) v = lj + 10.0 * r;
    else if (fu) v = uj - 10.0 * r;
    else v = 20.0 * r - 10.0;
    x[t] = fmin(fmax(v, lj), uj);
    z[t] = 2.0 * unit_dyadic(seed, 3ull * static_cast<uint64_t>(t) + 1) - 1.0;
  }
  if (t < m) y[t] = 2.0 * unit_dyadic(seed, 3ull * static_cast<uint64_t>(n) + static_cast<uint64_t>(t)) - 1.0;
}

// acc += sum_i splitmix64(bits(v[i]) ^ splitmix64(salt + i)) (mod 2^64).
__global__ void k_checksum_f64(const double* v, int32_t n, uint64_t salt, unsigned long long* acc) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t h = 0;
  if (i < n) {
    const uint64_t bits = static_cast<uint64_t>(__double_as_longlong(v[i]));
    h = splitmix64(bits ^ splitmix64(salt + static_cast<uint64_t>(i)));
  }
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) {
    const uint32_t lo = __shfl_down(static_cast<uint32_t>(h), d, 64);
    const uint32_t hi = __shfl_down(static_cast<uint32_t>(h >> 32), d, 64);
    h += (static_cast<uint64_t>(hi) << 32) | lo;
  }
  if (lane_id() == 0 && h != 0) atomicAdd(acc, static_cast<unsigned long long>(h));
}

}  // namespace

extern "C" {

hipError_t gpk_synthetic_point(uint64_t seed, int32_t n, int32_t m, const double* l, const double* u, double* x,
                               double* y, double* z, hipStream_t stream) {
  const int32_t t = n > m ? n : m;
  if (t <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_synthetic_point, dim3(grid_for(t)), dim3(kBlock), 0, stream, seed, n, m, l, u, x, y, z);
  return hipGetLastError();
}

hipError_t gpk_checksum_f64(const double* v, int32_t n, uint64_t salt, unsigned long long* acc, hipStream_t stream) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_checksum_f64, dim3(grid_for(n)), dim3(kBlock), 0, stream, v, n, salt, acc);
  return hipGetLastError();
}

}  // extern "C"
