// Primal recovery steps of the structure-specific reductions, applied by
// postsolve between the primal replay of tape segments (reference
// gpu_postsolve.cu:44-141 and 1048-1212): L1 splits, outer pairs, linked
// slacks, max slacks and antipodal components.  Each kernel rewrites x_org in
// place; the order of the steps within one recovery is the caller's.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "gp_device.hpp"
#include "gp_record_kernels.h"

using namespace gpdev;

namespace {

// t <- t + e, e <- (t - e) / rho, both from the values before the step.
__global__ void k_restore_splits(double* x, const int32_t* t_cols, const int32_t* e_cols, const double* rho,
                                 int32_t k) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  const int32_t t = t_cols[i], e = e_cols[i];
  const double tv = x[t], ev = x[e];
  x[t] = tv + ev;
  x[e] = (tv - ev) / rho[i];
}

__global__ void k_restore_outer_pairs(double* x, const int32_t* bound_cols, const int32_t* free_cols, int32_t k) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < k) x[free_cols[i]] = x[bound_cols[i]];
}

__global__ void k_restore_linked_slacks(double* x, const int32_t* slack_cols, const int32_t* t_cols,
                                        const double* factors, int32_t k) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < k) x[slack_cols[i]] = factors[i] * x[t_cols[i]];
}

// One workgroup per max slack s: x[s] = max(lower, max_j factor_j x[t_j]),
// lower = the original lower bound of s when finite, else 0.  A maximum is
// independent of the order it is formed in.
__global__ __launch_bounds__(kBlock) void k_restore_max_slacks(double* x, const int32_t* slack_cols,
                                                               const int32_t* starts, const int32_t* t_cols,
                                                               const double* factors, const double* original_l) {
  __shared__ double s_max[kBlock / 64];
  const int32_t s = blockIdx.x;
  const int32_t slack = slack_cols[s];
  double lo = 0.0;
  if (original_l != nullptr) {
    lo = original_l[slack];
    if (!isfinite(lo)) lo = 0.0;
  }
  double v = lo;
  for (int32_t j = starts[s] + threadIdx.x; j < starts[s + 1]; j += kBlock) v = fmax(v, factors[j] * x[t_cols[j]]);
  for (int d = 32; d >= 1; d >>= 1) v = fmax(v, __shfl_xor(v, d, 64));
  if (lane_id() == 0) s_max[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0) {
    double m = s_max[0];
    for (int w = 1; w < kBlock / 64; ++w) m = fmax(m, s_max[w]);
    x[slack] = m;
  }
}

__global__ void k_antipodal_cache(double* plus_cache, double* minus_cache, const double* x,
                                  const int32_t* root_plus, const int32_t* root_minus, int32_t k) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  plus_cache[i] = x[root_plus[i]];
  minus_cache[i] = x[root_minus[i]];
}

__global__ void k_antipodal_restore(double* x, const int32_t* plus_cols, const int32_t* minus_cols,
                                    const double* plus_lower, const double* minus_lower, const double* plus_cache,
                                    const double* minus_cache, int32_t k) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= k) return;
  x[plus_cols[i]] = plus_lower[i] + plus_cache[i];
  x[minus_cols[i]] = minus_lower[i] + minus_cache[i];
}

}  // namespace

#define GPK_REC_LAUNCH(kernel, n, ...)                                                                \
  do {                                                                                                \
    if ((n) > 0) hipLaunchKernelGGL(kernel, dim3(grid_for(n)), dim3(kBlock), 0, stream, __VA_ARGS__); \
    return hipGetLastError();                                                                         \
  } while (0)

extern "C" {

hipError_t gpk_restore_splits(double* x, const int32_t* t_cols, const int32_t* e_cols, const double* rho, int32_t k,
                              hipStream_t stream) {
  GPK_REC_LAUNCH(k_restore_splits, k, x, t_cols, e_cols, rho, k);
}

hipError_t gpk_restore_outer_pairs(double* x, const int32_t* bound_cols, const int32_t* free_cols, int32_t k,
                                   hipStream_t stream) {
  GPK_REC_LAUNCH(k_restore_outer_pairs, k, x, bound_cols, free_cols, k);
}

hipError_t gpk_restore_linked_slacks(double* x, const int32_t* slack_cols, const int32_t* t_cols,
                                     const double* factors, int32_t k, hipStream_t stream) {
  GPK_REC_LAUNCH(k_restore_linked_slacks, k, x, slack_cols, t_cols, factors, k);
}

hipError_t gpk_restore_max_slacks(double* x, const int32_t* slack_cols, const int32_t* starts, const int32_t* t_cols,
                                  const double* factors, int32_t slacks, const double* original_l,
                                  hipStream_t stream) {
  if (slacks <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_restore_max_slacks, dim3(slacks), dim3(kBlock), 0, stream, x, slack_cols, starts, t_cols,
                     factors, original_l);
  return hipGetLastError();
}

hipError_t gpk_restore_antipodal(double* x, const int32_t* plus_cols, const int32_t* minus_cols,
                                 const int32_t* root_plus, const int32_t* root_minus, const double* plus_lower,
                                 const double* minus_lower, double* plus_cache, double* minus_cache, int32_t k,
                                 hipStream_t stream) {
  if (k <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_antipodal_cache, dim3(grid_for(k)), dim3(kBlock), 0, stream, plus_cache, minus_cache, x,
                     root_plus, root_minus, k);
  hipLaunchKernelGGL(k_antipodal_restore, dim3(grid_for(k)), dim3(kBlock), 0, stream, x, plus_cols, minus_cols,
                     plus_lower, minus_lower, static_cast<const double*>(plus_cache),
                     static_cast<const double*>(minus_cache), k);
  return hipGetLastError();
}

}  // extern "C"
