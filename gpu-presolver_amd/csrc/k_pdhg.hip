// Kernels of the on-device LP solver that takes the reduced LP from presolve
// (restarted primal-dual hybrid gradient, PDHG, on the Ruiz-scaled LP):
// the two half steps of an iteration, Ruiz scaling, sparse products for the
// norm estimate, and the residual sums of the KKT check.  Every sum over a
// vector goes through per-block partial sums reduced in block order by one
// workgroup, so a solve is deterministic run to run.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

#include "gp_device.hpp"
#include "gp_kernels.h"

using namespace gpdev;

namespace {

__device__ inline double clampd(double v, double lo, double hi) { return fmin(fmax(v, lo), hi); }

// Sum of v over the block in a fixed order (lanes by shuffle tree, waves in
// index order); thread 0 returns it.
__device__ inline double block_sum(double v, double* s_w) {
#pragma unroll
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_down(v, d, 64);
  if (lane_id() == 0) s_w[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0)
    for (int w = 0; w < kBlock / 64; ++w) t += s_w[w];
  __syncthreads();
  return t;
}

// x <- proj_[l,u](x - tau (c - A^T y)); xbar = 2 x_new - x_old; x_sum += x_new.
// One thread per column (row j of AT).
__global__ void k_pdhg_primal(const int32_t* __restrict__ at_ptr, const int32_t* __restrict__ at_idx,
                              const double* __restrict__ at_val, const double* __restrict__ c,
                              const double* __restrict__ l, const double* __restrict__ u,
                              const double* __restrict__ y, double* __restrict__ x, double* __restrict__ xbar,
                              double* __restrict__ x_sum, const double* __restrict__ step, int32_t n) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  double g = c[j];
  for (int32_t p = at_ptr[j]; p < at_ptr[j + 1]; ++p) g -= at_val[p] * y[at_idx[p]];
  const double xo = x[j];
  const double xn = clampd(xo - step[0] * g, l[j], u[j]);
  x[j] = xn;
  xbar[j] = 2.0 * xn - xo;
  x_sum[j] += xn;
}

// y <- yt + sigma proj_[AL,AU](-yt / sigma) with yt = y - sigma A xbar (the
// Moreau form of the projected dual ascent); y_sum += y_new.  One thread per row.
__global__ void k_pdhg_dual(const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                            const double* __restrict__ val, const double* __restrict__ AL,
                            const double* __restrict__ AU, const double* __restrict__ xbar, double* __restrict__ y,
                            double* __restrict__ y_sum, const double* __restrict__ step, int32_t m) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= m) return;
  double ax = 0.0;
  for (int32_t p = ptr[i]; p < ptr[i + 1]; ++p) ax += val[p] * xbar[idx[p]];
  const double sigma = step[1];
  const double yt = y[i] - sigma * ax;
  const double yn = yt + sigma * clampd(-yt / sigma, AL[i], AU[i]);
  y[i] = yn;
  y_sum[i] += yn;
}

// out = a * x + b * y (any of the inputs may alias the output).
__global__ void k_pdhg_axpby(double a, const double* __restrict__ x, double b, const double* __restrict__ y,
                             double* out, int32_t n) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = a * x[i] + (y ? b * y[i] : 0.0);
}

__global__ void k_pdhg_mul(const double* __restrict__ x, const double* __restrict__ d, double* out, int32_t n) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = x[i] * d[i];
}

// y = M x for a CSR M.
__global__ void k_pdhg_spmv(const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                            const double* __restrict__ val, const double* __restrict__ x, double* __restrict__ y,
                            int32_t rows) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  double s = 0.0;
  for (int32_t p = ptr[i]; p < ptr[i + 1]; ++p) s += val[p] * x[idx[p]];
  y[i] = s;
}

// Ruiz step: the largest |entry| of every row of the scaled matrix; the row
// scale is divided by its square root.
__global__ void k_pdhg_ruiz(const int32_t* __restrict__ ptr, const double* __restrict__ val, double* d, int32_t rows) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  double mx = 0.0;
  for (int32_t p = ptr[i]; p < ptr[i + 1]; ++p) mx = fmax(mx, fabs(val[p]));
  if (mx > 0.0) d[i] /= sqrt(mx);
}

// out[p] = val[p] * drow[i] * dcol[idx[p]] for every entry p of row i.
__global__ void k_pdhg_scale_csr(const int32_t* __restrict__ ptr, const int32_t* __restrict__ idx,
                                 const double* __restrict__ val, const double* __restrict__ drow,
                                 const double* __restrict__ dcol, double* __restrict__ out, int32_t rows) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= rows) return;
  const double r = drow[i];
  for (int32_t p = ptr[i]; p < ptr[i + 1]; ++p) out[p] = val[p] * r * dcol[idx[p]];
}

// Scaled vectors: multiply (mul) or divide (!mul) by d; infinities stay.
__global__ void k_pdhg_scale_vec(const double* __restrict__ v, const double* __restrict__ d, double* __restrict__ out,
                                 int32_t n, int32_t mul) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = mul ? v[i] * d[i] : v[i] / d[i];
}

// Per-block partial sums, kPartials words per block.
//   rows (original A, unscaled x, y): [0] sum viol^2 where viol = Ax - proj_[AL,AU](Ax),
//     [1] sum of the rows' bound terms of the dual objective, [2] sum b^2 (the
//     larger finite |AL|, |AU| of each row), [3] sum of y^2 on rows whose bound of
//     y's sign is infinite (a dual infeasibility)
//   cols (original AT): [0] sum (g - lambda)^2 with g = c - A^T y and lambda its
//     projection on the reduced-cost cone of the column's bounds, [1] the
//     columns' bound terms of the dual objective, [2] sum c^2, [3] c^T x; z = lambda.
constexpr int kPartials = 4;

__global__ __launch_bounds__(kBlock) void k_pdhg_row_terms(const int32_t* __restrict__ ptr,
                                                           const int32_t* __restrict__ idx,
                                                           const double* __restrict__ val,
                                                           const double* __restrict__ AL,
                                                           const double* __restrict__ AU,
                                                           const double* __restrict__ x,
                                                           const double* __restrict__ y, int32_t m,
                                                           double* __restrict__ partial) {
  __shared__ double s_w[kBlock / 64];
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  double t[kPartials] = {0.0, 0.0, 0.0, 0.0};
  if (i < m) {
    double ax = 0.0;
    for (int32_t p = ptr[i]; p < ptr[i + 1]; ++p) ax += val[p] * x[idx[p]];
    const double lo = AL[i], hi = AU[i];
    const double viol = ax - clampd(ax, lo, hi);
    t[0] = viol * viol;
    const double yi = y[i];
    if (yi > 0.0) {
      if (isfinite(lo)) t[1] = yi * lo;
      else t[3] = yi * yi;
    } else if (yi < 0.0) {
      if (isfinite(hi)) t[1] = yi * hi;
      else t[3] = yi * yi;
    }
    const double b = fmax(isfinite(lo) ? fabs(lo) : 0.0, isfinite(hi) ? fabs(hi) : 0.0);
    t[2] = b * b;
  }
  for (int k = 0; k < kPartials; ++k) {
    const double v = block_sum(t[k], s_w);
    if (threadIdx.x == 0) partial[static_cast<int64_t>(blockIdx.x) * kPartials + k] = v;
  }
}

__global__ __launch_bounds__(kBlock) void k_pdhg_col_terms(const int32_t* __restrict__ at_ptr,
                                                           const int32_t* __restrict__ at_idx,
                                                           const double* __restrict__ at_val,
                                                           const double* __restrict__ c,
                                                           const double* __restrict__ l,
                                                           const double* __restrict__ u,
                                                           const double* __restrict__ x,
                                                           const double* __restrict__ y, int32_t n,
                                                           double* __restrict__ z, double* __restrict__ partial) {
  __shared__ double s_w[kBlock / 64];
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  double t[kPartials] = {0.0, 0.0, 0.0, 0.0};
  if (j < n) {
    double g = c[j];
    for (int32_t p = at_ptr[j]; p < at_ptr[j + 1]; ++p) g -= at_val[p] * y[at_idx[p]];
    const bool lo_fin = isfinite(l[j]), hi_fin = isfinite(u[j]);
    // lambda >= 0 needs a finite lower bound, lambda <= 0 a finite upper one
    const double lam = clampd(g, hi_fin ? -INFINITY : 0.0, lo_fin ? INFINITY : 0.0);
    if (z) z[j] = lam;
    t[0] = (g - lam) * (g - lam);
    if (lam > 0.0) t[1] = lam * l[j];
    else if (lam < 0.0) t[1] = lam * u[j];
    t[2] = c[j] * c[j];
    t[3] = c[j] * x[j];
  }
  for (int k = 0; k < kPartials; ++k) {
    const double v = block_sum(t[k], s_w);
    if (threadIdx.x == 0) partial[static_cast<int64_t>(blockIdx.x) * kPartials + k] = v;
  }
}

// sum of (a - b)^2 and of a^2 (b may be null) into partials [0], [1].
__global__ __launch_bounds__(kBlock) void k_pdhg_sq_terms(const double* __restrict__ a,
                                                          const double* __restrict__ b, int32_t n,
                                                          double* __restrict__ partial) {
  __shared__ double s_w[kBlock / 64];
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  double d2 = 0.0, a2 = 0.0;
  if (i < n) {
    const double d = a[i] - (b ? b[i] : 0.0);
    d2 = d * d;
    a2 = a[i] * a[i];
  }
  double v = block_sum(d2, s_w);
  if (threadIdx.x == 0) partial[static_cast<int64_t>(blockIdx.x) * kPartials] = v;
  v = block_sum(a2, s_w);
  if (threadIdx.x == 0) partial[static_cast<int64_t>(blockIdx.x) * kPartials + 1] = v;
  if (threadIdx.x == 0) {
    partial[static_cast<int64_t>(blockIdx.x) * kPartials + 2] = 0.0;
    partial[static_cast<int64_t>(blockIdx.x) * kPartials + 3] = 0.0;
  }
}

// out[k] = sum over blocks of partial[b * kPartials + k], in block order (one
// workgroup: thread t sums blocks t, t + 256, ..., then the block tree).
__global__ __launch_bounds__(kBlock) void k_pdhg_reduce(const double* __restrict__ partial, int32_t blocks,
                                                        double* __restrict__ out) {
  __shared__ double s_w[kBlock / 64];
  for (int k = 0; k < kPartials; ++k) {
    double t = 0.0;
    for (int32_t b = threadIdx.x; b < blocks; b += kBlock) t += partial[static_cast<int64_t>(b) * kPartials + k];
    const double v = block_sum(t, s_w);
    if (threadIdx.x == 0) out[k] = v;
  }
}

}  // namespace

extern "C" {

hipError_t gpk_pdhg_iterations(gpk_csr A, gpk_csr AT, const double* c, const double* l, const double* u,
                               const double* AL, const double* AU, double* x, double* xbar, double* x_sum, double* y,
                               double* y_sum, const double* step, int32_t iterations, hipStream_t s) {
  const int32_t m = A.rows, n = A.cols;
  for (int32_t k = 0; k < iterations; ++k) {
    if (n > 0)
      hipLaunchKernelGGL(k_pdhg_primal, dim3(grid_for(n)), dim3(kBlock), 0, s, AT.ptr, AT.idx, AT.val, c, l,
                         u, y, x, xbar, x_sum, step, n);
    if (m > 0)
      hipLaunchKernelGGL(k_pdhg_dual, dim3(grid_for(m)), dim3(kBlock), 0, s, A.ptr, A.idx, A.val, AL, AU,
                         xbar, y, y_sum, step, m);
  }
  return hipGetLastError();
}

hipError_t gpk_pdhg_axpby(double a, const double* x, double b, const double* y, double* out, int32_t n,
                          hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pdhg_axpby, dim3(grid_for(n)), dim3(kBlock), 0, s, a, x, b, y, out, n);
  return hipGetLastError();
}

hipError_t gpk_pdhg_mul(const double* x, const double* d, double* out, int32_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pdhg_mul, dim3(grid_for(n)), dim3(kBlock), 0, s, x, d, out, n);
  return hipGetLastError();
}

hipError_t gpk_pdhg_spmv(gpk_csr M, const double* x, double* y, hipStream_t s) {
  if (M.rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pdhg_spmv, dim3(grid_for(M.rows)), dim3(kBlock), 0, s, M.ptr, M.idx, M.val, x, y,
                     M.rows);
  return hipGetLastError();
}

hipError_t gpk_pdhg_ruiz(gpk_csr M, double* d, hipStream_t s) {
  if (M.rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pdhg_ruiz, dim3(grid_for(M.rows)), dim3(kBlock), 0, s, M.ptr, M.val, d, M.rows);
  return hipGetLastError();
}

hipError_t gpk_pdhg_scale_csr(gpk_csr M, const double* drow, const double* dcol, double* out, hipStream_t s) {
  if (M.rows <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pdhg_scale_csr, dim3(grid_for(M.rows)), dim3(kBlock), 0, s, M.ptr, M.idx, M.val, drow,
                     dcol, out, M.rows);
  return hipGetLastError();
}

hipError_t gpk_pdhg_scale_vec(const double* v, const double* d, double* out, int32_t n, int32_t mul, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pdhg_scale_vec, dim3(grid_for(n)), dim3(kBlock), 0, s, v, d, out, n, mul);
  return hipGetLastError();
}

int32_t gpk_pdhg_partials(int32_t n) { return n > 0 ? static_cast<int32_t>(grid_for(n)) * kPartials : kPartials; }

hipError_t gpk_pdhg_row_terms(gpk_csr A, const double* AL, const double* AU, const double* x, const double* y,
                              double* partial, double* out, hipStream_t s) {
  const unsigned blocks = grid_for(A.rows);
  hipLaunchKernelGGL(k_pdhg_row_terms, dim3(blocks), dim3(kBlock), 0, s, A.ptr, A.idx, A.val, AL, AU, x, y,
                     A.rows, partial);
  hipLaunchKernelGGL(k_pdhg_reduce, dim3(1), dim3(kBlock), 0, s, partial, static_cast<int32_t>(blocks), out);
  return hipGetLastError();
}

hipError_t gpk_pdhg_col_terms(gpk_csr AT, const double* c, const double* l, const double* u, const double* x,
                              const double* y, double* z, double* partial, double* out, hipStream_t s) {
  const unsigned blocks = grid_for(AT.rows);
  hipLaunchKernelGGL(k_pdhg_col_terms, dim3(blocks), dim3(kBlock), 0, s, AT.ptr, AT.idx, AT.val, c, l, u, x,
                     y, AT.rows, z, partial);
  hipLaunchKernelGGL(k_pdhg_reduce, dim3(1), dim3(kBlock), 0, s, partial, static_cast<int32_t>(blocks), out);
  return hipGetLastError();
}

hipError_t gpk_pdhg_sq_terms(const double* a, const double* b, int32_t n, double* partial, double* out,
                             hipStream_t s) {
  const unsigned blocks = grid_for(n);
  hipLaunchKernelGGL(k_pdhg_sq_terms, dim3(blocks), dim3(kBlock), 0, s, a, b, n, partial);
  hipLaunchKernelGGL(k_pdhg_reduce, dim3(1), dim3(kBlock), 0, s, partial, static_cast<int32_t>(blocks), out);
  return hipGetLastError();
}

}  // extern "C"

namespace {

// Start vector of the power iteration: 0.5 + frac(i * golden ratio).
__global__ void k_pdhg_seed(double* __restrict__ p, int32_t n) {
  const int32_t i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    const double t = static_cast<double>(i) * 0.6180339887498949;
    p[i] = 0.5 + (t - floor(t));
  }
}

// out = proj_[l,u](xs * d): a scaled primal point back in the LP's own
// variables, clamped so the rounding of l / d * d cannot leave a bound.
__global__ void k_pdhg_unscale_x(const double* __restrict__ xs, const double* __restrict__ d,
                                 const double* __restrict__ l, const double* __restrict__ u, double* __restrict__ out,
                                 int32_t n) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) out[j] = clampd(xs[j] * d[j], l[j], u[j]);
}

// x = projection of 0 on [l, u].
__global__ void k_pdhg_project_zero(const double* __restrict__ l, const double* __restrict__ u,
                                    double* __restrict__ x, int32_t n) {
  const int32_t j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < n) x[j] = clampd(0.0, l[j], u[j]);
}

}  // namespace

extern "C" {

hipError_t gpk_pdhg_seed(double* p, int32_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pdhg_seed, dim3(grid_for(n)), dim3(kBlock), 0, s, p, n);
  return hipGetLastError();
}

hipError_t gpk_pdhg_unscale_x(const double* xs, const double* d, const double* l, const double* u, double* out,
                              int32_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pdhg_unscale_x, dim3(grid_for(n)), dim3(kBlock), 0, s, xs, d, l, u, out, n);
  return hipGetLastError();
}

hipError_t gpk_pdhg_project_zero(const double* l, const double* u, double* x, int32_t n, hipStream_t s) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_pdhg_project_zero, dim3(grid_for(n)), dim3(kBlock), 0, s, l, u, x, n);
  return hipGetLastError();
}

}  // extern "C"
